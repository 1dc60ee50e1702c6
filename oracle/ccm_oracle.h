/*
 * CCMpred CPU ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU algorithms for the
 * pseudo-likelihood path.  It is the checker for the HIP product in
 * ccmpred_amd/: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it, and never as the thing measured or shipped.
 *
 * Every function names the reference file:line it follows.  How each one is
 * pinned (closed forms, brute force, finite differences, invariants):
 *   ora_calculate_weights   brute force + closed forms       (tests/test_oracle_weights.py)
 *   ora_init_bias           closed form                      (tests/test_oracle_misc.py)
 *   ora_evaluate_f32/_f64   brute-force FP64, x=0 closed forms, tied-pair
 *                           central differences, invariants  (tests/test_oracle_pll.py)
 *   ora_evaluate_omp_f32    agreement with ora_evaluate_f32  (tests/test_oracle_pll.py)
 *   ora_tied_fd_f64         full re-evaluation central differences (tests/test_oracle_pll.py)
 *   ora_forward_lse_f32/_f64
 *                           brute-force FP64 log-partitions, x=0 closed form
 *                                                            (tests/test_oracle_pll.py)
 *   ora_cg_pll_f64par       same run as ora_cg_pll_f64       (tests/test_golden_cg_maps.py)
 *   ora_sum_submatrices, ora_apc, ora_normalize
 *                           hand-computed cases              (tests/test_oracle_misc.py)
 *   ora_cg                  parity unpinned: the reference optimizer
 *                           (libconjugrad) is not vendored; checked only for
 *                           monotone descent and convergence on a quadratic
 *                           with a known minimiser           (tests/test_oracle_cg.py)
 *
 * The reference was not compiled: its sources need the libconjugrad header
 * (conjugrad.h), which is absent, so the oracle is pinned as listed above.
 */
#ifndef CCM_ORACLE_H
#define CCM_ORACLE_H

#include <stdint.h>

#define ORA_NA 21     /* 20 amino acids + gap; gap is index 20 (sequence.c:10-17) */
#define ORA_NAP 32    /* padded alphabet of the default PADDING build (ccmpred.h:5-9) */

/* x2 offset of W(b,k,a,j) (ccmpred.h:16) */
#define ORA_X2(b, k, a, j, L) ((((long)(b) * (L) + (k)) * ORA_NAP + (a)) * (L) + (j))

/* ccmpred.c:387-390: always pads by 1..32 entries */
static inline long ora_nsingle_padded(int ncol)
{
	long ns = (long)ncol * (ORA_NA - 1);
	return ns + ORA_NAP - (ns % ORA_NAP);
}

static inline long ora_nvar_padded(int ncol)
{
	return ora_nsingle_padded(ncol) + (long)ncol * ncol * ORA_NA * ORA_NAP;
}

#ifdef __cplusplus
extern "C" {
#endif

/* reweighting.c:9-58 (serial; the reference's OpenMP loop races on w[]) */
int ora_identity_threshold(int ncol, float threshold);
void ora_identity_counts(const unsigned char *msa, int ncol, int nrow, int idthres,
                         int32_t *counts);
void ora_calculate_weights(const unsigned char *msa, int ncol, int nrow, float threshold,
                           float *w);
/* reweighting.c:61-67 */
void ora_uniform_weights(int nrow, float *w);

/* ccmpred.c:90-131 */
void ora_init_bias(const unsigned char *msa, int ncol, int nrow, float *x);

/* evaluate_cpu.c:16-166, FP32 and FP64 */
float ora_evaluate_f32(const unsigned char *msa, const float *w, int ncol, int nrow,
                       float lambda_single, float lambda_pair, const float *x, float *g);
double ora_evaluate_f64(const unsigned char *msa, const double *w, int ncol, int nrow,
                        double lambda_single, double lambda_pair, const double *x,
                        double *g);
/* FP64 evaluator parallel over sequence blocks and coupling columns (same
 * model as ora_evaluate_f64, summation order differs only in rounding). */
double ora_evaluate_f64_par(const unsigned char *msa, const double *weights, int ncol, int nrow,
                            double lambda_single, double lambda_pair, const double *x, double *g);
void ora_forward_lse_f32(const unsigned char *msa, int ncol, int nrow, const float *x,
                         float *lse);
void ora_forward_lse_f64(const unsigned char *msa, int ncol, int nrow, const double *x,
                         double *lse);

/* numdif.c:5-57 methodology: FP64 central differences of f along
 * e_{idx_a} + e_{idx_b} (one unit vector when idx_a == idx_b) */
void ora_tied_fd_f64(const unsigned char *msa, const double *w, int ncol, int nrow,
                     double lambda_single, double lambda_pair, const double *x, const long *idx_a,
                     const long *idx_b, int count, double h, double *out);

/* evaluate_cpu_omp.c:18-189, the timed CPU baseline */
typedef struct ora_omp_ctx ora_omp_ctx;
ora_omp_ctx *ora_omp_create(const unsigned char *msa, const float *w, int ncol, int nrow,
                            float lambda_single, float lambda_pair, int nthreads);
float ora_evaluate_omp_f32(ora_omp_ctx *ctx, const float *x, float *g);
void ora_omp_destroy(ora_omp_ctx *ctx);
int ora_omp_max_threads(void);

/* util.c:16-64, util.c:66-108, util.c:110-139 */
void ora_sum_submatrices(const float *x, int ncol, float *out);
void ora_apc(float *mat, int ncol);
void ora_normalize(float *mat, int ncol);

/*
 * Host nonlinear conjugate gradient with the same algorithm as the device
 * CG in ccmpred_amd/csrc/host/cg_device.c (Polak-Ribiere+ beta, line search
 * with an Armijo test -- approximate, on the slope, below a relative
 * objective resolution of 1e-6 --, quadratic-interpolation backtracking,
 * a strong-Wolfe curvature test, extension of short steps and bisection
 * once bracketed; FP64 scalars, FP32 vectors).
 */
typedef double (*ora_eval_fn)(void *ctx, const float *x, float *g, int n);
typedef struct ora_cg_param {
	int max_iterations;
	double epsilon;
	int k;
	int max_linesearch;
	double alpha_mul;
	double ftol;
	double wolfe;
	double fres; /* relative resolution of f for the approximate decrease test; 0: 1e-6 */
} ora_cg_param;
int ora_cg(int n, float *x, double *fx, ora_eval_fn eval, void *ctx, const ora_cg_param *p,
           int *iterations, int *evaluations, double *fhist);

/* Drives ora_cg with the FP64 oracle evaluator (FP32 vectors as on the
 * device, f and g evaluated in FP64). */
int ora_cg_pll_f64(const unsigned char *msa, const float *w, int ncol, int nrow,
                   float lambda_single, float lambda_pair, float *x, const ora_cg_param *p,
                   double *fx, int *iterations, int *evaluations, double *fhist);

/* The same with ora_evaluate_f64_par (OpenMP), for the large alignments. */
int ora_cg_pll_f64par(const unsigned char *msa, const float *w, int ncol, int nrow,
                      float lambda_single, float lambda_pair, float *x, const ora_cg_param *p,
                      double *fx, int *iterations, int *evaluations, double *fhist);

/* Drives ora_cg with the FP32 oracle evaluator. */
int ora_cg_pll_f32(const unsigned char *msa, const float *w, int ncol, int nrow,
                   float lambda_single, float lambda_pair, float *x, const ora_cg_param *p,
                   double *fx, int *iterations, int *evaluations, double *fhist);

#ifdef __cplusplus
}
#endif

#endif
