/*
 * ccmpred_amd -- MI355X (gfx950) implementation of CCMpred's
 * pseudo-likelihood objective-and-gradient path, as a C ABI.
 *
 * The entry points mirror the reference's backend interface so that the
 * reference driver can select this backend the way it selects its CUDA one
 * (ccmpred.c:537-558): init / evaluate / destroy on a `userdata` instance,
 * device-resident x and g in the reference's padded layout, and a
 * conjugate-gradient driver with the libconjugrad call shape.
 *
 * Layout of x and g (ccmpred.h:11-17, ccmpred.c:387-390), nvar_padded floats:
 *   singles  V(j,a) = x[a*L + j], a < 20, padded to nsingle_padded entries;
 *   pairs    W(b,k,a,j) = x[nsingle_padded + ((b*L + k)*32 + a)*L + j].
 */
#ifndef CCMPRED_AMD_H
#define CCMPRED_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* conjugrad_float_t of the reference's default build (CONJUGRAD_FLOAT=32,
 * CMakeLists.txt:57-63) */
typedef float conjugrad_float_t;

#define CCM_N_ALPHA 21
#define CCM_N_ALPHA_PAD 32

/* ccmpred.h:39-92, field for field.  `weights` is a host array of nrow
 * floats that init_hip fills. */
typedef struct userdata {
	unsigned char *msa;
	conjugrad_float_t *weights;
	int ncol;
	int nrow;
	int nsingle;
	int nvar;
	void *extra;
	conjugrad_float_t lambda_single;
	conjugrad_float_t lambda_pair;
	conjugrad_float_t reweighting_threshold;
	void *meta_steps;
} userdata;

/* libconjugrad callback types (call sites ccmpred.c:50-61, 632) */
typedef conjugrad_float_t (*conjugrad_evaluate_t)(void *instance, const conjugrad_float_t *x,
                                                  conjugrad_float_t *g, const int nvar);
typedef int (*conjugrad_progress_t)(void *instance, const conjugrad_float_t *x,
                                    const conjugrad_float_t *g, const conjugrad_float_t fx,
                                    const conjugrad_float_t xnorm, const conjugrad_float_t gnorm,
                                    const conjugrad_float_t step, int n, int k, int ls);

/* ---- sizes (ccmpred.c:385-390) ---------------------------------------- */
long ccm_nsingle_padded(int ncol);
long ccm_nvar_padded(int ncol);

/* ---- backend lifecycle ------------------------------------------------ */

enum { CCM_PRECISION_FP32 = 0, CCM_PRECISION_BF16 = 1 };

typedef struct ccm_hip_config {
	int device;      /* HIP device ordinal */
	int precision;   /* CCM_PRECISION_FP32 (plain) or CCM_PRECISION_BF16 (MFMA) */
	int bf16_passes; /* gradient contraction: 1 = BF16 operand, 2 = hi+lo BF16 split */
	int bf16_fwd_passes; /* forward contraction, same meaning; 0 = as bf16_passes */
	int verbose;     /* 1: print the reweighting summary like the reference */
	int chunk_rows;  /* sequences per chunk (rounded up to 256 in FP32, 512 in BF16); 0 = all */
	int gemm_variant; /* MFMA tile variant (0..ccm_gemm_variants()-1); -1 = default */
	int sparse_fwd;  /* BF16: forward contraction on 2:4 structured-sparse MFMA (1, the
	                  * default, where the shape allows it) or dense (0) */
} ccm_hip_config;

void ccm_hip_default_config(ccm_hip_config *cfg);

/* Replaces init_cuda (evaluate_cuda.c:77-139 / evaluate_cuda.h:16).
 * Computes sequence weights on the device (reweighting.c semantics, bit
 * exact), fills ud->weights, and allocates per-instance device state in
 * ud->extra.  Returns 0 on success. */
int init_hip(void *instance);
int init_hip_ex(void *instance, const ccm_hip_config *cfg);
int init_hip_bf16(void *instance);

/* Replaces evaluate_cuda (evaluate_cuda.c:23-74 / evaluate_cuda.h:8-13).
 * d_x, d_g: device pointers to nvar_padded floats.  Overwrites all of d_g
 * with the tied gradient and returns f.  Returns NaN on error. */
conjugrad_float_t evaluate_hip(void *instance, const conjugrad_float_t *d_x,
                               conjugrad_float_t *d_g, const int nvar);

/* Replaces destroy_cuda (evaluate_cuda.c:141-151). */
int destroy_hip(void *instance);

/* Same as evaluate_hip with the FP64 objective and a status code. */
int ccm_hip_evaluate(void *instance, const float *d_x, float *d_g, int nvar, double *f);

/* The objective alone (no gradient, no stored probabilities); sequences are
 * processed in the instance's chunks.  No counterpart in the reference. */
int ccm_hip_objective(void *instance, const float *d_x, int nvar, double *f);

/* The evaluation split in two for an unchunked instance: forward computes f
 * and keeps the N x 21L probability operand on the device, backward turns it
 * into the tied gradient.  No counterpart in the reference. */
int ccm_hip_forward(void *instance, const float *d_x, int nvar, double *f);
int ccm_hip_backward(void *instance, const float *d_x, float *d_g, int nvar);

/* Relative resolution of the instance's objective: the size of f changes
 * below which its differences are rounding, 1e-5 for a single-pass BF16
 * forward (its logits sum L BF16-rounded couplings) and 1e-6 otherwise.  The
 * device CG's approximate decrease test uses it when the callback is
 * evaluate_hip.  No counterpart in the reference. */
double ccm_hip_objective_resolution(void *instance);

/* Device bytes the instance holds (all buffers allocated at init). */
size_t ccm_hip_device_bytes(void *instance);

/* HIP stream the instance runs on (hipStream_t). */
void *ccm_hip_stream(void *instance);

/* Per-step device timing with HIP events on the instance stream (no
 * counterpart in the reference).  set_timing(on=1) creates the events and
 * clears the accumulators, on=0 releases them (the events cost host time
 * per evaluation); get_timing fills up to n averages in ms and
 * returns the number of timed evaluations.  Steps, FP32 path: gather_fwd,
 * softmax, gather_grad, assembly (pair epilogue, singles and the objective's
 * sum); BF16 path: convert_wt, gemm_fwd (with the softmax when fused),
 * softmax (unfused only), gemm_grad, assembly; with the sparse forward
 * convert_ws, sparse_fwd_softmax, gemm_grad, assemble. */
int ccm_hip_set_timing(void *instance, int on);
int ccm_hip_get_timing(void *instance, double *avg_ms, int n);
/* Names of those steps (static strings; up to n filled); returns their
 * number.  BF16 with the softmax fused into the forward contraction:
 * convert_wt, gemm_fwd_softmax, gemm_grad, assemble. */
int ccm_hip_step_names(void *instance, const char **names, int n);

/* ---- device conjugate gradient (libconjugrad's conjugrad_gpu shape) --- */

/* Status codes; same names as libconjugrad's, own values (the library is
 * not vendored with the reference). */
#define CONJUGRAD_SUCCESS 0
#define CONJUGRAD_ALREADY_MINIMIZED 1
#define CONJUGRADERR_MAXIMUMITERATION 2
#define CONJUGRADERR_MAXIMUMLINESEARCH -1
#define CONJUGRADERR_DEVICE -2

typedef struct ccm_conjugrad_parameter {
	int max_iterations;
	double epsilon;
	int k;
	int max_linesearch;
	double alpha_mul;
	double ftol;
	double wolfe;
	int line_search; /* CCM_LINE_SEARCH_*: where the line search runs (no counterpart) */
} ccm_conjugrad_parameter;

/* CCM_LINE_SEARCH_DEVICE (the default): when the callback is evaluate_hip
 * on a BF16, unchunked instance, no progress callback is given and d_x is
 * 16-byte aligned, the whole CG -- vector steps, evaluations, dot products,
 * line-search and direction decisions -- runs on the device as the batched
 * form's graph-captured state machine on that instance, with one host
 * synchronisation per 8 rounds (a round: one evaluation); the results are
 * the host loop's bit for bit.  Otherwise, and with CCM_LINE_SEARCH_HOST,
 * the host makes the decisions after each evaluation. */
#define CCM_LINE_SEARCH_DEVICE 0
#define CCM_LINE_SEARCH_HOST 1

/* ccmpred.c:526-534 defaults */
void ccm_conjugrad_init(ccm_conjugrad_parameter *param);

/* Nonlinear CG on device vectors (conjugrad_gpu at ccmpred.c:632); see
 * CCM_LINE_SEARCH_* for where its decisions are made.
 * Polak-Ribiere+ direction, backtracking line search with an Armijo test
 * (ftol) and a curvature test (wolfe), FP64 scalars and dot products.
 * d_x: device pointer, n floats; *fx receives the final objective. */
int ccm_conjugrad_gpu(int n, conjugrad_float_t *d_x, conjugrad_float_t *fx,
                      conjugrad_evaluate_t evaluate, conjugrad_progress_t progress, void *instance,
                      const ccm_conjugrad_parameter *param);

/* Same with counters and the per-iteration objective history (fhist: at
 * least max_iterations doubles, may be NULL). */
int ccm_conjugrad_gpu_ex(int n, conjugrad_float_t *d_x, double *fx, conjugrad_evaluate_t evaluate,
                         conjugrad_progress_t progress, void *instance,
                         const ccm_conjugrad_parameter *param, int *iterations, int *evaluations,
                         double *fhist);

/* ---- batched form: many alignments optimised together ----------------- */
/* No counterpart in the reference (one alignment per ccmpred process).
 * Every alignment gets the BF16 evaluator of init_hip_ex (unchunked) and its
 * own device vectors x, g, gprev, s; each step of an evaluation or of the
 * conjugate gradient is one grouped launch over all alignments, and the line
 * search runs as a device-side state machine (the decisions of
 * ccm_conjugrad_gpu_ex), so that a whole CG round is host-free and rounds
 * are replayed from a captured hipGraph. */
typedef struct ccm_batch ccm_batch;

/* uds[0..nprob): msa, ncol, nrow, lambdas and reweighting threshold set by
 * the caller; weights (host, nrow floats) are filled as init_hip does.
 * cfg: precision must be CCM_PRECISION_BF16 (chunk_rows is ignored).
 * Returns 0, or non-zero with *out = NULL. */
int ccm_batch_create(userdata *uds, int nprob, const ccm_hip_config *cfg, ccm_batch **out);
void ccm_batch_destroy(ccm_batch *b);
int ccm_batch_size(const ccm_batch *b);
/* device vectors of alignment p (nvar_padded floats): x is the caller's to
 * fill before a solve and holds the result after it; g the last gradient */
float *ccm_batch_x(ccm_batch *b, int p);
float *ccm_batch_g(ccm_batch *b, int p);
/* host copies of alignment p's vectors (nvar_padded floats each; read:
 * either pointer may be NULL) */
int ccm_batch_write_x(ccm_batch *b, int p, const float *host_x);
int ccm_batch_read(ccm_batch *b, int p, float *host_x, float *host_g);
/* util.c:16-139 on the device for alignment p: the L x L contact map of its
 * current x (sum_submatrices, then apc and normalize if asked) into host_out,
 * so that only L*L floats leave the device instead of x. */
int ccm_batch_contact_map(ccm_batch *b, int p, int apc, int normalize, float *host_out);
void *ccm_batch_stream(ccm_batch *b);
size_t ccm_batch_device_bytes(const ccm_batch *b);

/* GEMM tile variant per alignment for the forward (fwd) and gradient (grad)
 * contractions; set: NULL or -1 entries take the default of the shape. */
int ccm_batch_set_variants(ccm_batch *b, const int *fwd, const int *grad);
int ccm_batch_get_variants(const ccm_batch *b, int *fwd, int *grad);
/* Autotuner: times the grouped contractions on the device (median of reps)
 * over assignments of the direct-to-LDS tile variants and keeps the
 * fastest; ms[0..1] receive the tuned forward and gradient times (may be
 * NULL). */
int ccm_batch_autotune(ccm_batch *b, int reps, double *ms);

/* One grouped evaluation of every alignment at its x: gradients into
 * ccm_batch_g, f[0..nprob) FP64 objectives. */
int ccm_batch_evaluate(ccm_batch *b, double *f);

/* Conjugate gradient on every alignment from its x, rounds_per_graph CG
 * rounds (vector step, evaluation, dots, control) per captured graph.
 * *rounds (may be NULL) receives the rounds run. */
int ccm_batch_solve(ccm_batch *b, const ccm_conjugrad_parameter *param, int rounds_per_graph,
                    int *rounds);
/* Outcome of alignment p after ccm_batch_solve: status (CONJUGRAD_* codes),
 * final f (as the single-alignment CG reports it), counters and the
 * per-iteration log (4 doubles per iteration: f, step, line-search trials,
 * <g,g>; room for max_iterations entries, may be NULL). */
int ccm_batch_result(ccm_batch *b, int p, int *status, double *fx, int *iterations,
                     int *evaluations, double *log);

/* ---- host model set-up and contact-map output ------------------------- */

/* sequence.c:48-80: PSICOV reader; returns a malloc'd N x L index matrix,
 * NULL on a FASTA/a3m header (the reference exits with status 100). */
unsigned char *ccm_read_msa(FILE *f, int *ncol, int *nrow);

/* ccmpred.c:90-131: singles from pseudo-counted column frequencies relative
 * to the gap; couplings and padding zero.  x: nvar_padded floats (host). */
void ccm_init_bias(float *x, const unsigned char *msa, int ncol, int nrow);

/* util.c:16-64: Frobenius norm of each mean-centred 21x21 block over the
 * 20x20 residue part, zero diagonal.  x: host, out: L x L. */
void ccm_sum_submatrices(const float *x, float *out, int ncol);
/* util.c:66-108 */
void ccm_apc(float *mat, int ncol);
/* util.c:110-139 */
void ccm_normalize(float *mat, int ncol);
/* io.c:20-28 */
void ccm_write_matrix(FILE *out, const float *mat, int ncol, int nrow);
/* io.c:31-60 */
void ccm_write_raw(FILE *out, const float *x, int ncol);
/* io.c:105-196 (write_raw_msgpack, the -b output): a MessagePack map
 * {format: "ccm-1", ncol, x_single: [20L float32, i outer], x_pair:
 * {"i/j": {i, j, x: [441 float32 W(b,j,a,i), a outer]} for i < j}}, without
 * the optional JSON metadata key.  Returns 0, or -1 on a write error. */
int ccm_write_raw_msgpack(FILE *out, const float *x, int ncol);
/* io.c:62-103; fills both mirrored couplings.  Returns 0, or -1 on a
 * malformed file (the reference asserts). */
int ccm_read_raw(const char *filename, int ncol, float *x);

#ifdef __cplusplus
}
#endif

#endif
