/*
 * CCMpred CPU ORACLE -- TEST INFRASTRUCTURE ONLY (see ccm_oracle.h).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "ccm_oracle.h"

/* ---- PLL evaluator, FP32 and FP64 instantiations (evaluate_cpu.c) ------ */

#define REAL float
#define SFX f32
#define EXPR(v) expf(v)
#define LOGR(v) logf(v)
#include "pll_tmpl.h"
#undef REAL
#undef SFX
#undef EXPR
#undef LOGR

#define REAL double
#define SFX f64
#define EXPR(v) exp(v)
#define LOGR(v) log(v)
#include "pll_tmpl.h"
#undef REAL
#undef SFX
#undef EXPR
#undef LOGR

/*
 * FP64 evaluator for large alignments, OpenMP-parallel: the model of
 * evaluate_cpu.c:16-166 (same logits, log-partition, objective, node and
 * pair gradient, symmetrisation, self-edge zeroing and L2 terms) with the
 * work split the way evaluate_cpu_omp.c:18-189 splits it -- per-sequence
 * logits in parallel over sequences, the raw pair-gradient rows in parallel
 * over the coupling column k -- over blocks of 256 sequences.  Only the
 * FP64 summation order differs from ora_evaluate_f64, which it matches to
 * rounding; it exists so that FP64 reference values at L = 600, N = 20000
 * take seconds on a multi-core host.
 */
double ora_evaluate_f64_par(const unsigned char *msa, const double *weights, int ncol, int nrow,
                            double lambda_single, double lambda_pair, const double *x, double *g)
{
	enum { BLK = 256 };
	const long L = ncol;
	const long C = ORA_NA * L;
	const long nsingle = L * (ORA_NA - 1);
	const long nsp = ora_nsingle_padded(ncol);
	const long nvp = ora_nvar_padded(ncol);
	const double *x1 = x;
	const double *x2 = x + nsp;
	double *g1 = g;
	double *g2 = g + nsp;
	double *raw = (double *)calloc((size_t)(nvp - nsp), sizeof(double));
	double *D = (double *)malloc(sizeof(double) * BLK * (size_t)C);
	double f = 0;
	memset(g, 0, sizeof(double) * (size_t)nvp);

	for (long n0 = 0; n0 < nrow; n0 += BLK) {
		const long nb = nrow - n0 < BLK ? nrow - n0 : BLK;
		double fb = 0;
		/* D[i][c] = w_n (P_n(c) - [x_nj = a]), c = a*L + j */
		#pragma omp parallel for reduction(+ : fb) schedule(dynamic, 4)
		for (long i = 0; i < nb; i++) {
			const long n = n0 + i;
			const double wn = weights[n];
			const unsigned char *seq = msa + n * L;
			double *lg = D + i * C;
			for (long a = 0; a < ORA_NA - 1; a++)
				for (long s = 0; s < L; s++)
					lg[a * L + s] = x1[a * L + s];
			for (long s = 0; s < L; s++)
				lg[(ORA_NA - 1) * L + s] = 0;
			for (long k = 0; k < L; k++) {
				const double *row = x2 + ORA_X2(seq[k], k, 0, 0, L);
				for (long c = 0; c < C; c++)
					lg[c] += row[c];
			}
			for (long s = 0; s < L; s++) {
				double z = 0;
				for (long a = 0; a < ORA_NA; a++)
					z += exp(lg[a * L + s]);
				const double lse = log(z);
				fb += wn * (lse - lg[seq[s] * L + s]);
				for (long a = 0; a < ORA_NA; a++)
					lg[a * L + s] = wn * exp(lg[a * L + s] - lse);
				lg[seq[s] * L + s] -= wn;
			}
		}
		f += fb;
		/* node gradient: column sums over the block, a < 20 */
		#pragma omp parallel for
		for (long c = 0; c < nsingle; c++) {
			double t = 0;
			for (long i = 0; i < nb; i++)
				t += D[i * C + c];
			g1[c] += t;
		}
		/* raw pair gradient rows (x_nk, k) */
		#pragma omp parallel for schedule(dynamic, 1)
		for (long k = 0; k < L; k++)
			for (long i = 0; i < nb; i++) {
				double *rg = raw + ORA_X2(msa[(n0 + i) * L + k], k, 0, 0, L);
				const double *d = D + i * C;
				for (long c = 0; c < C; c++)
					rg[c] += d[c];
			}
	}

	/* g2 = raw + raw^T, zero self-edge blocks */
	#pragma omp parallel for
	for (long b = 0; b < ORA_NA; b++)
		for (long k = 0; k < L; k++)
			for (long a = 0; a < ORA_NA; a++)
				for (long j = 0; j < L; j++)
					g2[ORA_X2(b, k, a, j, L)] =
					    k == j ? 0.0 : raw[ORA_X2(b, k, a, j, L)] + raw[ORA_X2(a, j, b, k, L)];

	/* L2: real singles, then the whole padded pair block */
	double reg = 0;
	for (long v = 0; v < nsingle; v++) {
		reg += lambda_single * x[v] * x[v];
		g[v] += 2 * lambda_single * x[v];
	}
	#pragma omp parallel for reduction(+ : reg)
	for (long v = nsp; v < nvp; v++) {
		reg += 0.5 * lambda_pair * x[v] * x[v];
		g[v] += 2 * lambda_pair * x[v];
	}
	free(raw);
	free(D);
	return f + reg;
}

/* ---- reweighting (reweighting.c:9-67) ---------------------------------- */

/* reweighting.c:10: threshold (float) times ncol is a float product; its
 * ceiling is taken in double and truncated to int. */
int ora_identity_threshold(int ncol, float threshold)
{
	float prod = threshold * (float)(uint64_t)ncol;
	return (int)ceil((double)prod);
}

/* counts[i] = #{ j != i : identities(i, j) > idthres } */
void ora_identity_counts(const unsigned char *msa, int ncol, int nrow, int idthres,
                         int32_t *counts)
{
	memset(counts, 0, sizeof(int32_t) * (size_t)nrow);
	for (long i = 0; i < nrow; i++) {
		const unsigned char *si = msa + i * ncol;
		for (long j = i + 1; j < nrow; j++) {
			const unsigned char *sj = msa + j * ncol;
			int ids = 0;
			for (int k = 0; k < ncol; k++)
				ids += si[k] == sj[k];
			if (ids > idthres) {
				counts[i]++;
				counts[j]++;
			}
		}
	}
}

/*
 * reweighting.c:9-58.  The reference loops over the N(N+1)/2 pairs i <= j,
 * so the diagonal pair adds 2 to w[i] when ncol > idthres; w accumulates in
 * float and is finally set to 1./(w - 1) in double, rounded to float.
 */
void ora_calculate_weights(const unsigned char *msa, int ncol, int nrow, float threshold,
                           float *w)
{
	int idthres = ora_identity_threshold(ncol, threshold);
	int32_t *cnt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nrow > 0 ? nrow : 1));
	ora_identity_counts(msa, ncol, nrow, idthres, cnt);
	int diag = ncol > idthres ? 2 : 0;
	for (long i = 0; i < nrow; i++) {
		float acc = (float)(cnt[i] + diag);
		w[i] = (float)(1. / (acc - 1));
	}
	free(cnt);
}

void ora_uniform_weights(int nrow, float *w)
{
	for (long i = 0; i < nrow; i++)
		w[i] = 1.0f;
}

/* ---- init_bias (ccmpred.c:90-131) -------------------------------------- */

void ora_init_bias(const unsigned char *msa, int ncol, int nrow, float *x)
{
	memset(x, 0, sizeof(float) * (size_t)ora_nvar_padded(ncol));
	for (long j = 0; j < ncol; j++) {
		int cnt[ORA_NA];
		for (int a = 0; a < ORA_NA; a++)
			cnt[a] = 1;
		for (long i = 0; i < nrow; i++)
			cnt[msa[i * ncol + j]]++;
		int total = nrow + ORA_NA;
		float frac[ORA_NA];
		for (int a = 0; a < ORA_NA; a++)
			frac[a] = (float)cnt[a] / (float)total;
		float gapnorm = logf(frac[ORA_NA - 1]);
		for (int a = 0; a < ORA_NA - 1; a++)
			x[a * ncol + j] = logf(frac[a]) - gapnorm;
	}
}

/* ---- OpenMP evaluator (evaluate_cpu_omp.c:18-189) ---------------------- */

struct ora_omp_ctx {
	const unsigned char *msa;
	const float *w;
	int ncol, nrow;
	float lambda_single, lambda_pair;
	float *raw; /* init-time pair scratch, like omp_extra_userdata.g2 */
};

int ora_omp_max_threads(void)
{
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}

ora_omp_ctx *ora_omp_create(const unsigned char *msa, const float *w, int ncol, int nrow,
                            float lambda_single, float lambda_pair, int nthreads)
{
	ora_omp_ctx *c = (ora_omp_ctx *)calloc(1, sizeof(*c));
	c->msa = msa;
	c->w = w;
	c->ncol = ncol;
	c->nrow = nrow;
	c->lambda_single = lambda_single;
	c->lambda_pair = lambda_pair;
	c->raw = (float *)malloc(sizeof(float) * (size_t)ncol * ncol * ORA_NA * ORA_NAP);
#ifdef _OPENMP
	if (nthreads > 0)
		omp_set_num_threads(nthreads);
#else
	(void)nthreads;
#endif
	return c;
}

void ora_omp_destroy(ora_omp_ctx *c)
{
	if (!c)
		return;
	free(c->raw);
	free(c);
}

/*
 * Same structure as evaluate_cpu_omp.c: per-call buffers of N*32*L floats
 * (:50-52), row-parallel logits (:57-99), serial f (:101-107), column-parallel
 * node gradient (:109-121), a serial loop over sequences with a parallel
 * loop over columns for the pair gradient (:124-144), serial symmetrisation
 * and self-edge reset (:147-165), parallel single and serial pair L2 (:170-179).
 */
float ora_evaluate_omp_f32(ora_omp_ctx *c, const float *x, float *g)
{
	const long L = c->ncol, N = c->nrow;
	const long nsingle = L * (ORA_NA - 1);
	const long nsp = ora_nsingle_padded(c->ncol);
	const long nvp = ora_nvar_padded(c->ncol);
	const unsigned char *msa = c->msa;
	const float *x1 = x, *x2 = x + nsp;
	float *g1 = g, *g2 = g + nsp;
	float *raw = c->raw;
	const float lam1 = c->lambda_single, lam2 = c->lambda_pair;

	float f = 0;
	memset(g, 0, sizeof(float) * (size_t)nvp);
	memset(raw, 0, sizeof(float) * (size_t)(nvp - nsp));

	float *logit = (float *)malloc(sizeof(float) * (size_t)N * ORA_NAP * L);
	float *lz = (float *)malloc(sizeof(float) * (size_t)N * L);
	float *prob = (float *)malloc(sizeof(float) * (size_t)N * ORA_NAP * L);
	memset(lz, 0, sizeof(float) * (size_t)N * L);

#define PCI(i, a, s) (((i) * ORA_NA + (a)) * L + (s))

	#pragma omp parallel for
	for (long i = 0; i < N; i++) {
		const unsigned char *seq = msa + i * L;
		for (long a = 0; a < ORA_NA - 1; a++)
			for (long s = 0; s < L; s++)
				logit[PCI(i, a, s)] = x1[a * L + s];
		for (long s = 0; s < L; s++)
			logit[PCI(i, ORA_NA - 1, s)] = 0;
		for (long k = 0; k < L; k++) {
			float *pp = logit + PCI(i, 0, 0);
			const float *pw = x2 + ORA_X2(seq[k], k, 0, 0, L);
			for (long t = 0; t < ORA_NA * L; t++)
				pp[t] += pw[t];
		}
		for (long a = 0; a < ORA_NA; a++)
			for (long s = 0; s < L; s++)
				lz[i * L + s] += expf(logit[PCI(i, a, s)]);
		for (long s = 0; s < L; s++)
			lz[i * L + s] = logf(lz[i * L + s]);
		for (long a = 0; a < ORA_NA; a++)
			for (long s = 0; s < L; s++)
				prob[PCI(i, a, s)] = expf(logit[PCI(i, a, s)] - lz[i * L + s]);
	}

	for (long i = 0; i < N; i++) {
		const float wi = c->w[i];
		for (long k = 0; k < L; k++) {
			int xk = msa[i * L + k];
			f += wi * (lz[i * L + k] - logit[PCI(i, xk, k)]);
		}
	}

	#pragma omp parallel for
	for (long k = 0; k < L; k++) {
		for (long i = 0; i < N; i++) {
			const float wi = c->w[i];
			int xk = msa[i * L + k];
			if (xk < ORA_NA - 1)
				g1[xk * L + k] -= wi;
			for (long a = 0; a < ORA_NA - 1; a++)
				g1[a * L + k] += wi * prob[PCI(i, a, k)];
		}
	}

	for (long i = 0; i < N; i++) {
		const float wi = c->w[i];
		const unsigned char *seq = msa + i * L;
		#pragma omp parallel for
		for (long k = 0; k < L; k++) {
			float *pg = raw + ORA_X2(seq[k], k, 0, 0, L);
			for (long j = 0; j < L; j++)
				pg[seq[j] * L + j] -= wi;
			const float *pp = prob + PCI(i, 0, 0);
			for (long t = 0; t < ORA_NA * L; t++)
				pg[t] += wi * pp[t];
		}
	}
#undef PCI

	for (long b = 0; b < ORA_NA; b++)
		for (long k = 0; k < L; k++)
			for (long a = 0; a < ORA_NA; a++)
				for (long j = 0; j < L; j++)
					g2[ORA_X2(b, k, a, j, L)] =
					    raw[ORA_X2(b, k, a, j, L)] + raw[ORA_X2(a, j, b, k, L)];
	for (long b = 0; b < ORA_NA; b++)
		for (long k = 0; k < L; k++)
			for (long a = 0; a < ORA_NA; a++)
				g2[ORA_X2(b, k, a, k, L)] = 0;

	float reg = 0;
	#pragma omp parallel for reduction(+ : reg)
	for (long v = 0; v < nsingle; v++) {
		reg += lam1 * x[v] * x[v];
		g[v] += 2.0f * lam1 * x[v];
	}
	for (long v = nsp; v < nvp; v++) {
		reg += 0.5f * lam2 * x[v] * x[v];
		g[v] += 2.0f * lam2 * x[v];
	}
	f += reg;

	free(logit);
	free(lz);
	free(prob);
	return f;
}

/* ---- contact scores (util.c:16-139) ------------------------------------ */

void ora_sum_submatrices(const float *x, int ncol, float *out)
{
	const long L = ncol;
	const float *x2 = x + ora_nsingle_padded(ncol);
	memset(out, 0, sizeof(float) * (size_t)L * L);
	for (long k = 0; k < L; k++) {
		for (long j = 0; j < L; j++) {
			float mean = 0;
			for (int a = 0; a < ORA_NA; a++)
				for (int b = 0; b < ORA_NA; b++)
					mean += x2[ORA_X2(b, k, a, j, L)];
			mean /= (float)(ORA_NA * ORA_NA);
			for (int a = 0; a < ORA_NA - 1; a++)
				for (int b = 0; b < ORA_NA - 1; b++) {
					float d = x2[ORA_X2(b, k, a, j, L)] - mean;
					out[k * L + j] += d * d;
				}
		}
	}
	for (long k = 0; k < L; k++) {
		for (long j = 0; j < L; j++)
			out[k * L + j] = (float)sqrt((double)out[k * L + j]);
		out[k * L + k] = 0;
	}
}

void ora_apc(float *mat, int ncol)
{
	const long L = ncol;
	float *means = (float *)calloc((size_t)L, sizeof(float));
	float meansum = 0;
	for (long i = 0; i < L; i++)
		for (long j = 0; j < L; j++) {
			float v = mat[i * L + j];
			means[j] += v / (float)L;
			meansum += v;
		}
	meansum /= (float)(L * L);
	for (long i = 0; i < L; i++)
		for (long j = 0; j < L; j++)
			mat[i * L + j] -= (means[i] * means[j]) / meansum;
	float lo = INFINITY;
	for (long i = 0; i < L; i++)
		for (long j = i + 1; j < L; j++)
			if (mat[i * L + j] < lo)
				lo = mat[i * L + j];
	for (long i = 0; i < L; i++) {
		for (long j = 0; j < L; j++)
			mat[i * L + j] -= lo;
		mat[i * L + i] = 0;
	}
	free(means);
}

void ora_normalize(float *mat, int ncol)
{
	const long L = ncol;
	float lo = mat[1], hi = mat[1];
	for (long i = 0; i < L; i++)
		for (long j = 0; j < L; j++) {
			if (i == j)
				continue;
			float v = mat[i * L + j];
			if (v < lo)
				lo = v;
			if (v > hi)
				hi = v;
		}
	float range = hi - lo;
	for (long i = 0; i < L; i++)
		for (long j = 0; j < L; j++)
			mat[i * L + j] = i == j ? 0.0f : (mat[i * L + j] - lo) / range;
}

/* ---- host conjugate gradient ------------------------------------------- */

static double ora_dot(const float *a, const float *b, long n)
{
	/* fixed-order pairwise FP64 summation over blocks of 4096 */
	double total = 0;
	for (long s = 0; s < n; s += 4096) {
		long e = s + 4096 < n ? s + 4096 : n;
		double part = 0;
		for (long i = s; i < e; i++)
			part += (double)a[i] * (double)b[i];
		total += part;
	}
	return total;
}

/* Sufficient decrease of a trial at step alpha (f0, dg0: objective and
 * slope at the start of the search; ft, dgt: at the trial).  The Armijo
 * test, or, where f differences sink below the objective's resolution
 * (fres relative; 1e-6 when the parameter leaves it 0), the approximate
 * form of Hager and Zhang (SIAM J. Optim. 16(1), 2005, eq. 4.1): the slope
 * at the trial has not risen above (1 - 2 ftol) |dg0| -- the Armijo test of
 * a quadratic, written with derivatives -- and f has not risen by more than
 * that resolution. */
static int ora_cg_decrease(double ft, double f0, double alpha, double dg0, double dgt, double ftol,
                           double fres)
{
	if (ft <= f0 + ftol * alpha * dg0)
		return 1;
	return ft <= f0 + (fres > 0.0 ? fres : 1e-6) * fabs(f0) && dgt <= (2.0 * ftol - 1.0) * dg0;
}

/* Next step after a trial that failed the decrease test with no lower
 * bracket: the minimiser of the quadratic through f0, dg0 and ft
 * (Nocedal and Wright, Numerical Optimization, 2nd ed., eq. 3.57),
 * kept within [0.1 alpha, alpha_mul alpha]. */
static double ora_cg_shrink(double ft, double f0, double alpha, double dg0, double alpha_mul)
{
	const double den = 2.0 * ((ft - f0) - dg0 * alpha);
	const double aq = (-dg0 * alpha) * alpha / den;
	const double amin = 0.1 * alpha, amax = alpha_mul * alpha;
	return !(aq >= amin) ? amin : (aq > amax ? amax : aq);
}

enum { ORA_CG_SUCCESS = 0, ORA_CG_ALREADY_MINIMIZED = 1, ORA_CG_MAXITER = 2,
       ORA_CG_ERR_LINESEARCH = -1 };

int ora_cg(int n, float *x, double *fx, ora_eval_fn eval, void *ctx, const ora_cg_param *p,
           int *iterations, int *evaluations, double *fhist)
{
	float *g = (float *)malloc(sizeof(float) * (size_t)n);
	float *gprev = (float *)malloc(sizeof(float) * (size_t)n);
	float *s = (float *)malloc(sizeof(float) * (size_t)n);
	double *hist = (double *)malloc(sizeof(double) * (size_t)(p->max_iterations + 1));
	int nevals = 0, it = 0, ret = ORA_CG_MAXITER;

	double f = eval(ctx, x, g, n);
	nevals++;
	double gg = ora_dot(g, g, n);
	if (gg == 0.0) {
		ret = ORA_CG_ALREADY_MINIMIZED;
		goto done;
	}
	for (long i = 0; i < n; i++)
		s[i] = -g[i];
	double dg0 = -gg;
	double alpha = 1.0 / sqrt(gg);
	int steepest = 1;

	while (it < p->max_iterations) {
		const double f0 = f;
		float *tmp = gprev;
		gprev = g;
		g = tmp;
		/* line search along s; x moves in place, `moved` is its offset */
		double moved = 0.0, ft = 0.0, lo = 0.0, hi = INFINITY;
		int accepted = 0, ls;
		for (ls = 1; ls <= p->max_linesearch; ls++) {
			const float step = (float)(alpha - moved);
			for (long i = 0; i < n; i++)
				x[i] += step * s[i];
			moved = alpha;
			ft = eval(ctx, x, g, n);
			nevals++;
			const double dgt = ora_dot(g, s, n);
			if (!ora_cg_decrease(ft, f0, alpha, dg0, dgt, p->ftol, p->fres)) {
				/* not enough decrease: shrink, bisecting once bracketed */
				hi = alpha;
				alpha = lo > 0.0 ? 0.5 * (lo + hi)
				                 : ora_cg_shrink(ft, f0, alpha, dg0, p->alpha_mul);
				continue;
			}
			if (fabs(dgt) <= -p->wolfe * dg0 || ls == p->max_linesearch) {
				accepted = 1;
				break;
			}
			if (dgt > 0.0) {
				hi = alpha; /* past the minimum along s */
			} else {
				lo = alpha; /* still descending steeply: extend */
			}
			alpha = isinf(hi) ? alpha / p->alpha_mul : 0.5 * (lo + hi);
		}
		if (!accepted && lo > 0.0) {
			/* last trial failed the decrease test: return to the longest
			 * step that passed it */
			const float step = (float)(lo - moved);
			for (long i = 0; i < n; i++)
				x[i] += step * s[i];
			moved = alpha = lo;
			ft = eval(ctx, x, g, n);
			nevals++;
			accepted = 1;
		}
		if (!accepted) {
			const float back = (float)(-moved);
			for (long i = 0; i < n; i++)
				x[i] += back * s[i];
			tmp = gprev;
			gprev = g;
			g = tmp;
			f = f0;
			if (steepest) {
				ret = ORA_CG_ERR_LINESEARCH;
				goto done;
			}
			for (long i = 0; i < n; i++)
				s[i] = -g[i];
			dg0 = -gg;
			alpha = 1.0 / sqrt(gg);
			steepest = 1;
			continue;
		}
		f = ft;
		hist[it] = f;
		if (fhist)
			fhist[it] = f;
		it++;
		if (p->epsilon > 0.0 && p->k > 0 && it > p->k) {
			double before = hist[it - 1 - p->k];
			if ((before - f) / fabs(f) < p->epsilon) {
				ret = ORA_CG_SUCCESS;
				break;
			}
		}
		const double dgs = ora_dot(g, s, n);
		double ggn = ora_dot(g, g, n);
		double ggp = ora_dot(g, gprev, n);
		double beta = (ggn - ggp) / gg;
		if (beta < 0)
			beta = 0;
		/* slope of the new direction -g + beta s, from the dots above */
		double dgn = -ggn + beta * dgs;
		if (dgn >= 0) {
			for (long i = 0; i < n; i++)
				s[i] = -g[i];
			dgn = -ggn;
			steepest = 1;
		} else {
			const float fb = (float)beta;
			for (long i = 0; i < n; i++)
				s[i] = -g[i] + fb * s[i];
			steepest = beta == 0.0;
		}
		alpha = alpha * dg0 / dgn;
		dg0 = dgn;
		gg = ggn;
	}
done:
	*fx = f;
	if (iterations)
		*iterations = it;
	if (evaluations)
		*evaluations = nevals;
	free(g);
	free(gprev);
	free(s);
	free(hist);
	return ret;
}

typedef struct {
	const unsigned char *msa;
	const float *w;
	int ncol, nrow;
	float l1, l2;
} ora_pll_ctx;

static double ora_pll_eval(void *vctx, const float *x, float *g, int n)
{
	(void)n;
	ora_pll_ctx *c = (ora_pll_ctx *)vctx;
	return ora_evaluate_f32(c->msa, c->w, c->ncol, c->nrow, c->l1, c->l2, x, g);
}

int ora_cg_pll_f32(const unsigned char *msa, const float *w, int ncol, int nrow,
                   float lambda_single, float lambda_pair, float *x, const ora_cg_param *p,
                   double *fx, int *iterations, int *evaluations, double *fhist)
{
	ora_pll_ctx c = {msa, w, ncol, nrow, lambda_single, lambda_pair};
	return ora_cg((int)ora_nvar_padded(ncol), x, fx, ora_pll_eval, &c, p, iterations,
	              evaluations, fhist);
}

/* ---- tied central differences (numdif.c:5-57 methodology) -------------- */

/* affected logits of one sequence: column, residue row, shift */
typedef struct {
	long col;
	int res;
	double delta;
} ora_shift;

static double ora_term_f64(const double *z21, int xobs)
{
	double m = z21[0];
	for (int a = 1; a < ORA_NA; a++)
		if (z21[a] > m)
			m = z21[a];
	double s = 0;
	for (int a = 0; a < ORA_NA; a++)
		s += exp(z21[a] - m);
	return m + log(s) - z21[xobs];
}

/*
 * out[i] = (f(x + h u) - f(x - h u)) / (2h) in FP64 with u = e_{idx_a[i]} +
 * e_{idx_b[i]} (a single unit vector when idx_a[i] == idx_b[i]), the
 * perturbation numdif.c applies to a coupling and its mirror.  Only the
 * pseudo-likelihood terms whose logits move are recomputed, so each probe
 * costs O(N * 21) after one O(N * L * 21L) logit pass.
 */
void ora_tied_fd_f64(const unsigned char *msa, const double *w, int ncol, int nrow,
                     double lambda_single, double lambda_pair, const double *x, const long *idx_a,
                     const long *idx_b, int count, double h, double *out)
{
	const long L = ncol, N = nrow;
	const long nsingle = L * (ORA_NA - 1);
	const long nsp = ora_nsingle_padded(ncol);
	const double *x1 = x, *x2 = x + nsp;
	double *Z = (double *)malloc(sizeof(double) * (size_t)N * ORA_NA * L); /* [n][a][j] */
	for (long n = 0; n < N; n++) {
		double *zn = Z + n * ORA_NA * L;
		const unsigned char *seq = msa + n * L;
		for (long a = 0; a < ORA_NA; a++)
			for (long j = 0; j < L; j++)
				zn[a * L + j] = a < ORA_NA - 1 ? x1[a * L + j] : 0.0;
		for (long k = 0; k < L; k++) {
			const double *row = x2 + ORA_X2(seq[k], k, 0, 0, L);
			for (long c = 0; c < ORA_NA * L; c++)
				zn[c] += row[c];
		}
	}
	for (int p = 0; p < count; p++) {
		long ids[2] = {idx_a[p], idx_b[p]};
		int nid = idx_a[p] == idx_b[p] ? 1 : 2;
		double total = 0.0;
		/* regularisation change */
		for (int q = 0; q < nid; q++) {
			const long v = ids[q];
			const double xv = x[v];
			const double dplus = (xv + h) * (xv + h) - xv * xv;
			const double dminus = (xv - h) * (xv - h) - xv * xv;
			if (v < nsingle)
				total += lambda_single * (dplus - dminus);
			else if (v >= nsp)
				total += 0.5 * lambda_pair * (dplus - dminus);
		}
		/* decode the likelihood effect of each index */
		int kind[2];
		long colv[2], kcol[2];
		int resv[2], bres[2];
		for (int q = 0; q < nid; q++) {
			const long v = ids[q];
			kind[q] = 0;
			if (v < nsingle) {
				kind[q] = 1;
				resv[q] = (int)(v / L);
				colv[q] = v % L;
			} else if (v >= nsp) {
				long off = v - nsp;
				const long j = off % L;
				off /= L;
				const long a = off % ORA_NAP;
				off /= ORA_NAP;
				const long k = off % L;
				const long b = off / L;
				if (a < ORA_NA) {
					kind[q] = 2;
					resv[q] = (int)a;
					colv[q] = j;
					kcol[q] = k;
					bres[q] = (int)b;
				}
			}
		}
		for (long n = 0; n < N; n++) {
			const unsigned char *seq = msa + n * L;
			ora_shift sh[2];
			int ns = 0;
			for (int q = 0; q < nid; q++) {
				if (kind[q] == 1 || (kind[q] == 2 && seq[kcol[q]] == bres[q])) {
					sh[ns].col = colv[q];
					sh[ns].res = resv[q];
					sh[ns].delta = 1.0;
					ns++;
				}
			}
			if (!ns)
				continue;
			const double *zn = Z + n * ORA_NA * L;
			/* distinct affected columns */
			long cols[2];
			int nc = 0;
			for (int s = 0; s < ns; s++) {
				int dup = 0;
				for (int c = 0; c < nc; c++)
					dup |= cols[c] == sh[s].col;
				if (!dup)
					cols[nc++] = sh[s].col;
			}
			for (int c = 0; c < nc; c++) {
				double zp[ORA_NA], zm[ORA_NA];
				for (int a = 0; a < ORA_NA; a++)
					zp[a] = zm[a] = zn[a * L + cols[c]];
				for (int s = 0; s < ns; s++)
					if (sh[s].col == cols[c]) {
						zp[sh[s].res] += h * sh[s].delta;
						zm[sh[s].res] -= h * sh[s].delta;
					}
				const int xo = seq[cols[c]];
				total += w[n] * (ora_term_f64(zp, xo) - ora_term_f64(zm, xo));
			}
		}
		out[p] = total / (2.0 * h);
	}
	free(Z);
}

typedef struct {
	const unsigned char *msa;
	const double *w;
	int ncol, nrow;
	double l1, l2;
	double *xd, *gd;
	int par;
} ora_pll64_ctx;

static double ora_pll_eval64(void *vctx, const float *x, float *g, int n)
{
	ora_pll64_ctx *c = (ora_pll64_ctx *)vctx;
	#pragma omp parallel for if (c->par)
	for (long i = 0; i < n; i++)
		c->xd[i] = x[i];
	double f = c->par ? ora_evaluate_f64_par(c->msa, c->w, c->ncol, c->nrow, c->l1, c->l2, c->xd,
	                                         c->gd)
	                  : ora_evaluate_f64(c->msa, c->w, c->ncol, c->nrow, c->l1, c->l2, c->xd, c->gd);
	#pragma omp parallel for if (c->par)
	for (long i = 0; i < n; i++)
		g[i] = (float)c->gd[i];
	return f;
}

static int cg_pll_f64(const unsigned char *msa, const float *w, int ncol, int nrow,
                      float lambda_single, float lambda_pair, float *x, const ora_cg_param *p,
                      double *fx, int *iterations, int *evaluations, double *fhist, int par)
{
	const long n = ora_nvar_padded(ncol);
	double *wd = (double *)malloc(sizeof(double) * (size_t)nrow);
	for (long i = 0; i < nrow; i++)
		wd[i] = w[i];
	ora_pll64_ctx c = {msa, wd, ncol, nrow, lambda_single, lambda_pair,
	                   (double *)malloc(sizeof(double) * (size_t)n),
	                   (double *)malloc(sizeof(double) * (size_t)n), par};
	int r = ora_cg((int)n, x, fx, ora_pll_eval64, &c, p, iterations, evaluations, fhist);
	free(c.xd);
	free(c.gd);
	free(wd);
	return r;
}

int ora_cg_pll_f64(const unsigned char *msa, const float *w, int ncol, int nrow,
                   float lambda_single, float lambda_pair, float *x, const ora_cg_param *p,
                   double *fx, int *iterations, int *evaluations, double *fhist)
{
	return cg_pll_f64(msa, w, ncol, nrow, lambda_single, lambda_pair, x, p, fx, iterations,
	                  evaluations, fhist, 0);
}

int ora_cg_pll_f64par(const unsigned char *msa, const float *w, int ncol, int nrow,
                      float lambda_single, float lambda_pair, float *x, const ora_cg_param *p,
                      double *fx, int *iterations, int *evaluations, double *fhist)
{
	return cg_pll_f64(msa, w, ncol, nrow, lambda_single, lambda_pair, x, p, fx, iterations,
	                  evaluations, fhist, 1);
}
