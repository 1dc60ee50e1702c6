/*
 * evaluate_hip: the PLL objective and tied gradient on one MI355X.
 *
 * Drop-in for the reference's CUDA backend (evaluate_cuda.c:23-151): same
 * callback signature, same padded x/g layout, device-resident x and g.
 * Differences by design: per-instance state lives in ud->extra (no
 * globals), nothing is allocated per evaluation, all steps run on one
 * stream with a single synchronisation for the f read-back, and init fills
 * the host weight array.
 *
 * Plain FP32 path per evaluation:
 *   gather_fwd   Z[n][a*L+j] = V + sum_k W(x_nk,k,a,j)
 *   softmax_rows D = w (P - onehot) in place; f and g1 partials
 *   gather_grad  G[(b,k)][(a,j)] = sum_n [x_nk=b] D[n][(a,j)]
 *   epilogue     g2 = G + G^T, self-edges 0, + 2 lambda_pair x
 * BF16 path per evaluation:
 *   convert_wt   Wt = bf16(W)^T in the padded (21*Lp)^2 index space
 *   gemm_fwd     Zt[c][n] on MFMA with the one-hot operand built in registers
 *   softmax_cols Dt = bf16(w P) (+ lo half); f and g1 partials
 *   gemm_grad    Gt[c][r] = sum_n Dt[c][n] [x_nk=b] on MFMA
 *   epilogue     g2 = Gt + Gt^T + hsym + 2 lambda_pair x, hsym = -2 H2 exact
 */
#define __HIP_PLATFORM_AMD__
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccmpred_amd.h"
#include "ccmpred_amd_kernels.h"

#define CCM_TIMING_STEPS 6

typedef struct ccm_hip_state {
	ccm_hip_config cfg;
	hipStream_t stream;
	int L, N, Lp, P, Np;
	long nsp, nvp;
	float lam1, lam2;
	unsigned char *d_msap;  /* Np x Lp, 255 past the real rows/columns */
	unsigned char *d_msapt; /* Lp x Np, 255 past the real rows/columns */
	float *d_w;
	/* plain FP32 path */
	float *d_zd;
	long ldz;
	float *d_graw;
	long ldg;
	/* BF16 path */
	uint16_t *d_wt;
	long ldwt;
	float *d_zt;
	uint16_t *d_dt;
	long lddt;
	float *d_gt;
	float *d_hsym;
	/* reductions */
	double *d_fpart;
	int n_fpart;
	double *d_g1part;
	int n_g1blk;
	double *d_regpair;
	int n_regpair;
	double *d_regsingle;
	int n_regsingle;
	double *d_f;
	double *h_f;
	/* optional per-step timing with HIP events on the instance stream */
	int timing;
	hipEvent_t ev[CCM_TIMING_STEPS + 1];
	double step_ms[CCM_TIMING_STEPS];
	int timed_calls;
} ccm_hip_state;

#define HIPCHK(call)                                                                       \
	do {                                                                               \
		hipError_t e_ = (call);                                                    \
		if (e_ != hipSuccess) {                                                    \
			fprintf(stderr, "ccmpred_amd: HIP error %d (%s) at %s:%d\n", (int)e_,   \
			        hipGetErrorString(e_), __FILE__, __LINE__);                \
			goto fail;                                                         \
		}                                                                          \
	} while (0)

#define KCHK(call)                                                                         \
	do {                                                                               \
		int r_ = (call);                                                           \
		if (r_ != 0) {                                                             \
			fprintf(stderr, "ccmpred_amd: launch failed (%d) at %s:%d\n", r_,       \
			        __FILE__, __LINE__);                                       \
			goto fail;                                                         \
		}                                                                          \
	} while (0)

long ccm_nsingle_padded(int ncol)
{
	long ns = (long)ncol * (CCM_N_ALPHA - 1);
	return ns + CCM_N_ALPHA_PAD - (ns % CCM_N_ALPHA_PAD);
}

long ccm_nvar_padded(int ncol)
{
	return ccm_nsingle_padded(ncol) + (long)ncol * ncol * CCM_N_ALPHA * CCM_N_ALPHA_PAD;
}

void ccm_hip_default_config(ccm_hip_config *cfg)
{
	cfg->device = 0;
	cfg->precision = CCM_PRECISION_FP32;
	cfg->bf16_passes = 1;
	cfg->verbose = 1;
}

static void free_state(ccm_hip_state *st)
{
	if (!st)
		return;
	void *bufs[] = {st->d_msap, st->d_msapt, st->d_w,  st->d_zd,     st->d_graw,
	                st->d_wt,   st->d_zt,    st->d_dt, st->d_gt,     st->d_hsym,
	                st->d_fpart, st->d_g1part, st->d_regpair, st->d_regsingle, st->d_f};
	for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++)
		if (bufs[i])
			(void)hipFree(bufs[i]);
	if (st->h_f)
		(void)hipHostFree(st->h_f);
	if (st->timing)
		for (int i = 0; i <= CCM_TIMING_STEPS; i++)
			(void)hipEventDestroy(st->ev[i]);
	if (st->stream)
		(void)hipStreamDestroy(st->stream);
	free(st);
}

static long round_up(long v, long m) { return (v + m - 1) / m * m; }

/* reweighting.c:10 -- float product, ceil in double, truncate */
static int identity_threshold(int ncol, float theta)
{
	float prod = theta * (float)(uint64_t)ncol;
	return (int)ceil((double)prod);
}

int init_hip_ex(void *instance, const ccm_hip_config *cfg_in)
{
	userdata *ud = (userdata *)instance;
	ccm_hip_config cfg;
	if (cfg_in)
		cfg = *cfg_in;
	else
		ccm_hip_default_config(&cfg);
	if (cfg.bf16_passes < 1 || cfg.bf16_passes > 2 || ud->ncol <= 0 || ud->nrow <= 0) {
		fprintf(stderr, "ccmpred_amd: invalid configuration\n");
		return 1;
	}
	ccm_hip_state *st = (ccm_hip_state *)calloc(1, sizeof(*st));
	unsigned char *h_msap = NULL, *h_msapt = NULL;
	uint32_t *d_planes = NULL;
	int32_t *d_counts = NULL;
	float *d_tmp_d = NULL, *d_tmp_g = NULL;
	if (!st)
		return 1;
	st->cfg = cfg;
	const int L = ud->ncol, N = ud->nrow;
	st->L = L;
	st->N = N;
	st->Lp = (int)round_up(L, 32);
	st->Np = (int)round_up(N, 128);
	st->P = CCM_N_ALPHA * st->Lp;
	st->nsp = ccm_nsingle_padded(L);
	st->nvp = ccm_nvar_padded(L);
	st->lam1 = ud->lambda_single;
	st->lam2 = ud->lambda_pair;
	const int Lp = st->Lp, Np = st->Np, P = st->P;

	HIPCHK(hipSetDevice(cfg.device));
	HIPCHK(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
	hipStream_t s = st->stream;

	/* padded MSA in both orientations */
	h_msap = (unsigned char *)malloc((size_t)Np * Lp);
	h_msapt = (unsigned char *)malloc((size_t)Lp * Np);
	if (!h_msap || !h_msapt)
		goto fail;
	memset(h_msap, 255, (size_t)Np * Lp);
	memset(h_msapt, 255, (size_t)Lp * Np);
	for (long n = 0; n < N; n++)
		for (long k = 0; k < L; k++) {
			unsigned char v = ud->msa[n * L + k];
			h_msap[n * Lp + k] = v;
			h_msapt[k * Np + n] = v;
		}
	HIPCHK(hipMalloc((void **)&st->d_msap, (size_t)Np * Lp));
	HIPCHK(hipMalloc((void **)&st->d_msapt, (size_t)Lp * Np));
	HIPCHK(hipMemcpyAsync(st->d_msap, h_msap, (size_t)Np * Lp, hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(st->d_msapt, h_msapt, (size_t)Lp * Np, hipMemcpyHostToDevice, s));

	/* sequence weights */
	HIPCHK(hipMalloc((void **)&st->d_w, sizeof(float) * (size_t)N));
	if (ud->reweighting_threshold != 1.0f) {
		int idthres = identity_threshold(L, ud->reweighting_threshold);
		unsigned char *d_msa_rows = NULL;
		/* the identity kernel reads an unpadded N x L matrix */
		HIPCHK(hipMalloc((void **)&d_msa_rows, (size_t)N * L));
		HIPCHK(hipMemcpyAsync(d_msa_rows, ud->msa, (size_t)N * L, hipMemcpyHostToDevice, s));
		HIPCHK(hipMalloc((void **)&d_planes,
		                 sizeof(uint32_t) * (size_t)N * 5 * ccm_reweight_words(L)));
		HIPCHK(hipMalloc((void **)&d_counts, sizeof(int32_t) * (size_t)N));
		int rc = ccm_launch_identity_counts(d_msa_rows, L, N, idthres, d_planes, d_counts, s);
		if (rc == 0)
			rc = ccm_launch_weights_from_counts(d_counts, N, L, idthres, st->d_w, s);
		HIPCHK(hipStreamSynchronize(s));
		(void)hipFree(d_msa_rows);
		KCHK(rc);
	} else {
		KCHK(ccm_launch_fill_f32(st->d_w, N, 1.0f, s));
	}
	if (ud->weights) {
		HIPCHK(hipMemcpyAsync(ud->weights, st->d_w, sizeof(float) * (size_t)N,
		                      hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		if (cfg.verbose) {
			double wsum = 0;
			float wmin = ud->weights[0], wmax = ud->weights[0];
			for (int i = 0; i < N; i++) {
				wsum += ud->weights[i];
				if (ud->weights[i] < wmin)
					wmin = ud->weights[i];
				if (ud->weights[i] > wmax)
					wmax = ud->weights[i];
			}
			if (ud->reweighting_threshold != 1.0f)
				printf("Reweighted %d sequences with threshold %.1f to Beff=%g weight mean=%g, "
				       "min=%g, max=%g\n",
				       N, ud->reweighting_threshold, wsum, wsum / N, wmin, wmax);
			else
				printf("Using uniform weights on %d sequences\n", N);
		}
	}

	/* reduction scratch */
	st->n_regpair = ccm_pair_epilogue_blocks(L);
	st->n_regsingle = ccm_singles_blocks(L);
	HIPCHK(hipMalloc((void **)&st->d_regpair, sizeof(double) * (size_t)st->n_regpair));
	HIPCHK(hipMalloc((void **)&st->d_regsingle, sizeof(double) * (size_t)st->n_regsingle));
	HIPCHK(hipMalloc((void **)&st->d_f, sizeof(double) * 4));
	HIPCHK(hipHostMalloc((void **)&st->h_f, sizeof(double) * 4, hipHostMallocDefault));

	if (cfg.precision == CCM_PRECISION_FP32) {
		st->ldz = round_up((long)CCM_N_ALPHA * L, 64);
		st->ldg = st->ldz;
		HIPCHK(hipMalloc((void **)&st->d_zd, sizeof(float) * (size_t)N * st->ldz));
		HIPCHK(hipMalloc((void **)&st->d_graw,
		                 sizeof(float) * (size_t)CCM_N_ALPHA * L * st->ldg));
		st->n_fpart = ccm_softmax_rows_blocks(L, N);
		st->n_g1blk = ccm_softmax_rows_nblk(N);
	} else {
		const int passes = cfg.bf16_passes;
		st->ldwt = (long)passes * P;
		st->lddt = (long)passes * Np;
		HIPCHK(hipMalloc((void **)&st->d_wt, sizeof(uint16_t) * (size_t)P * st->ldwt));
		HIPCHK(hipMalloc((void **)&st->d_zt, sizeof(float) * (size_t)P * Np));
		HIPCHK(hipMalloc((void **)&st->d_dt, sizeof(uint16_t) * (size_t)P * st->lddt));
		HIPCHK(hipMemsetAsync(st->d_dt, 0, sizeof(uint16_t) * (size_t)P * st->lddt, s));
		HIPCHK(hipMalloc((void **)&st->d_gt, sizeof(float) * (size_t)P * P));
		HIPCHK(hipMalloc((void **)&st->d_hsym,
		                 sizeof(float) * (size_t)CCM_N_ALPHA * L * CCM_N_ALPHA_PAD * L));
		st->n_fpart = ccm_softmax_cols_blocks(L, N);
		st->n_g1blk = ccm_softmax_cols_nblk(N);

		/* exact data term: hsym = -(H2 + H2^T), H2 = A^T diag(w) A, computed
		 * once with the FP32 scatter kernel (cu:350-387 analogue) */
		long ld = round_up((long)CCM_N_ALPHA * L, 64);
		HIPCHK(hipMalloc((void **)&d_tmp_d, sizeof(float) * (size_t)N * ld));
		HIPCHK(hipMalloc((void **)&d_tmp_g, sizeof(float) * (size_t)CCM_N_ALPHA * L * ld));
		KCHK(ccm_launch_neg_weighted_onehot(st->d_msap, Lp, st->d_w, L, N, d_tmp_d, ld, s));
		KCHK(ccm_launch_gather_grad_f32(d_tmp_d, ld, st->d_msap, Lp, L, N, d_tmp_g, ld, s));
		KCHK(ccm_launch_hsym(d_tmp_g, ld, st->d_hsym, L, s));
		HIPCHK(hipStreamSynchronize(s));
	}
	HIPCHK(hipMalloc((void **)&st->d_fpart, sizeof(double) * (size_t)st->n_fpart));
	HIPCHK(hipMalloc((void **)&st->d_g1part,
	                 sizeof(double) * (size_t)st->n_g1blk * 20 * (size_t)L));
	HIPCHK(hipStreamSynchronize(s));

	if (d_planes)
		(void)hipFree(d_planes);
	if (d_counts)
		(void)hipFree(d_counts);
	if (d_tmp_d)
		(void)hipFree(d_tmp_d);
	if (d_tmp_g)
		(void)hipFree(d_tmp_g);
	free(h_msap);
	free(h_msapt);
	ud->extra = st;
	return 0;

fail:
	if (d_planes)
		(void)hipFree(d_planes);
	if (d_counts)
		(void)hipFree(d_counts);
	if (d_tmp_d)
		(void)hipFree(d_tmp_d);
	if (d_tmp_g)
		(void)hipFree(d_tmp_g);
	free(h_msap);
	free(h_msapt);
	free_state(st);
	ud->extra = NULL;
	return 1;
}

int init_hip(void *instance) { return init_hip_ex(instance, NULL); }

int init_hip_bf16(void *instance)
{
	ccm_hip_config cfg;
	ccm_hip_default_config(&cfg);
	cfg.precision = CCM_PRECISION_BF16;
	return init_hip_ex(instance, &cfg);
}

int destroy_hip(void *instance)
{
	userdata *ud = (userdata *)instance;
	free_state((ccm_hip_state *)ud->extra);
	ud->extra = NULL;
	return 0;
}

void *ccm_hip_stream(void *instance)
{
	userdata *ud = (userdata *)instance;
	ccm_hip_state *st = (ccm_hip_state *)ud->extra;
	return st ? (void *)st->stream : NULL;
}

#define MARK(i)                                                            \
	do {                                                               \
		if (st->timing)                                            \
			HIPCHK(hipEventRecord(st->ev[(i)], s));            \
	} while (0)

int ccm_hip_evaluate(void *instance, const float *d_x, float *d_g, int nvar, double *f)
{
	userdata *ud = (userdata *)instance;
	ccm_hip_state *st = (ccm_hip_state *)ud->extra;
	if (!st || nvar != st->nvp) {
		fprintf(stderr, "ccmpred_amd: evaluate called with nvar=%d, expected %ld\n", nvar,
		        st ? st->nvp : -1L);
		return -1;
	}
	hipStream_t s = st->stream;
	const int L = st->L, N = st->N, Lp = st->Lp, Np = st->Np;
	int step = 0;
	if (st->cfg.precision == CCM_PRECISION_FP32) {
		MARK(step++);
		KCHK(ccm_launch_gather_fwd_f32(st->d_msap, Lp, d_x, L, N, st->d_zd, st->ldz, s));
		MARK(step++);
		KCHK(ccm_launch_softmax_rows(st->d_zd, st->ldz, st->d_msap, Lp, st->d_w, L, N,
		                             st->d_fpart, st->d_g1part, s));
		MARK(step++);
		KCHK(ccm_launch_gather_grad_f32(st->d_zd, st->ldz, st->d_msap, Lp, L, N, st->d_graw,
		                                st->ldg, s));
		MARK(step++);
		KCHK(ccm_launch_pair_epilogue(st->d_graw, st->ldg, L, NULL, d_x, d_g, L, st->lam2,
		                              st->d_regpair, s));
	} else {
		const int passes = st->cfg.bf16_passes;
		MARK(step++);
		KCHK(ccm_launch_convert_wt(d_x, L, Lp, st->d_wt, st->ldwt, passes, s));
		MARK(step++);
		KCHK(ccm_launch_gemm_fwd(st->d_wt, st->ldwt, Lp, passes, st->d_msap, Np, st->d_zt,
		                         (long)Np, s));
		MARK(step++);
		KCHK(ccm_launch_softmax_cols(st->d_zt, (long)Np, d_x, st->d_msapt, Np, st->d_w, L, Lp,
		                             N, st->d_dt, st->lddt, passes, st->d_fpart, st->d_g1part,
		                             s));
		MARK(step++);
		KCHK(ccm_launch_gemm_grad(st->d_dt, st->lddt, Lp, passes, st->d_msapt, Np, st->d_gt,
		                          (long)st->P, s));
		MARK(step++);
		KCHK(ccm_launch_pair_epilogue(st->d_gt, (long)st->P, Lp, st->d_hsym, d_x, d_g, L,
		                              st->lam2, st->d_regpair, s));
	}
	MARK(step++);
	KCHK(ccm_launch_singles_finalize(st->d_g1part, st->n_g1blk, d_x, d_g, L, st->lam1,
	                                 st->d_regsingle, s));
	KCHK(ccm_launch_sum3(st->d_fpart, st->n_fpart, st->d_regpair, st->n_regpair,
	                     st->d_regsingle, st->n_regsingle, st->d_f, s));
	MARK(step);
	HIPCHK(hipMemcpyAsync(st->h_f, st->d_f, sizeof(double), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	*f = st->h_f[0];
	if (st->timing) {
		for (int i = 0; i < step; i++) {
			float ms = 0.f;
			HIPCHK(hipEventElapsedTime(&ms, st->ev[i], st->ev[i + 1]));
			st->step_ms[i] += ms;
		}
		st->timed_calls++;
	}
	return 0;
fail:
	return -1;
}

int ccm_hip_set_timing(void *instance, int on)
{
	userdata *ud = (userdata *)instance;
	ccm_hip_state *st = (ccm_hip_state *)ud->extra;
	if (!st)
		return -1;
	if (on && !st->timing) {
		for (int i = 0; i <= CCM_TIMING_STEPS; i++)
			HIPCHK(hipEventCreate(&st->ev[i]));
		st->timing = 1;
	}
	for (int i = 0; i < CCM_TIMING_STEPS; i++)
		st->step_ms[i] = 0.0;
	st->timed_calls = 0;
	return 0;
fail:
	return -1;
}

int ccm_hip_get_timing(void *instance, double *avg_ms, int n)
{
	userdata *ud = (userdata *)instance;
	ccm_hip_state *st = (ccm_hip_state *)ud->extra;
	if (!st)
		return -1;
	for (int i = 0; i < n && i < CCM_TIMING_STEPS; i++)
		avg_ms[i] = st->timed_calls ? st->step_ms[i] / st->timed_calls : 0.0;
	return st->timed_calls;
}

conjugrad_float_t evaluate_hip(void *instance, const conjugrad_float_t *d_x, conjugrad_float_t *d_g,
                               const int nvar)
{
	double f = 0;
	if (ccm_hip_evaluate(instance, d_x, d_g, nvar, &f) != 0)
		return NAN;
	return (conjugrad_float_t)f;
}
