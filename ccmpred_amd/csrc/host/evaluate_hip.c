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
 * Sequences are processed in chunks (all of them by default): per chunk the
 * logits and probabilities are formed and the pair gradient accumulated, so
 * the N x 21L intermediate never exceeds one chunk.
 *
 * Plain FP32 path, per chunk:
 *   gather_fwd   Z[n][a*L+j] = V + sum_k W(x_nk,k,a,j)
 *   softmax_rows D = w (P - onehot) in place; f and g1 partials
 *   gather_grad  G[(b,k)][(a,j)] (+)= sum_n [x_nk=b] D[n][(a,j)]
 * then epilogue  g2 = G + G^T, self-edges 0, + 2 lambda_pair x.
 * BF16 path: convert_wt Wt = bf16(W)^T once, then per chunk
 *   gemm_fwd     Zt[c][n] on MFMA with the one-hot operand built in registers
 *   softmax_cols Dt = bf16(w P) (+ lo half); f and g1 partials
 *   gemm_grad    Gt[c][r] (+)= sum_n Dt[c][n] [x_nk=b] on MFMA
 * then epilogue  g2 = Gt + Gt^T + hsym + 2 lambda_pair x, hsym = -2 H2 exact.
 */
#define __HIP_PLATFORM_AMD__
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccmpred_amd.h"
#include "ccmpred_amd_kernels.h"

#include "state.h"

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
	cfg->bf16_fwd_passes = 0;
	cfg->verbose = 1;
	cfg->chunk_rows = 0;
	cfg->gemm_variant = -1;
	cfg->sparse_fwd = 1;
}

static hipError_t dalloc(ccm_hip_state *st, void **p, size_t bytes)
{
	hipError_t e = hipMalloc(p, bytes);
	if (e == hipSuccess)
		st->bytes += bytes;
	return e;
}

static void free_state(ccm_hip_state *st)
{
	if (!st)
		return;
	void *bufs[] = {st->d_msap, st->d_msapt, st->d_segmeta, st->d_w, st->d_zd, st->d_graw,
	                st->d_wt,   st->d_sidx,  st->d_zt, st->d_dt,     st->d_gt, st->d_hsym,
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

/* Records of the segmented-gather pair gradient: for every chunk of 128
 * sequences and every column k, the chunk's rows grouped by residue
 * (ascending, each group padded to a multiple of 4 with the zero row 128)
 * in bytes [0, 192) and the group starts in bytes [192, 214). */
static void build_seg_meta(const unsigned char *msa, int L, int N, unsigned char *meta)
{
	const int nch = (N + 127) / 128;
	for (int c = 0; c < nch; c++) {
		const int r0 = c * 128, nr = N - r0 < 128 ? N - r0 : 128;
		for (int k = 0; k < L; k++) {
			unsigned char *rec = meta + ((size_t)c * L + k) * 256;
			memset(rec, 0, 256);
			int pos = 0;
			for (int b = 0; b < CCM_N_ALPHA; b++) {
				rec[192 + b] = (unsigned char)pos;
				for (int r = 0; r < nr; r++)
					if (msa[(size_t)(r0 + r) * L + k] == b)
						rec[pos++] = (unsigned char)r;
				while (pos & 3)
					rec[pos++] = 128;
			}
			rec[192 + CCM_N_ALPHA] = (unsigned char)pos;
		}
	}
}

/* reweighting.c:10 -- float product, ceil in double, truncate */
static int identity_threshold(int ncol, float theta)
{
	float prod = theta * (float)(uint64_t)ncol;
	return (int)ceil((double)prod);
}

static const unsigned char *seg_chunk(const ccm_hip_state *st, int n0)
{
	return st->d_segmeta + (size_t)(n0 / 128) * st->L * 256;
}

int init_hip_ex(void *instance, const ccm_hip_config *cfg_in)
{
	return ccm_init_hip_state(instance, cfg_in, 1, 1);
}

int ccm_init_hip_state(void *instance, const ccm_hip_config *cfg_in, int allow_fused, int allow_sparse)
{
	userdata *ud = (userdata *)instance;
	ccm_hip_config cfg;
	if (cfg_in)
		cfg = *cfg_in;
	else
		ccm_hip_default_config(&cfg);
	if (cfg.bf16_fwd_passes == 0)
		cfg.bf16_fwd_passes = cfg.bf16_passes;
	if (cfg.bf16_passes < 1 || cfg.bf16_passes > 2 || cfg.bf16_fwd_passes < 1 ||
	    cfg.bf16_fwd_passes > 2 || cfg.chunk_rows < 0 || ud->ncol <= 0 || ud->nrow <= 0) {
		fprintf(stderr, "ccmpred_amd: invalid configuration\n");
		return 1;
	}
	ccm_hip_state *st = (ccm_hip_state *)calloc(1, sizeof(*st));
	unsigned char *h_msap = NULL, *h_msapt = NULL, *d_msa_rows = NULL;
	uint32_t *d_planes = NULL;
	int32_t *d_counts = NULL;
	float *d_tmp_d = NULL, *d_tmp_g = NULL, *h_w = NULL;
	double *h_c = NULL;
	if (!st)
		return 1;
	st->cfg = cfg;
	const int L = ud->ncol, N = ud->nrow;
	st->L = L;
	st->N = N;
	st->Lp = (int)round_up(L, 64);
	st->Np = (int)round_up(N, 128);
	st->P = CCM_N_ALPHA * st->Lp;
	st->nsp = ccm_nsingle_padded(L);
	st->nvp = ccm_nvar_padded(L);
	st->lam1 = ud->lambda_single;
	st->lam2 = ud->lambda_pair;
	/* chunks are whole blocks of the softmax kernel of the precision form */
	const int gran = cfg.precision == CCM_PRECISION_BF16 ? ccm_softmax_cols_granule() : 256;
	st->chunk = cfg.chunk_rows > 0 ? (int)round_up(cfg.chunk_rows, gran) : (int)round_up(N, gran);
	if (st->chunk >= N)
		st->chunk = (int)round_up(N, gran);
	st->ccols = (int)round_up(st->chunk < N ? st->chunk : N, 128);
	const int Lp = st->Lp, Np = st->Np, P = st->P;
	const int chunk_rows = st->chunk < N ? st->chunk : N;

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
	HIPCHK(dalloc(st, (void **)&st->d_msap, (size_t)Np * Lp));
	HIPCHK(dalloc(st, (void **)&st->d_msapt, (size_t)Lp * Np));
	HIPCHK(hipMemcpyAsync(st->d_msap, h_msap, (size_t)Np * Lp, hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(st->d_msapt, h_msapt, (size_t)Lp * Np, hipMemcpyHostToDevice, s));

	{
		const size_t mb = (size_t)ccm_seg_meta_bytes(L, N);
		unsigned char *h_meta = (unsigned char *)malloc(mb);
		if (!h_meta)
			goto fail;
		build_seg_meta(ud->msa, L, N, h_meta);
		if (dalloc(st, (void **)&st->d_segmeta, mb) != hipSuccess ||
		    hipMemcpy(st->d_segmeta, h_meta, mb, hipMemcpyHostToDevice) != hipSuccess) {
			free(h_meta);
			goto fail;
		}
		free(h_meta);
	}

	/* sequence weights (reweighting.c:9-67, bit exact) */
	HIPCHK(dalloc(st, (void **)&st->d_w, sizeof(float) * (size_t)N));
	if (ud->reweighting_threshold != 1.0f) {
		int idthres = identity_threshold(L, ud->reweighting_threshold);
		/* the identity kernel reads an unpadded N x L matrix */
		HIPCHK(hipMalloc((void **)&d_msa_rows, (size_t)N * L));
		HIPCHK(hipMemcpyAsync(d_msa_rows, ud->msa, (size_t)N * L, hipMemcpyHostToDevice, s));
		HIPCHK(hipMalloc((void **)&d_planes,
		                 sizeof(uint32_t) * (size_t)N * 5 * ccm_reweight_words(L)));
		HIPCHK(hipMalloc((void **)&d_counts, sizeof(int32_t) * (size_t)N));
		KCHK(ccm_launch_identity_counts(d_msa_rows, L, N, idthres, d_planes, d_counts, s));
		KCHK(ccm_launch_weights_from_counts(d_counts, N, L, idthres, st->d_w, s));
	} else {
		KCHK(ccm_launch_fill_f32(st->d_w, N, 1.0f, s));
	}
	HIPCHK(hipStreamSynchronize(s));
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
	if (st->n_regpair < ccm_reg_only_blocks())
		st->n_regpair = ccm_reg_only_blocks();
	/* the sparse form's conversion leaves its own partials (either pass count) */
	for (int pf = 1; pf <= 2; pf++)
		if (st->n_regpair < ccm_convert_ws_reg_parts(L, st->Lp, pf))
			st->n_regpair = ccm_convert_ws_reg_parts(L, st->Lp, pf);
	st->n_regsingle = ccm_singles_blocks(L);
	HIPCHK(dalloc(st, (void **)&st->d_regpair, sizeof(double) * (size_t)st->n_regpair));
	HIPCHK(dalloc(st, (void **)&st->d_regsingle, sizeof(double) * (size_t)st->n_regsingle));
	HIPCHK(dalloc(st, (void **)&st->d_f, sizeof(double) * 4));
	HIPCHK(hipHostMalloc((void **)&st->h_f, sizeof(double) * 4, hipHostMallocMapped | hipHostMallocCoherent));
	HIPCHK(hipHostGetDevicePointer((void **)&st->dh_f, st->h_f, 0));

	if (cfg.precision == CCM_PRECISION_FP32) {
		st->ldz = round_up((long)CCM_N_ALPHA * L, 64);
		st->ldg = st->ldz;
		HIPCHK(dalloc(st, (void **)&st->d_zd, sizeof(float) * (size_t)chunk_rows * st->ldz));
		HIPCHK(dalloc(st, (void **)&st->d_graw,
		              sizeof(float) * (size_t)CCM_N_ALPHA * L * st->ldg));
		st->n_fpart = ccm_softmax_rows_blocks(L, N);
		st->n_g1blk = ccm_softmax_rows_nblk(N);
	} else {
		const int passes = cfg.bf16_passes, fpasses = cfg.bf16_fwd_passes;
		st->lddt = (long)passes * st->ccols;
		/* the fused forward addresses the gathered rows of Wt with 32-bit
		 * offsets; the sparse one exists in the fused form */
		st->fused = allow_fused && ccm_fwd_softmax_fits(Lp, fpasses, passes, st->ccols);
		st->sparse = st->fused && allow_sparse && cfg.sparse_fwd &&
		             ccm_sparse_fwd_fits(Lp, fpasses, passes, st->ccols) &&
		             (long)ccm_sparse_steps(Lp, fpasses) * 2 * ccm_sparse_index_ld(Np) * 4 <= 0x7FFFFFFFL;
		st->ldwt = st->sparse ? ccm_sparse_ws_len(Lp, fpasses) : (long)fpasses * P;
		HIPCHK(dalloc(st, (void **)&st->d_wt, sizeof(uint16_t) * (size_t)P * st->ldwt));
		/* the rows of padding columns j >= L stay zero: the conversions skip them */
		HIPCHK(hipMemsetAsync(st->d_wt, 0, sizeof(uint16_t) * (size_t)P * st->ldwt, s));
		if (st->sparse) {
			/* the one-hot positions of the sparse forward: constant per MSA */
			st->ldsidx = ccm_sparse_index_ld(Np);
			HIPCHK(dalloc(st, (void **)&st->d_sidx,
			              sizeof(uint32_t) * 2 * (size_t)ccm_sparse_steps(Lp, fpasses) * st->ldsidx));
			KCHK(ccm_launch_sparse_index(st->d_msap, Lp, Np, fpasses, st->d_sidx, s));
		}
		if (!st->fused)
			HIPCHK(dalloc(st, (void **)&st->d_zt, sizeof(float) * (size_t)P * st->ccols));
		HIPCHK(dalloc(st, (void **)&st->d_dt, sizeof(uint16_t) * (size_t)P * st->lddt));
		HIPCHK(hipMemsetAsync(st->d_dt, 0, sizeof(uint16_t) * (size_t)P * st->lddt, s));
		HIPCHK(dalloc(st, (void **)&st->d_gt, sizeof(float) * (size_t)P * P));
		HIPCHK(dalloc(st, (void **)&st->d_hsym,
		              sizeof(float) * (size_t)CCM_N_ALPHA * L * CCM_N_ALPHA_PAD * L));
		if (st->fused) {
			/* one block per 128 sequences, then the constant data term */
			st->n_fpart = ccm_fwd_softmax_blocks(N) * L;
			st->n_g1blk = ccm_fwd_softmax_blocks(N) + 1;
		} else {
			st->n_fpart = ccm_softmax_cols_blocks(L, N);
			st->n_g1blk = ccm_softmax_cols_nblk(N);
		}

		/* exact data term: hsym = -(H2 + H2^T), H2 = A^T diag(w) A, computed
		 * once with the FP32 scatter kernel over sequence chunks (the
		 * reference's init-time histograms, cu:350-387) */
		const long ld = round_up((long)CCM_N_ALPHA * L, 64);
		HIPCHK(hipMalloc((void **)&d_tmp_d, sizeof(float) * (size_t)chunk_rows * ld));
		HIPCHK(hipMalloc((void **)&d_tmp_g, sizeof(float) * (size_t)CCM_N_ALPHA * L * ld));
		for (int n0 = 0; n0 < N; n0 += st->chunk) {
			const int nc = N - n0 < st->chunk ? N - n0 : st->chunk;
			KCHK(ccm_launch_neg_weighted_onehot(st->d_msap + (size_t)n0 * Lp, Lp, st->d_w + n0,
			                                    L, nc, d_tmp_d, ld, s));
			KCHK(ccm_launch_gather_grad_seg_f32(d_tmp_d, ld, seg_chunk(st, n0), L, nc, d_tmp_g, ld,
			                                    n0 > 0, s));
		}
		KCHK(ccm_launch_hsym(d_tmp_g, ld, st->d_hsym, L, s));
	}
	HIPCHK(dalloc(st, (void **)&st->d_fpart, sizeof(double) * (size_t)st->n_fpart));
	HIPCHK(dalloc(st, (void **)&st->d_g1part,
	              sizeof(double) * (size_t)st->n_g1blk * 20 * (size_t)L));
	if (st->fused) {
		/* The fused forward sums only sum_n w_n P_n(a, j) per block; the data
		 * term -sum_n w_n [x_nj = a] of the node gradient is constant and
		 * sits in the last block of the partials (sequences in order, FP64) */
		h_w = (float *)malloc(sizeof(float) * (size_t)N);
		h_c = (double *)calloc((size_t)20 * L, sizeof(double));
		if (!h_w || !h_c)
			goto fail;
		HIPCHK(hipMemcpy(h_w, st->d_w, sizeof(float) * (size_t)N, hipMemcpyDeviceToHost));
		for (long n = 0; n < N; n++)
			for (long j = 0; j < L; j++) {
				const int a = ud->msa[n * L + j];
				if (a < 20)
					h_c[(long)a * L + j] -= (double)h_w[n];
			}
		HIPCHK(hipMemcpy(st->d_g1part + (size_t)(st->n_g1blk - 1) * 20 * L, h_c,
		                 sizeof(double) * (size_t)20 * L, hipMemcpyHostToDevice));
	}
	HIPCHK(hipStreamSynchronize(s));

	(void)hipFree(d_msa_rows);
	(void)hipFree(d_planes);
	(void)hipFree(d_counts);
	(void)hipFree(d_tmp_d);
	(void)hipFree(d_tmp_g);
	free(h_msap);
	free(h_msapt);
	free(h_w);
	free(h_c);
	ud->extra = st;
	return 0;

fail:
	if (st->stream)
		(void)hipStreamSynchronize(st->stream);
	(void)hipFree(d_msa_rows);
	(void)hipFree(d_planes);
	(void)hipFree(d_counts);
	(void)hipFree(d_tmp_d);
	(void)hipFree(d_tmp_g);
	free(h_msap);
	free(h_msapt);
	free(h_w);
	free(h_c);
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

double ccm_hip_objective_resolution(void *instance)
{
	const userdata *ud = (const userdata *)instance;
	const ccm_hip_state *st = ud ? (const ccm_hip_state *)ud->extra : NULL;
	if (st && st->cfg.precision == CCM_PRECISION_BF16 && st->cfg.bf16_fwd_passes == 1)
		return 1e-5;
	return 1e-6;
}

size_t ccm_hip_device_bytes(void *instance)
{
	userdata *ud = (userdata *)instance;
	ccm_hip_state *st = (ccm_hip_state *)ud->extra;
	return st ? st->bytes : 0;
}

#define MARK(i)                                                            \
	do {                                                               \
		if (st->timing)                                            \
			HIPCHK(hipEventRecord(st->ev[(i)], s));            \
	} while (0)

enum { RUN_OBJECTIVE = 0, RUN_FORWARD = 1, RUN_FULL = 2 };

static ccm_hip_state *get_state(void *instance, int nvar)
{
	userdata *ud = (userdata *)instance;
	ccm_hip_state *st = (ccm_hip_state *)ud->extra;
	if (!st || nvar != st->nvp) {
		fprintf(stderr, "ccmpred_amd: evaluate called with nvar=%d, expected %ld\n", nvar,
		        st ? st->nvp : -1L);
		return NULL;
	}
	return st;
}

/* The forward's coupling operand from x: Wt (ccm_launch_convert_wt), or the
 * sparse layout Ws (ccm_launch_convert_ws) */
static int convert_couplings(ccm_hip_state *st, const float *d_x, hipStream_t s)
{
	return st->sparse
	           ? ccm_launch_convert_ws(d_x, st->L, st->Lp, st->d_wt, st->ldwt, st->cfg.bf16_fwd_passes, s)
	           : ccm_launch_convert_wt(d_x, st->L, st->Lp, st->d_wt, st->ldwt, st->cfg.bf16_fwd_passes, s);
}

/* Logits, softmax and (RUN_FULL) pair-gradient accumulation over all
 * sequence chunks.  *step advances through the timing marks: one mark
 * before each of the three per-chunk kernels of the first chunk.  d_g_pad
 * (sparse form): the first chunk's forward also runs the coupling
 * conversion's padding tiles, writing g's a >= 21 padding. */
static int run_chunks(ccm_hip_state *st, const float *d_x, int what, int *step, float *d_g_pad)
{
	hipStream_t s = st->stream;
	const int L = st->L, N = st->N, Lp = st->Lp, Np = st->Np, P = st->P;
	const int mode = what == RUN_OBJECTIVE ? 0 : 1;
	for (int n0 = 0; n0 < N; n0 += st->chunk) {
		const int nc = N - n0 < st->chunk ? N - n0 : st->chunk;
		const int first = n0 == 0;
		const unsigned char *msap = st->d_msap + (size_t)n0 * Lp;
		if (st->cfg.precision == CCM_PRECISION_FP32) {
			if (first)
				MARK((*step)++);
			KCHK(ccm_launch_gather_fwd_f32(msap, Lp, d_x, L, nc, st->d_zd, st->ldz, s));
			if (first)
				MARK((*step)++);
			KCHK(ccm_launch_softmax_rows(st->d_zd, st->ldz, msap, Lp, st->d_w + n0, L, nc, n0,
			                             mode, st->d_fpart, st->d_g1part, s));
			if (first)
				MARK((*step)++);
			if (what == RUN_FULL)
				KCHK(ccm_launch_gather_grad_seg_f32(st->d_zd, st->ldz, seg_chunk(st, n0), L, nc,
				                                    st->d_graw, st->ldg, !first, s));
		} else {
			const int passes = st->cfg.bf16_passes;
			const int ncols = (int)round_up(nc, 128);
			if (first)
				MARK((*step)++);
			if (st->sparse) {
				const int pad = first && d_g_pad;
				KCHK(ccm_launch_sparse_fwd_softmax_pad(
				    st->d_wt, st->ldwt, Lp, st->cfg.bf16_fwd_passes, st->d_sidx + n0, st->ldsidx, ncols,
				    d_x, st->d_msapt + n0, (long)Np, st->d_w + n0, L, nc, st->d_dt, (long)passes * ncols,
				    (long)ncols, passes, mode, n0 / 128, st->d_fpart, st->d_g1part, pad ? d_x : NULL,
				    pad ? d_g_pad : NULL, st->lam2, pad ? st->d_regpair : NULL, s));
			} else if (st->fused) {
				KCHK(ccm_launch_gemm_fwd_softmax(st->d_wt, st->ldwt, Lp, st->cfg.bf16_fwd_passes,
				                                 msap, ncols, d_x, st->d_msapt + n0, (long)Np,
				                                 st->d_w + n0, L, nc, st->d_dt,
				                                 (long)passes * ncols, (long)ncols, passes, mode,
				                                 n0 / 128, st->d_fpart, st->d_g1part,
				                                 st->cfg.gemm_variant, s));
			} else {
				KCHK(ccm_launch_gemm_fwd(st->d_wt, st->ldwt, Lp, st->cfg.bf16_fwd_passes, msap,
				                         ncols, st->d_zt, (long)ncols, st->cfg.gemm_variant, s));
				if (first)
					MARK((*step)++);
				KCHK(ccm_launch_softmax_cols(st->d_zt, (long)ncols, d_x, st->d_msapt + n0,
				                             (long)Np, st->d_w + n0, L, Lp, nc, ncols, n0,
				                             st->d_dt, (long)passes * ncols, (long)ncols, passes,
				                             mode, st->d_fpart, st->d_g1part, s));
			}
			if (first)
				MARK((*step)++);
			if (what == RUN_FULL)
				KCHK(ccm_launch_gemm_grad(st->d_dt, (long)passes * ncols, Lp, passes,
				                          st->d_msapt + n0, (long)Np, ncols, st->d_gt, (long)P,
				                          !first, st->cfg.gemm_variant, s));
		}
	}
	return 0;
fail:
	return -1;
}

static int pair_gradient_from_forward(ccm_hip_state *st)
{
	/* the single-chunk forward left D / Dt in place */
	hipStream_t s = st->stream;
	if (st->cfg.precision == CCM_PRECISION_FP32)
		KCHK(ccm_launch_gather_grad_seg_f32(st->d_zd, st->ldz, st->d_segmeta, st->L, st->N,
		                                    st->d_graw, st->ldg, 0, s));
	else
		KCHK(ccm_launch_gemm_grad(st->d_dt, st->lddt, st->Lp, st->cfg.bf16_passes, st->d_msapt,
		                          (long)st->Np, st->ccols, st->d_gt, (long)st->P, 0,
		                          st->cfg.gemm_variant, s));
	return 0;
fail:
	return -1;
}

/* g from the raw pair gradient and the node-gradient partials, and the
 * objective (softmax partials + regularisation) into *f_out (device
 * address).  reg_done: the coupling conversion of this evaluation
 * (ccm_launch_convert_ws_reg) wrote the regularisation partials and the
 * a >= 21 padding of g, and the assembly forms f in its own launch. */
static int assemble_gradient(ccm_hip_state *st, const float *d_x, float *d_g, double *f_out, int reg_done)
{
	hipStream_t s = st->stream;
	const int L = st->L;
	const int fp32 = st->cfg.precision == CCM_PRECISION_FP32;
	if (reg_done) {
		KCHK(ccm_launch_assemble_fs(st->d_gt, (long)st->P, st->Lp, st->d_hsym, d_x, d_g, L, st->lam1,
		                            st->lam2, st->d_g1part, st->n_g1blk, st->d_fpart, st->n_fpart,
		                            st->d_regpair,
		                            ccm_convert_ws_reg_parts(L, st->Lp, st->cfg.bf16_fwd_passes),
		                            st->d_regsingle, 1, f_out, s));
		return 0;
	}
	KCHK(ccm_launch_assemble(fp32 ? st->d_graw : st->d_gt, fp32 ? st->ldg : (long)st->P,
	                         fp32 ? L : st->Lp, fp32 ? NULL : st->d_hsym, d_x, d_g, L, st->lam1,
	                         st->lam2, st->d_g1part, st->n_g1blk, st->d_fpart, st->n_fpart,
	                         st->d_regpair, st->d_regsingle, f_out, s));
	return 0;
fail:
	return -1;
}

static int read_f(ccm_hip_state *st, int nreg_pair, int nreg_single, double *f)
{
	hipStream_t s = st->stream;
	KCHK(ccm_launch_sum3(st->d_fpart, st->n_fpart, st->d_regpair, nreg_pair, st->d_regsingle,
	                     nreg_single, st->d_f, s));
	HIPCHK(hipMemcpyAsync(st->h_f, st->d_f, sizeof(double), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	*f = st->h_f[0];
	return 0;
fail:
	return -1;
}

int ccm_hip_evaluate(void *instance, const float *d_x, float *d_g, int nvar, double *f)
{
	ccm_hip_state *st = get_state(instance, nvar);
	if (!st)
		return -1;
	hipStream_t s = st->stream;
	int step = 0;
	st->forward_valid = 0;
	/* the sparse form's conversion also writes the regularisation partials
	 * and g's a >= 21 padding */
	const int reg_done = st->cfg.precision != CCM_PRECISION_FP32 && st->sparse;
	if (st->cfg.precision != CCM_PRECISION_FP32) {
		MARK(step++);
		if (reg_done)
			KCHK(ccm_launch_convert_ws_reg(d_x, d_g, st->L, st->Lp, st->d_wt, st->ldwt,
			                               st->cfg.bf16_fwd_passes, st->lam1, st->lam2, st->d_regpair,
			                               st->d_regsingle, 0, s));
		else
			KCHK(convert_couplings(st, d_x, s));
	}
	if (run_chunks(st, d_x, RUN_FULL, &step, reg_done ? d_g : NULL))
		goto fail;
	MARK(step++);
	/* the objective's last kernel writes f into mapped host memory: no
	 * device-to-host copy before the synchronisation */
	if (assemble_gradient(st, d_x, d_g, st->dh_f, reg_done))
		goto fail;
	MARK(step);
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

int ccm_hip_objective(void *instance, const float *d_x, int nvar, double *f)
{
	ccm_hip_state *st = get_state(instance, nvar);
	if (!st)
		return -1;
	hipStream_t s = st->stream;
	int step = 0, timing = st->timing;
	st->timing = 0;
	st->forward_valid = 0;
	if (st->cfg.precision != CCM_PRECISION_FP32)
		KCHK(convert_couplings(st, d_x, s));
	if (run_chunks(st, d_x, RUN_OBJECTIVE, &step, NULL))
		goto fail;
	KCHK(ccm_launch_reg_only(d_x, st->L, st->lam1, st->lam2, st->d_regpair, s));
	if (read_f(st, ccm_reg_only_blocks(), 0, f))
		goto fail;
	st->timing = timing;
	return 0;
fail:
	st->timing = timing;
	return -1;
}

int ccm_hip_forward(void *instance, const float *d_x, int nvar, double *f)
{
	ccm_hip_state *st = get_state(instance, nvar);
	if (!st)
		return -1;
	if (st->chunk < st->N) {
		fprintf(stderr, "ccmpred_amd: ccm_hip_forward needs an unchunked instance\n");
		return -1;
	}
	hipStream_t s = st->stream;
	int step = 0, timing = st->timing;
	st->timing = 0;
	st->forward_valid = 0;
	if (st->cfg.precision != CCM_PRECISION_FP32)
		KCHK(convert_couplings(st, d_x, s));
	if (run_chunks(st, d_x, RUN_FORWARD, &step, NULL))
		goto fail;
	KCHK(ccm_launch_reg_only(d_x, st->L, st->lam1, st->lam2, st->d_regpair, s));
	if (read_f(st, ccm_reg_only_blocks(), 0, f))
		goto fail;
	st->forward_valid = 1;
	st->timing = timing;
	return 0;
fail:
	st->timing = timing;
	return -1;
}

int ccm_hip_backward(void *instance, const float *d_x, float *d_g, int nvar)
{
	ccm_hip_state *st = get_state(instance, nvar);
	if (!st)
		return -1;
	if (!st->forward_valid) {
		fprintf(stderr, "ccmpred_amd: ccm_hip_backward without a preceding ccm_hip_forward\n");
		return -1;
	}
	if (pair_gradient_from_forward(st) || assemble_gradient(st, d_x, d_g, st->d_f, 0))
		return -1;
	if (hipStreamSynchronize(st->stream) != hipSuccess)
		return -1;
	return 0;
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
	} else if (!on && st->timing) {
		HIPCHK(hipStreamSynchronize(st->stream));
		for (int i = 0; i <= CCM_TIMING_STEPS; i++)
			(void)hipEventDestroy(st->ev[i]);
		st->timing = 0;
	}
	for (int i = 0; i < CCM_TIMING_STEPS; i++)
		st->step_ms[i] = 0.0;
	st->timed_calls = 0;
	return 0;
fail:
	return -1;
}

int ccm_hip_step_names(void *instance, const char **names, int n)
{
	static const char *fp32[] = {"gather_fwd", "softmax_rows", "gather_grad", "assemble"};
	static const char *bf16[] = {"convert_wt", "gemm_fwd", "softmax_cols", "gemm_grad", "assemble"};
	static const char *fused[] = {"convert_wt", "gemm_fwd_softmax", "gemm_grad", "assemble"};
	static const char *sparse[] = {"convert_ws", "sparse_fwd_softmax", "gemm_grad", "assemble"};
	userdata *ud = (userdata *)instance;
	ccm_hip_state *st = (ccm_hip_state *)ud->extra;
	if (!st)
		return -1;
	const char **v = st->cfg.precision == CCM_PRECISION_FP32 ? fp32
	                 : st->sparse                             ? sparse
	                 : st->fused                              ? fused
	                                                          : bf16;
	const int cnt = st->cfg.precision == CCM_PRECISION_FP32 ? 4 : st->fused ? 4 : 5;
	for (int i = 0; i < n && i < cnt; i++)
		names[i] = v[i];
	return cnt;
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
