/*
 * Batched form: many independent alignments optimised together on one GPU.
 *
 * No counterpart in the reference, which optimises one alignment per
 * process (ccmpred.c:537-648).  Each alignment gets the BF16 evaluator of
 * init_hip_ex (whole alignment as one chunk) and four device vectors; a CG
 * round is a fixed sequence of grouped launches, each covering every
 * alignment that still has work:
 *
 *   cg_vec       x += c_pre s; s per the control block; x += c_trial s
 *   convert_wt   Wt = bf16(W)^T
 *   gemm_fwd     one launch per forward tile variant in use
 *   softmax_cols
 *   gemm_grad    one launch per gradient tile variant in use
 *   assemble     pair epilogue and singles; sum3: f into the control block
 *   dots         <g,s>, <g,g>, <g,gprev>
 *   cg_ctrl      line-search / direction decisions of ccm_conjugrad_gpu_ex
 *
 * Nothing in a round needs the host, so rounds are captured into a
 * hipGraph once and replayed until every alignment is done; the host reads
 * one counter per replay.  The grouped kernels run the same device code as
 * the single-alignment launchers, and the control kernel makes the host
 * loop's decisions in the same FP64 arithmetic, so an alignment solved in a
 * batch follows the trajectory ccm_conjugrad_gpu_ex(evaluate_hip) takes on
 * the same BF16 instance.
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

#define NVAR_MAX 16 /* GEMM launches (one per tile variant) per contraction */

typedef struct {
	ccm_group g;
	int variant;     /* GEMM groups only */
	size_t qmap_pos; /* GEMM groups: table position of the tile map */
} bgroup;

struct ccm_batch {
	int n;
	int fused; /* forward contractions with the softmax fused (every alignment fits) */
	int sparse; /* and on structured-sparse MFMA (every alignment fits) */
	int borrowed; /* one alignment whose instance and x belong to the caller */
	ccm_hip_config cfg;
	hipStream_t s;
	userdata *uds;
	ccm_bprob *h_probs;
	ccm_bprob *d_probs;
	ccm_cg_ctl *d_ctl;
	double *d_log;
	int log_cap;
	int *var_fwd, *var_grad;
	/* group tables (device) and the groups that index them */
	uint32_t *d_tab;
	size_t tab_len;
	bgroup g_conv, g_soft, g_asm, g_all, g_dots, g_vec;
	bgroup g_fwd[NVAR_MAX], g_grad[NVAR_MAX];
	int n_fwd, n_grad;
	int32_t *d_nactive, *h_nactive;
	float *d_cmap; /* contact map + column means, allocated on first use */
	size_t cmap_cap;
	size_t bytes;
};

#define HIPCHK(call)                                                                       \
	do {                                                                               \
		hipError_t e_ = (call);                                                    \
		if (e_ != hipSuccess) {                                                    \
			fprintf(stderr, "ccmpred_amd batch: HIP error %d (%s) at %s:%d\n",      \
			        (int)e_, hipGetErrorString(e_), __FILE__, __LINE__);       \
			goto fail;                                                         \
		}                                                                          \
	} while (0)

#define KCHK(call)                                                                         \
	do {                                                                               \
		int r_ = (call);                                                           \
		if (r_ != 0) {                                                             \
			fprintf(stderr, "ccmpred_amd batch: launch failed (%d) at %s:%d\n",      \
			        r_, __FILE__, __LINE__);                                   \
			goto fail;                                                         \
		}                                                                          \
	} while (0)

static hipError_t balloc(ccm_batch *b, void **p, size_t bytes)
{
	hipError_t e = hipMalloc(p, bytes);
	if (e == hipSuccess)
		b->bytes += bytes;
	return e;
}

/* ---- group tables ------------------------------------------------------ */

enum { K_CONV, K_SOFT, K_ASM, K_ALL, K_DOTS, K_VEC };

static uint32_t blocks_of(const ccm_bprob *p, int kind)
{
	switch (kind) {
	case K_CONV:
		return (uint32_t)(p->sparse ? ccm_convert_ws_blocks(p->L, p->Lp, p->passes_fwd)
		                            : ccm_convert_wt_blocks(p->L, p->Lp));
	case K_SOFT: return (uint32_t)ccm_softmax_cols_blocks(p->L, p->N);
	case K_ASM:
		return (uint32_t)(p->sparse ? ccm_assemble_fs_blocks(p->L) : ccm_assemble_blocks(p->L));
	case K_ALL: return 1;
	case K_DOTS: return (uint32_t)p->n_dots;
	default: return (uint32_t)ccm_cg_vec_blocks(p->nvp);
	}
}

/* host image of the tables: per group np+1 offsets then np indices */
typedef struct {
	uint32_t *h;
	size_t len, cap;
} tabbuf;

static int tab_push(tabbuf *t, uint32_t v)
{
	if (t->len == t->cap) {
		size_t nc = t->cap ? 2 * t->cap : 1024;
		uint32_t *nh = (uint32_t *)realloc(t->h, nc * sizeof(uint32_t));
		if (!nh)
			return -1;
		t->h = nh;
		t->cap = nc;
	}
	t->h[t->len++] = v;
	return 0;
}

/* A row panel of a grouped contraction: the tiles of one alignment that
 * share a band of output rows (and so the same rows of the dense operand). */
typedef struct {
	double work;   /* tiles x K steps */
	int prob, row; /* position in the group, tile row */
} panel;

static int panel_cmp(const void *a, const void *b)
{
	const panel *x = (const panel *)a, *y = (const panel *)b;
	if (x->work != y->work)
		return x->work > y->work ? -1 : 1;
	if (x->prob != y->prob)
		return x->prob < y->prob ? -1 : 1;
	return x->row < y->row ? -1 : (x->row > y->row);
}

/* row panels of a grouped contraction's tiles for one alignment */
static int panel_rows(const ccm_bprob *p, int mode, int variant, int tm)
{
	if (mode == 3)
		return (int)ccm_sparse_fwd_tile_rows(p->L);
	if (mode == 2)
		return ccm_fwd_softmax_tile_rows(variant, p->L);
	return (p->P + tm - 1) / tm;
}

/* Tile map of a grouped contraction.  Workgroups are dispatched to the 8
 * XCDs round-robin (workgroup b to XCD b % 8), and the tiles of a group
 * differ in cost (their K length is the alignment's).  Whole row panels go
 * to XCDs, longest first onto the least-loaded XCD, so that the XCDs get
 * equal work, a panel's dense rows stay in one L2, and each XCD runs its
 * heaviest tiles first; qmap[8 i + x] is XCD x's i-th tile (CCM_QMAP_IDLE
 * past its list).  Returns the map length (the grid), 0 on failure. */
static uint32_t push_tile_map(tabbuf *t, const ccm_bprob *probs, const int *list, int np, int mode,
                              int variant, const uint32_t *offs)
{
	int tm = 256, tn = 256;
	if (mode != 3 && ccm_gemm_tile_shape(variant, &tm, &tn))
		return 0;
	/* mode 2: the fused forward, whose row tiles cover tm / 21 alignment
	 * columns each; mode 3: the sparse fused forward (panels of 12
	 * columns, 256 sequences per tile) */
	size_t npan = 0;
	for (int i = 0; i < np; i++)
		npan += (size_t)panel_rows(&probs[list[i]], mode, variant, tm);
	const size_t ntiles = offs[np];
	panel *pan = (panel *)malloc(sizeof(panel) * (npan ? npan : 1));
	uint32_t *xl = (uint32_t *)malloc(sizeof(uint32_t) * 8 * (ntiles ? ntiles : 1));
	uint32_t len = 0;
	size_t cnt[8] = {0};
	double load[8] = {0};
	if (!pan || !xl)
		goto out;
	npan = 0;
	for (int i = 0; i < np; i++) {
		const ccm_bprob *p = &probs[list[i]];
		const int ncol = mode != 1 ? p->ncols : p->P;
		const int cols = (ncol + tn - 1) / tn;
		const int rows = panel_rows(p, mode, variant, tm);
		const double steps = mode == 3   ? (double)ccm_sparse_steps(p->Lp, p->passes_fwd)
		                     : mode != 1 ? (double)p->passes_fwd * p->P / 64.0
		                                 : (double)p->passes * p->ncols / 64.0;
		for (int r = 0; r < rows; r++) {
			pan[npan].work = steps * cols;
			pan[npan].prob = i;
			pan[npan].row = r;
			npan++;
		}
	}
	qsort(pan, npan, sizeof(panel), panel_cmp);
	for (size_t k = 0; k < npan; k++) {
		int x = 0;
		for (int y = 1; y < 8; y++)
			if (load[y] < load[x])
				x = y;
		const ccm_bprob *p = &probs[list[pan[k].prob]];
		const int ncol = mode != 1 ? p->ncols : p->P;
		const int cols = (ncol + tn - 1) / tn;
		for (int c = 0; c < cols; c++)
			xl[(size_t)x * ntiles + cnt[x]++] = offs[pan[k].prob] + (uint32_t)(pan[k].row * cols + c);
		load[x] += pan[k].work;
	}
	size_t mx = 0;
	for (int x = 0; x < 8; x++)
		mx = cnt[x] > mx ? cnt[x] : mx;
	for (size_t i = 0; i < mx; i++)
		for (int x = 0; x < 8; x++)
			if (tab_push(t, i < cnt[x] ? xl[(size_t)x * ntiles + i] : CCM_QMAP_IDLE))
				goto out;
	len = (uint32_t)(8 * mx);
out:
	free(pan);
	free(xl);
	return len;
}

/* appends a group over the listed problems; offsets and indices are stored
 * as table positions and turned into device pointers once uploaded */
static int add_group(tabbuf *t, bgroup *bg, const ccm_bprob *probs, const int *list, int np,
                     int kind, int mode, int variant)
{
	size_t o = t->len;
	uint32_t acc = 0;
	if (tab_push(t, 0))
		return -1;
	for (int i = 0; i < np; i++) {
		const ccm_bprob *p = &probs[list[i]];
		uint32_t nb;
		if (kind < 0)
			nb = mode == 3   ? ccm_sparse_fwd_tile_count(p->L, p->ncols)
			     : mode == 2 ? ccm_fwd_softmax_tile_count(variant, p->L, p->ncols)
			     : mode == 0 ? ccm_gemm_tile_count(variant, p->P, p->ncols)
			                 : ccm_gemm_tile_count(variant, p->P, p->P);
		else
			nb = blocks_of(p, kind);
		acc += nb;
		if (tab_push(t, acc))
			return -1;
	}
	size_t ix = t->len;
	for (int i = 0; i < np; i++)
		if (tab_push(t, (uint32_t)list[i]))
			return -1;
	bg->g.offs = (const uint32_t *)(uintptr_t)o;
	bg->g.pidx = (const int32_t *)(uintptr_t)ix;
	bg->g.np = np;
	bg->g.total = acc;
	bg->g.qmap = NULL;
	bg->variant = variant;
	bg->qmap_pos = 0;
	if (kind < 0 && acc > 0) {
		/* a copy of the offsets: pushing the map may move the table */
		uint32_t *ofs = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(np + 1));
		if (!ofs)
			return -1;
		memcpy(ofs, t->h + o, sizeof(uint32_t) * (size_t)(np + 1));
		const size_t qp = t->len;
		const uint32_t grid = push_tile_map(t, probs, list, np, mode, variant, ofs);
		free(ofs);
		if (!grid)
			return -1;
		bg->qmap_pos = qp + 1; /* 0: no map */
		bg->g.total = grid;
	}
	return 0;
}

static void fix_group(bgroup *bg, const uint32_t *base)
{
	bg->g.offs = base + (uintptr_t)bg->g.offs;
	bg->g.pidx = (const int32_t *)(base + (uintptr_t)bg->g.pidx);
	bg->g.qmap = bg->qmap_pos ? base + (bg->qmap_pos - 1) : NULL;
}

static int build_tables(ccm_batch *b)
{
	tabbuf t = {0};
	int *all = (int *)malloc(sizeof(int) * (size_t)b->n);
	int *list = (int *)malloc(sizeof(int) * (size_t)b->n);
	int rc = -1;
	if (!all || !list)
		goto out;
	for (int i = 0; i < b->n; i++)
		all[i] = i;
	if (add_group(&t, &b->g_conv, b->h_probs, all, b->n, K_CONV, 0, 0) ||
	    add_group(&t, &b->g_soft, b->h_probs, all, b->n, K_SOFT, 0, 0) ||
	    add_group(&t, &b->g_asm, b->h_probs, all, b->n, K_ASM, 0, 0) ||
	    add_group(&t, &b->g_all, b->h_probs, all, b->n, K_ALL, 0, 0) ||
	    add_group(&t, &b->g_dots, b->h_probs, all, b->n, K_DOTS, 0, 0) ||
	    add_group(&t, &b->g_vec, b->h_probs, all, b->n, K_VEC, 0, 0))
		goto out;
	b->n_fwd = b->n_grad = 0;
	/* the sparse forward: one launch over every alignment */
	if (b->sparse && add_group(&t, &b->g_fwd[b->n_fwd++], b->h_probs, all, b->n, -1, 3, -1))
		goto out;
	/* one GEMM launch per tile variant in use */
	for (int mode = b->sparse ? 1 : 0; mode < 2; mode++) {
		const int *var = mode == 0 ? b->var_fwd : b->var_grad;
		for (int v = 0; v < ccm_gemm_variants(); v++) {
			int np = 0;
			for (int i = 0; i < b->n; i++)
				if (var[i] == v)
					list[np++] = i;
			if (!np)
				continue;
			if ((mode == 0 ? b->n_fwd : b->n_grad) >= NVAR_MAX)
				goto out;
			bgroup *bg = mode == 0 ? &b->g_fwd[b->n_fwd++] : &b->g_grad[b->n_grad++];
			if (add_group(&t, bg, b->h_probs, list, np, -1, mode == 0 && b->fused ? 2 : mode, v))
				goto out;
		}
	}
	if (b->d_tab && b->tab_len < t.len) {
		(void)hipFree(b->d_tab);
		b->d_tab = NULL;
		b->bytes -= sizeof(uint32_t) * b->tab_len;
	}
	if (!b->d_tab) {
		if (balloc(b, (void **)&b->d_tab, sizeof(uint32_t) * t.len) != hipSuccess)
			goto out;
		b->tab_len = t.len;
	}
	if (hipMemcpy(b->d_tab, t.h, sizeof(uint32_t) * t.len, hipMemcpyHostToDevice) != hipSuccess)
		goto out;
	bgroup *fixed[] = {&b->g_conv, &b->g_soft, &b->g_asm, &b->g_all, &b->g_dots, &b->g_vec};
	for (size_t i = 0; i < sizeof(fixed) / sizeof(fixed[0]); i++)
		fix_group(fixed[i], b->d_tab);
	for (int i = 0; i < b->n_fwd; i++)
		fix_group(&b->g_fwd[i], b->d_tab);
	for (int i = 0; i < b->n_grad; i++)
		fix_group(&b->g_grad[i], b->d_tab);
	rc = 0;
out:
	free(t.h);
	free(all);
	free(list);
	return rc;
}

/* ---- lifecycle --------------------------------------------------------- */

static void bprob_from_state(ccm_bprob *p, const ccm_hip_state *st)
{
	memset(p, 0, sizeof(*p));
	p->L = st->L;
	p->Lp = st->Lp;
	p->P = st->P;
	p->N = st->N;
	p->ncols = st->ccols;
	p->passes_fwd = st->cfg.bf16_fwd_passes;
	p->passes = st->cfg.bf16_passes;
	p->n_fpart = st->n_fpart;
	p->n_g1blk = st->n_g1blk;
	/* the sparse form's conversion writes the regularisation partials */
	p->n_regpair = st->sparse ? ccm_convert_ws_reg_parts(st->L, st->Lp, st->cfg.bf16_fwd_passes)
	                          : ccm_pair_epilogue_blocks(st->L);
	p->n_regsingle = st->sparse ? 1 : st->n_regsingle;
	p->n_dots = ccm_dots_blocks(st->nvp);
	p->nsp = st->nsp;
	p->nvp = st->nvp;
	p->ldwt = st->ldwt;
	p->lddt = st->lddt;
	p->ldmt = st->Np;
	p->lam1 = st->lam1;
	p->lam2 = st->lam2;
	p->msap = st->d_msap;
	p->msapt = st->d_msapt;
	p->w = st->d_w;
	p->wt = st->d_wt;
	p->dt = st->d_dt;
	p->zt = st->d_zt;
	p->gt = st->d_gt;
	p->hsym = st->d_hsym;
	p->fpart = st->d_fpart;
	p->g1part = st->d_g1part;
	p->regpair = st->d_regpair;
	p->regsingle = st->d_regsingle;
	p->sparse = st->sparse;
	p->sidx = st->d_sidx;
	p->ldsidx = st->ldsidx;
}

int ccm_batch_size(const ccm_batch *b) { return b ? b->n : 0; }

void *ccm_batch_stream(ccm_batch *b) { return b ? (void *)b->s : NULL; }

size_t ccm_batch_device_bytes(const ccm_batch *b)
{
	if (!b)
		return 0;
	size_t t = b->bytes;
	for (int i = 0; i < b->n; i++)
		if (b->uds[i].extra)
			t += ccm_hip_device_bytes(&b->uds[i]);
	return t;
}

float *ccm_batch_x(ccm_batch *b, int p) { return b && p >= 0 && p < b->n ? b->h_probs[p].x : NULL; }

float *ccm_batch_g(ccm_batch *b, int p) { return b && p >= 0 && p < b->n ? b->h_probs[p].g : NULL; }

int ccm_batch_write_x(ccm_batch *b, int p, const float *host_x)
{
	if (!b || p < 0 || p >= b->n || !host_x)
		return -1;
	const size_t vb = sizeof(float) * (size_t)b->h_probs[p].nvp;
	return hipMemcpy(b->h_probs[p].x, host_x, vb, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}

int ccm_batch_read(ccm_batch *b, int p, float *host_x, float *host_g)
{
	if (!b || p < 0 || p >= b->n)
		return -1;
	const size_t vb = sizeof(float) * (size_t)b->h_probs[p].nvp;
	if (hipStreamSynchronize(b->s) != hipSuccess)
		return -1;
	if (host_x && hipMemcpy(host_x, b->h_probs[p].x, vb, hipMemcpyDeviceToHost) != hipSuccess)
		return -1;
	if (host_g && hipMemcpy(host_g, b->h_probs[p].g, vb, hipMemcpyDeviceToHost) != hipSuccess)
		return -1;
	return 0;
}

int ccm_batch_contact_map(ccm_batch *b, int p, int apc, int normalize, float *host_out)
{
	if (!b || p < 0 || p >= b->n || !host_out)
		return -1;
	const int L = b->uds[p].ncol;
	const size_t need = (size_t)L * L + (size_t)ccm_contact_map_work_len(L);
	if (need > b->cmap_cap) {
		if (hipStreamSynchronize(b->s) != hipSuccess)
			return -1;
		if (b->d_cmap)
			(void)hipFree(b->d_cmap);
		b->d_cmap = NULL;
		b->cmap_cap = 0;
		if (hipMalloc((void **)&b->d_cmap, need * sizeof(float)) != hipSuccess)
			return -1;
		b->cmap_cap = need;
	}
	float *d_out = b->d_cmap, *d_work = b->d_cmap + (size_t)L * L;
	if (ccm_launch_contact_map(b->h_probs[p].x, L, apc, normalize, d_out, d_work, b->s) != 0)
		return -1;
	if (hipMemcpyAsync(host_out, d_out, sizeof(float) * (size_t)L * L, hipMemcpyDeviceToHost,
	                   b->s) != hipSuccess ||
	    hipStreamSynchronize(b->s) != hipSuccess)
		return -1;
	return 0;
}

void ccm_batch_destroy(ccm_batch *b)
{
	if (!b)
		return;
	if (b->s)
		(void)hipStreamSynchronize(b->s);
	for (int i = 0; i < b->n; i++) {
		ccm_bprob *p = &b->h_probs[i];
		float *v[] = {b->borrowed ? NULL : p->x, p->g, p->gprev, p->s};
		for (int k = 0; k < 4; k++)
			if (v[k])
				(void)hipFree(v[k]);
		if (p->dpart)
			(void)hipFree(p->dpart);
		if (!b->borrowed && b->uds[i].extra)
			destroy_hip(&b->uds[i]);
	}
	if (b->d_probs)
		(void)hipFree(b->d_probs);
	if (b->d_ctl)
		(void)hipFree(b->d_ctl);
	if (b->d_log)
		(void)hipFree(b->d_log);
	if (b->d_tab)
		(void)hipFree(b->d_tab);
	if (b->d_nactive)
		(void)hipFree(b->d_nactive);
	if (b->d_cmap)
		(void)hipFree(b->d_cmap);
	if (b->h_nactive)
		(void)hipHostFree(b->h_nactive);
	if (b->s)
		(void)hipStreamDestroy(b->s);
	free(b->h_probs);
	free(b->var_fwd);
	free(b->var_grad);
	free(b);
}

/* The CG of ccm_conjugrad_gpu_ex on the device (CCM_LINE_SEARCH_DEVICE):
 * returns 0 with the results, 1 when the instance does not qualify or the
 * set-up failed (nothing has run), -1 when the solve failed. */
/* A batch of one alignment over an existing evaluate_hip instance (BF16,
 * unchunked) and the caller's x (16-byte aligned, nvar_padded floats): the
 * batched form's device-side CG then drives that instance's kernels.  The
 * batch owns only its work vectors and tables; destroying it leaves the
 * instance and x alone. */
static int batch_borrow(userdata *ud, float *d_x, ccm_batch **out)
{
	*out = NULL;
	const ccm_hip_state *st = (const ccm_hip_state *)ud->extra;
	if (!st || st->cfg.precision != CCM_PRECISION_BF16 || st->chunk < st->N ||
	    ((uintptr_t)d_x & 15u) != 0)
		return 1;
	ccm_batch *b = (ccm_batch *)calloc(1, sizeof(*b));
	if (!b)
		return 1;
	b->n = 1;
	b->borrowed = 1;
	b->cfg = st->cfg;
	b->uds = ud;
	b->fused = st->fused;
	b->sparse = st->sparse;
	b->h_probs = (ccm_bprob *)calloc(1, sizeof(ccm_bprob));
	b->var_fwd = (int *)malloc(sizeof(int));
	b->var_grad = (int *)malloc(sizeof(int));
	if (!b->h_probs || !b->var_fwd || !b->var_grad)
		goto fail;
	HIPCHK(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
	HIPCHK(balloc(b, (void **)&b->d_ctl, sizeof(ccm_cg_ctl)));
	HIPCHK(hipMemset(b->d_ctl, 0, sizeof(ccm_cg_ctl)));
	ccm_bprob *p = &b->h_probs[0];
	bprob_from_state(p, st);
	p->x = d_x;
	const size_t vb = sizeof(float) * (size_t)p->nvp;
	HIPCHK(balloc(b, (void **)&p->g, vb));
	HIPCHK(balloc(b, (void **)&p->gprev, vb));
	HIPCHK(balloc(b, (void **)&p->s, vb));
	HIPCHK(hipMemset(p->g, 0, vb));
	HIPCHK(hipMemset(p->gprev, 0, vb));
	HIPCHK(hipMemset(p->s, 0, vb));
	HIPCHK(balloc(b, (void **)&p->dpart, sizeof(double) * 3 * (size_t)p->n_dots));
	p->ctl = b->d_ctl;
	b->var_fwd[0] = st->cfg.gemm_variant >= 0 ? st->cfg.gemm_variant
	                                          : ccm_gemm_default_variant_for(0, p->Lp, p->ncols);
	b->var_grad[0] = st->cfg.gemm_variant >= 0 ? st->cfg.gemm_variant
	                                           : ccm_gemm_default_variant_for(1, p->Lp, p->ncols);
	HIPCHK(balloc(b, (void **)&b->d_probs, sizeof(ccm_bprob)));
	HIPCHK(hipMemcpy(b->d_probs, b->h_probs, sizeof(ccm_bprob), hipMemcpyHostToDevice));
	HIPCHK(balloc(b, (void **)&b->d_nactive, sizeof(int32_t)));
	HIPCHK(hipHostMalloc((void **)&b->h_nactive, sizeof(int32_t), hipHostMallocDefault));
	if (build_tables(b))
		goto fail;
	*out = b;
	return 0;
fail:
	ccm_batch_destroy(b);
	return 1;
}

int ccm_conjugrad_instance_device(void *instance, float *d_x, const ccm_conjugrad_parameter *param,
                                  double *fx, int *iterations, int *evaluations, double *fhist,
                                  int *status)
{
	ccm_batch *b = NULL;
	double *log = NULL;
	int it = 0, rc = -1;
	if (param->max_iterations <= 0 || param->max_linesearch <= 0)
		return 1;
	log = (double *)malloc(sizeof(double) * 4 * (size_t)param->max_iterations);
	if (!log || batch_borrow((userdata *)instance, d_x, &b)) {
		free(log);
		return 1; /* nothing has run: x is unchanged */
	}
	if (ccm_batch_solve(b, param, 8, NULL) || ccm_batch_result(b, 0, status, fx, &it, evaluations, log))
		goto out;
	if (iterations)
		*iterations = it;
	if (fhist)
		for (int i = 0; i < it && i < param->max_iterations; i++)
			fhist[i] = log[4L * i];
	rc = 0;
out:
	free(log);
	ccm_batch_destroy(b);
	return rc;
}

int ccm_batch_create(userdata *uds, int nprob, const ccm_hip_config *cfg_in, ccm_batch **out)
{
	*out = NULL;
	if (!uds || nprob <= 0)
		return 1;
	ccm_hip_config cfg;
	if (cfg_in)
		cfg = *cfg_in;
	else
		ccm_hip_default_config(&cfg);
	if (cfg.precision != CCM_PRECISION_BF16) {
		fprintf(stderr, "ccmpred_amd batch: the batched form runs the BF16 evaluator\n");
		return 1;
	}
	cfg.chunk_rows = 0;
	cfg.gemm_variant = -1;
	ccm_batch *b = (ccm_batch *)calloc(1, sizeof(*b));
	if (!b)
		return 1;
	b->n = nprob;
	b->cfg = cfg;
	b->uds = uds;
	{
		const int fp = cfg.bf16_fwd_passes ? cfg.bf16_fwd_passes : cfg.bf16_passes;
		b->fused = 1;
		b->sparse = cfg.sparse_fwd != 0;
		for (int i = 0; i < nprob; i++) {
			const int Lp = (uds[i].ncol + 63) / 64 * 64, Np = (uds[i].nrow + 127) / 128 * 128;
			b->fused &= ccm_fwd_softmax_fits(Lp, fp, cfg.bf16_passes, Np);
			b->sparse &= ccm_sparse_fwd_fits(Lp, fp, cfg.bf16_passes, Np) &&
			             (long)ccm_sparse_steps(Lp, fp) * 2 * ccm_sparse_index_ld(Np) * 4 <= 0x7FFFFFFFL;
		}
		b->sparse &= b->fused;
	}
	b->h_probs = (ccm_bprob *)calloc((size_t)nprob, sizeof(ccm_bprob));
	b->var_fwd = (int *)malloc(sizeof(int) * (size_t)nprob);
	b->var_grad = (int *)malloc(sizeof(int) * (size_t)nprob);
	if (!b->h_probs || !b->var_fwd || !b->var_grad)
		goto fail;
	for (int i = 0; i < nprob; i++)
		uds[i].extra = NULL;
	HIPCHK(hipSetDevice(cfg.device));
	HIPCHK(hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking));
	HIPCHK(balloc(b, (void **)&b->d_ctl, sizeof(ccm_cg_ctl) * (size_t)nprob));
	HIPCHK(hipMemset(b->d_ctl, 0, sizeof(ccm_cg_ctl) * (size_t)nprob));
	for (int i = 0; i < nprob; i++) {
		if (ccm_init_hip_state(&uds[i], &cfg, b->fused, b->sparse) != 0) {
			fprintf(stderr, "ccmpred_amd batch: alignment %d failed to initialise\n", i);
			goto fail;
		}
		const ccm_hip_state *st = (const ccm_hip_state *)uds[i].extra;
		ccm_bprob *p = &b->h_probs[i];
		bprob_from_state(p, st);
		if (p->sparse != b->sparse) {
			fprintf(stderr, "ccmpred_amd batch: alignment %d chose another forward form\n", i);
			goto fail;
		}
		const size_t vb = sizeof(float) * (size_t)p->nvp;
		HIPCHK(balloc(b, (void **)&p->x, vb));
		HIPCHK(balloc(b, (void **)&p->g, vb));
		HIPCHK(balloc(b, (void **)&p->gprev, vb));
		HIPCHK(balloc(b, (void **)&p->s, vb));
		HIPCHK(hipMemset(p->x, 0, vb));
		HIPCHK(hipMemset(p->g, 0, vb));
		HIPCHK(hipMemset(p->gprev, 0, vb));
		HIPCHK(hipMemset(p->s, 0, vb));
		HIPCHK(balloc(b, (void **)&p->dpart, sizeof(double) * 3 * (size_t)p->n_dots));
		p->ctl = b->d_ctl + i;
		b->var_fwd[i] = ccm_gemm_default_variant_for(0, p->Lp, p->ncols);
		b->var_grad[i] = ccm_gemm_default_variant_for(1, p->Lp, p->ncols);
	}
	HIPCHK(balloc(b, (void **)&b->d_probs, sizeof(ccm_bprob) * (size_t)nprob));
	HIPCHK(hipMemcpy(b->d_probs, b->h_probs, sizeof(ccm_bprob) * (size_t)nprob,
	                 hipMemcpyHostToDevice));
	HIPCHK(balloc(b, (void **)&b->d_nactive, sizeof(int32_t)));
	HIPCHK(hipHostMalloc((void **)&b->h_nactive, sizeof(int32_t), hipHostMallocDefault));
	if (build_tables(b))
		goto fail;
	*out = b;
	return 0;
fail:
	ccm_batch_destroy(b);
	return 1;
}

int ccm_batch_set_variants(ccm_batch *b, const int *fwd, const int *grad)
{
	for (int i = 0; i < b->n; i++) {
		const ccm_bprob *p = &b->h_probs[i];
		int vf = fwd ? fwd[i] : -1, vg = grad ? grad[i] : -1;
		if (vf < 0)
			vf = ccm_gemm_default_variant_for(0, p->Lp, p->ncols);
		if (vg < 0)
			vg = ccm_gemm_default_variant_for(1, p->Lp, p->ncols);
		if (!ccm_gemm_variant_ok(vf, 0, p->Lp, p->ncols) ||
		    !ccm_gemm_variant_ok(vg, 1, p->Lp, p->ncols)) {
			fprintf(stderr, "ccmpred_amd batch: tile variant %d/%d does not fit alignment %d\n",
			        vf, vg, i);
			return -1;
		}
		b->var_fwd[i] = vf;
		b->var_grad[i] = vg;
	}
	return build_tables(b);
}

int ccm_batch_get_variants(const ccm_batch *b, int *fwd, int *grad)
{
	for (int i = 0; i < b->n; i++) {
		if (fwd)
			fwd[i] = b->var_fwd[i];
		if (grad)
			grad[i] = b->var_grad[i];
	}
	return 0;
}

/* ---- rounds ------------------------------------------------------------ */

static int enqueue_gemms(ccm_batch *b, int mode)
{
	const int ng = mode == 0 ? b->n_fwd : b->n_grad;
	const bgroup *gs = mode == 0 ? b->g_fwd : b->g_grad;
	/* one alignment with the default gradient tile: the banded launch of
	 * the single-alignment evaluation where its size takes one (a grouped
	 * launch has one tile shape per alignment) */
	if (mode == 1 && b->n == 1 && ng == 1 &&
	    gs[0].variant == ccm_gemm_default_variant_for(1, b->h_probs[0].Lp, b->h_probs[0].ncols)) {
		const int r = ccm_launch_gemm_grad_banded_desc(b->d_probs, 0, b->h_probs[0].P, b->s);
		if (r != 1)
			return r;
	}
	for (int i = 0; i < ng; i++) {
		int r = mode == 1   ? ccm_launch_gemm_grad_grouped(b->d_probs, gs[i].g, gs[i].variant, b->s)
		        : b->sparse ? ccm_launch_sparse_fwd_softmax_grouped(b->d_probs, gs[i].g, b->s)
		        : b->fused  ? ccm_launch_gemm_fwd_softmax_grouped(b->d_probs, gs[i].g, gs[i].variant, b->s)
		                    : ccm_launch_gemm_fwd_grouped(b->d_probs, gs[i].g, gs[i].variant, b->s);
		if (r)
			return r;
	}
	return 0;
}

static int enqueue_eval(ccm_batch *b)
{
	if (b->sparse)
		KCHK(ccm_launch_convert_ws_grouped(b->d_probs, b->g_conv.g, b->h_probs[0].passes_fwd, b->s));
	else
		KCHK(ccm_launch_convert_wt_grouped(b->d_probs, b->g_conv.g, b->s));
	KCHK(enqueue_gemms(b, 0));
	if (!b->fused)
		KCHK(ccm_launch_softmax_cols_grouped(b->d_probs, b->g_soft.g, b->s));
	KCHK(enqueue_gemms(b, 1));
	KCHK(ccm_launch_assemble_grouped(b->d_probs, b->g_asm.g, b->s));
	/* (the sparse form's assembly forms f itself) */
	if (!b->sparse)
		KCHK(ccm_launch_sum3_grouped(b->d_probs, b->g_all.g, b->s));
	return 0;
fail:
	return -1;
}

static int enqueue_round(ccm_batch *b, const ccm_cg_prm *prm)
{
	KCHK(ccm_launch_cg_vec_grouped(b->d_probs, b->g_vec.g, b->s));
	if (enqueue_eval(b))
		goto fail;
	KCHK(ccm_launch_dots_grouped(b->d_probs, b->g_dots.g, b->g_all.g, b->s));
	KCHK(ccm_launch_cg_ctrl(b->d_probs, b->n, prm, b->d_nactive, b->s));
	return 0;
fail:
	return -1;
}

/* control blocks for a plain evaluation of every alignment */
static int arm_eval(ccm_batch *b)
{
	ccm_cg_ctl *h = (ccm_cg_ctl *)calloc((size_t)b->n, sizeof(ccm_cg_ctl));
	if (!h)
		return -1;
	for (int i = 0; i < b->n; i++)
		h[i].eval_on = 1;
	hipError_t e = hipMemcpy(b->d_ctl, h, sizeof(ccm_cg_ctl) * (size_t)b->n, hipMemcpyHostToDevice);
	free(h);
	return e == hipSuccess ? 0 : -1;
}

int ccm_batch_evaluate(ccm_batch *b, double *f)
{
	ccm_cg_ctl *h = NULL;
	if (arm_eval(b) || enqueue_eval(b))
		return -1;
	h = (ccm_cg_ctl *)malloc(sizeof(ccm_cg_ctl) * (size_t)b->n);
	if (!h)
		return -1;
	HIPCHK(hipMemcpyAsync(h, b->d_ctl, sizeof(ccm_cg_ctl) * (size_t)b->n, hipMemcpyDeviceToHost,
	                      b->s));
	HIPCHK(hipStreamSynchronize(b->s));
	for (int i = 0; i < b->n; i++)
		f[i] = h[i].fx;
	free(h);
	return 0;
fail:
	free(h);
	return -1;
}

/* ---- autotuner --------------------------------------------------------- */

static int time_gemms(ccm_batch *b, int mode, int reps, double *ms)
{
	hipEvent_t e0 = NULL, e1 = NULL;
	double *t = (double *)malloc(sizeof(double) * (size_t)reps);
	int rc = -1;
	if (!t)
		return -1;
	HIPCHK(hipEventCreate(&e0));
	HIPCHK(hipEventCreate(&e1));
	KCHK(enqueue_gemms(b, mode)); /* warm-up */
	for (int r = 0; r < reps; r++) {
		HIPCHK(hipEventRecord(e0, b->s));
		KCHK(enqueue_gemms(b, mode));
		HIPCHK(hipEventRecord(e1, b->s));
		HIPCHK(hipEventSynchronize(e1));
		float v = 0.f;
		HIPCHK(hipEventElapsedTime(&v, e0, e1));
		t[r] = v;
	}
	/* median */
	for (int i = 1; i < reps; i++)
		for (int j = i; j > 0 && t[j - 1] > t[j]; j--) {
			double x = t[j];
			t[j] = t[j - 1];
			t[j - 1] = x;
		}
	*ms = t[reps / 2];
	rc = 0;
fail:
	if (e0)
		(void)hipEventDestroy(e0);
	if (e1)
		(void)hipEventDestroy(e1);
	free(t);
	return rc;
}

/* Per contraction: every variant for all alignments (the default where it
 * does not fit), then one pass of coordinate descent over four size
 * classes (quartiles of the output tile area), keeping a change only when
 * the grouped launches get faster by more than 1 %. */
int ccm_batch_autotune(ccm_batch *b, int reps, double *ms_out)
{
	const int n = b->n, nv = ccm_gemm_variants();
	int *cur = (int *)malloc(sizeof(int) * (size_t)n), *trial = (int *)malloc(sizeof(int) * (size_t)n);
	int *order = (int *)malloc(sizeof(int) * (size_t)n);
	double *area = (double *)malloc(sizeof(double) * (size_t)n);
	int rc = -1;
	if (!cur || !trial || !order || !area)
		goto out;
	if (reps < 1)
		reps = 1;
	if (arm_eval(b) || enqueue_eval(b) || hipStreamSynchronize(b->s) != hipSuccess)
		goto out;
	for (int mode = 0; mode < 2; mode++) {
		int *var = mode == 0 ? b->var_fwd : b->var_grad;
		double best = 0;
		memcpy(cur, var, sizeof(int) * (size_t)n);
		if (time_gemms(b, mode, reps, &best))
			goto out;
		if (mode == 0 && b->sparse) {
			/* the sparse forward has one tile shape */
			if (ms_out)
				ms_out[0] = best;
			continue;
		}
		for (int v = 0; v < nv; v++) {
			int any = 0;
			if (!ccm_gemm_variant_direct(v))
				continue;
			for (int i = 0; i < n; i++) {
				const ccm_bprob *p = &b->h_probs[i];
				const int ok = ccm_gemm_variant_ok(v, mode, p->Lp, p->ncols);
				trial[i] = ok ? v : ccm_gemm_default_variant_for(mode, p->Lp, p->ncols);
				any |= ok && trial[i] != cur[i];
			}
			if (!any)
				continue;
			memcpy(var, trial, sizeof(int) * (size_t)n);
			double t;
			if (build_tables(b) || time_gemms(b, mode, reps, &t))
				goto out;
			if (t < 0.99 * best) {
				best = t;
				memcpy(cur, trial, sizeof(int) * (size_t)n);
			}
		}
		/* size classes */
		for (int i = 0; i < n; i++) {
			const ccm_bprob *p = &b->h_probs[i];
			order[i] = i;
			area[i] = (double)p->P * (mode == 0 ? p->ncols : p->P);
		}
		for (int i = 1; i < n; i++)
			for (int j = i; j > 0 && area[order[j - 1]] > area[order[j]]; j--) {
				int x = order[j];
				order[j] = order[j - 1];
				order[j - 1] = x;
			}
		for (int c = 0; c < 4; c++) {
			const int lo = c * n / 4, hi = (c + 1) * n / 4;
			if (lo == hi)
				continue;
			for (int v = 0; v < nv; v++) {
				if (!ccm_gemm_variant_direct(v))
					continue;
				memcpy(trial, cur, sizeof(int) * (size_t)n);
				int any = 0;
				for (int k = lo; k < hi; k++) {
					const ccm_bprob *p = &b->h_probs[order[k]];
					if (ccm_gemm_variant_ok(v, mode, p->Lp, p->ncols) && cur[order[k]] != v) {
						trial[order[k]] = v;
						any = 1;
					}
				}
				if (!any)
					continue;
				memcpy(var, trial, sizeof(int) * (size_t)n);
				double t;
				if (build_tables(b) || time_gemms(b, mode, reps, &t))
					goto out;
				if (t < 0.99 * best) {
					best = t;
					memcpy(cur, trial, sizeof(int) * (size_t)n);
				}
			}
		}
		memcpy(var, cur, sizeof(int) * (size_t)n);
		if (build_tables(b))
			goto out;
		if (ms_out)
			ms_out[mode] = best;
	}
	rc = 0;
out:
	free(cur);
	free(trial);
	free(order);
	free(area);
	return rc;
}

/* ---- solve ------------------------------------------------------------- */

int ccm_batch_solve(ccm_batch *b, const ccm_conjugrad_parameter *param, int rounds_per_graph,
                    int *rounds_out)
{
	hipGraph_t graph = NULL;
	hipGraphExec_t exec = NULL;
	ccm_cg_ctl *h = NULL;
	int rc = -1, rounds = 0, capturing = 0;
	if (!param || param->max_iterations <= 0 || param->max_linesearch <= 0)
		return -1;
	if (rounds_per_graph < 1)
		rounds_per_graph = 1;
	ccm_cg_prm prm;
	memset(&prm, 0, sizeof(prm));
	prm.max_iterations = param->max_iterations;
	prm.k = param->k;
	prm.max_linesearch = param->max_linesearch;
	prm.epsilon = param->epsilon;
	prm.alpha_mul = param->alpha_mul;
	prm.ftol = param->ftol;
	prm.wolfe = param->wolfe;
	prm.fres = ccm_hip_objective_resolution(&b->uds[0]);

	/* per-iteration logs */
	if (b->log_cap < param->max_iterations) {
		if (b->d_log) {
			(void)hipFree(b->d_log);
			b->d_log = NULL;
		}
		HIPCHK(balloc(b, (void **)&b->d_log,
		              sizeof(double) * 4 * (size_t)param->max_iterations * (size_t)b->n));
		b->log_cap = param->max_iterations;
	}
	for (int i = 0; i < b->n; i++) {
		b->h_probs[i].log = b->d_log + 4L * b->log_cap * i;
		b->h_probs[i].log_cap = b->log_cap;
	}
	HIPCHK(hipMemcpy(b->d_probs, b->h_probs, sizeof(ccm_bprob) * (size_t)b->n,
	                 hipMemcpyHostToDevice));
	/* initial control: evaluate at x, no vector step */
	h = (ccm_cg_ctl *)calloc((size_t)b->n, sizeof(ccm_cg_ctl));
	if (!h)
		goto fail;
	for (int i = 0; i < b->n; i++) {
		h[i].eval_on = 1;
		h[i].phase = CCM_CG_INIT;
		h[i].hi = INFINITY;
		h[i].pad_scanned = 1;
	}
	HIPCHK(hipMemcpy(b->d_ctl, h, sizeof(ccm_cg_ctl) * (size_t)b->n, hipMemcpyHostToDevice));
	/* g too: where x's padding is zero the sparse form's evaluation leaves
	 * g's padding (2 lambda_pair x = 0 there) as it is */
	for (int i = 0; i < b->n; i++) {
		const size_t vb = sizeof(float) * (size_t)b->h_probs[i].nvp;
		HIPCHK(hipMemsetAsync(b->h_probs[i].s, 0, vb, b->s));
		HIPCHK(hipMemsetAsync(b->h_probs[i].gprev, 0, vb, b->s));
		HIPCHK(hipMemsetAsync(b->h_probs[i].g, 0, vb, b->s));
	}
	/* the vector step and dots skip the a >= 21 padding of every x that
	 * starts with zeros there */
	KCHK(ccm_launch_pad_scan(b->d_probs, b->n, b->s));
	HIPCHK(hipStreamSynchronize(b->s));

	HIPCHK(hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
	capturing = 1;
	for (int r = 0; r < rounds_per_graph; r++)
		if (enqueue_round(b, &prm))
			goto fail;
	capturing = 0;
	HIPCHK(hipStreamEndCapture(b->s, &graph));
	HIPCHK(hipGraphInstantiate(&exec, graph, NULL, NULL, 0));

	/* every evaluation is one round; a CG iteration takes at most
	 * max_linesearch + 1 of them, a restarted one twice that */
	const long max_rounds = 2L * (param->max_iterations + 1) * (param->max_linesearch + 2) + 2;
	for (;;) {
		HIPCHK(hipGraphLaunch(exec, b->s));
		HIPCHK(hipMemcpyAsync(b->h_nactive, b->d_nactive, sizeof(int32_t), hipMemcpyDeviceToHost,
		                      b->s));
		HIPCHK(hipStreamSynchronize(b->s));
		rounds += rounds_per_graph;
		if (*b->h_nactive == 0)
			break;
		if (rounds >= max_rounds) {
			fprintf(stderr, "ccmpred_amd batch: %d alignments still active after %d rounds\n",
			        *b->h_nactive, rounds);
			goto fail;
		}
	}
	rc = 0;
fail:
	if (capturing) {
		hipGraph_t g2 = NULL;
		(void)hipStreamEndCapture(b->s, &g2);
		if (g2)
			(void)hipGraphDestroy(g2);
	}
	if (exec)
		(void)hipGraphExecDestroy(exec);
	if (graph)
		(void)hipGraphDestroy(graph);
	free(h);
	if (rounds_out)
		*rounds_out = rounds;
	return rc;
}

int ccm_batch_result(ccm_batch *b, int p, int *status, double *fx, int *iterations,
                     int *evaluations, double *log)
{
	if (!b || p < 0 || p >= b->n)
		return -1;
	ccm_cg_ctl c;
	if (hipMemcpy(&c, b->d_ctl + p, sizeof(c), hipMemcpyDeviceToHost) != hipSuccess)
		return -1;
	if (status)
		*status = c.status;
	if (fx)
		*fx = c.f;
	if (iterations)
		*iterations = c.it;
	if (evaluations)
		*evaluations = c.nevals;
	if (log && b->d_log && c.it > 0) {
		const int nit = c.it < b->log_cap ? c.it : b->log_cap;
		if (hipMemcpy(log, b->d_log + 4L * b->log_cap * p, sizeof(double) * 4 * (size_t)nit,
		              hipMemcpyDeviceToHost) != hipSuccess)
			return -1;
	}
	return 0;
}
