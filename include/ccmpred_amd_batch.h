/*
 * Device-side descriptors of the batched form: many independent alignments
 * evaluated and optimised together, one grouped launch per step.
 *
 * Shared by the C host layer and the HIP kernels (plain C layout: the
 * host compiler and the gfx950 device compiler agree on it).  No
 * counterpart in the reference, which evaluates one alignment per process
 * (evaluate_cuda.c:15-21 keeps its state in globals).
 */
#ifndef CCMPRED_AMD_BATCH_H
#define CCMPRED_AMD_BATCH_H

#include <stdint.h>

/* what the vector step of a round does to the direction s before the trial
 * step x += c_trial s (after x += c_pre s with the old direction) */
enum {
	CCM_VEC_KEEP = 0,      /* s unchanged */
	CCM_VEC_NEGATE = 1,    /* s = -g, gprev = g (steepest descent from the new point) */
	CCM_VEC_DIRECTION = 2, /* s = -g + beta s, gprev = g (conjugate direction) */
	CCM_VEC_RESTART = 3    /* s = -gprev (line search failed: restart from the old point) */
};

/* control phases of the per-alignment conjugate-gradient state machine */
enum { CCM_CG_INIT = 0, CCM_CG_SEARCH = 1, CCM_CG_FORCED = 2 };

/* Per-alignment control block.  The control kernel writes the first group
 * (read by the next round's kernels) and the scalars; the evaluation writes
 * fx and the dot kernel writes dots. */
typedef struct ccm_cg_ctl {
	int32_t vec_on;   /* the vector step runs this round */
	int32_t eval_on;  /* the evaluation and the dots run this round */
	int32_t vec_mode; /* CCM_VEC_* */
	int32_t done;
	double c_pre, c_trial, beta;
	double f, f0, alpha, moved, lo, hi, dg0, gg;
	int32_t phase, it, ls, steepest;
	int32_t status, nevals;
	int32_t x_pad_nz; /* set by ccm_launch_pad_scan: x holds nonzero a >= 21 padding */
	int32_t pad_scanned; /* set before ccm_launch_pad_scan: x_pad_nz == 0 then means zero padding
	                      * (and g's padding, cleared at the solve's start, stays zero) */
	double fx;      /* objective of the last evaluation (FP64 sum) */
	double dots[3]; /* <g,s>, <g,g>, <g,gprev> at the last evaluation */
} ccm_cg_ctl;

/* One alignment of a batch: geometry and device buffers (BF16 form,
 * unchunked, index spaces as in the single-alignment path). */
typedef struct ccm_bprob {
	int32_t L, Lp, P, N;
	int32_t ncols;              /* N rounded up to 128 */
	int32_t passes_fwd, passes; /* BF16 passes of the forward / gradient operand */
	int32_t n_fpart, n_g1blk, n_regpair, n_regsingle, n_dots;
	int64_t nsp, nvp, ldwt, lddt, ldmt;
	float lam1, lam2;
	const uint8_t *msap, *msapt;
	const float *w;
	uint16_t *wt, *dt;
	float *zt, *gt;
	const float *hsym;
	double *fpart, *g1part, *regpair, *regsingle, *dpart;
	float *x, *g, *gprev, *s;
	ccm_cg_ctl *ctl;
	double *log;     /* per accepted iteration: f, step, line-search trials, <g,g> */
	int32_t log_cap; /* iterations the log holds */
	int32_t sparse;  /* wt holds the sparse forward's layout; sidx its positions */
	const uint32_t *sidx;
	int64_t ldsidx;
} ccm_bprob;

/* A grouped launch: problems pidx[0..np) get consecutive block ranges
 * [offs[i], offs[i+1]); total is the grid size (offs[np], or more for
 * launches whose grid is padded).  Contractions may carry qmap: workgroup
 * b then takes logical tile qmap[b] (CCM_QMAP_IDLE: none) instead of the
 * XCD-contiguous order, and total is the length of qmap. */
#define CCM_QMAP_IDLE 0xFFFFFFFFu
typedef struct ccm_group {
	const uint32_t *offs;
	const int32_t *pidx;
	int32_t np;
	uint32_t total;
	const uint32_t *qmap;
} ccm_group;

/* conjugate-gradient parameters as the control kernel reads them */
typedef struct ccm_cg_prm {
	int32_t max_iterations, k, max_linesearch, pad;
	double epsilon, alpha_mul, ftol, wolfe;
	double fres; /* relative resolution of f (ccm_hip_objective_resolution) */
} ccm_cg_prm;

#endif
