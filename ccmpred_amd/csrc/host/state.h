/*
 * Per-instance state of the single-alignment evaluator (ud->extra), shared
 * by the host files of the library (evaluate_hip.c owns it; batch.c reuses
 * its device buffers).
 */
#ifndef CCMPRED_AMD_STATE_H
#define CCMPRED_AMD_STATE_H

#define __HIP_PLATFORM_AMD__
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "ccmpred_amd.h"

#define CCM_TIMING_STEPS 5

typedef struct ccm_hip_state {
	ccm_hip_config cfg;
	hipStream_t stream;
	int L, N, Lp, P, Np;
	int chunk;  /* sequences per chunk: whole softmax blocks (>= N when not chunked) */
	int ccols;  /* padded width of one chunk (multiple of 128) */
	long nsp, nvp;
	float lam1, lam2;
	size_t bytes; /* device bytes held by this instance */
	unsigned char *d_msap;  /* Np x Lp, 255 past the real rows/columns */
	unsigned char *d_msapt; /* Lp x Np, 255 past the real rows/columns */
	unsigned char *d_segmeta; /* per 128-sequence chunk and column: rows grouped by residue */
	float *d_w;
	/* plain FP32 path */
	float *d_zd; /* chunk x ldz */
	long ldz;
	float *d_graw;
	long ldg;
	/* BF16 path */
	uint16_t *d_wt; /* the forward's coupling operand: Wt, or Ws when sparse */
	long ldwt;
	int sparse;        /* forward on structured-sparse MFMA (fused only) */
	uint32_t *d_sidx;  /* its one-hot positions, [step][half][ldsidx sequences] */
	long ldsidx;
	int fused;      /* forward contraction with the softmax as its epilogue (no d_zt) */
	float *d_zt;    /* P x ccols, unfused only */
	uint16_t *d_dt; /* P x passes*ccols */
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
	double *h_f;  /* pinned, mapped and coherent host memory */
	double *dh_f; /* its device address: the evaluation writes f there */
	int forward_valid; /* ccm_hip_forward left probabilities for ccm_hip_backward */
	/* optional per-step timing with HIP events on the instance stream */
	int timing;
	hipEvent_t ev[CCM_TIMING_STEPS + 1];
	double step_ms[CCM_TIMING_STEPS];
	int timed_calls;
} ccm_hip_state;

/* init_hip_ex with the choice of the fused and of the sparse forward (the
 * batched form runs its grouped kernels, which take the dense operand, on
 * the instance's buffers) */
int ccm_init_hip_state(void *instance, const ccm_hip_config *cfg, int allow_fused, int allow_sparse);

/* batch.c: the batched form's device-side CG on one evaluate_hip instance
 * and the caller's x (ccm_conjugrad_gpu_ex with CCM_LINE_SEARCH_DEVICE):
 * 0 done, 1 the instance does not qualify (nothing ran), -1 failed */
int ccm_conjugrad_instance_device(void *instance, float *d_x, const ccm_conjugrad_parameter *param,
                                  double *fx, int *iterations, int *evaluations, double *fhist,
                                  int *status);

#endif
