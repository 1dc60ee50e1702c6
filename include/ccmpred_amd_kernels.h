/*
 * ccmpred_amd kernel launchers: one extern "C" function per device step of
 * the PLL path, replacing the reference's gpu_* wrappers
 * (evaluate_cuda_kernels.h:16-82, evaluate_cuda_kernels.cu:556-830).
 *
 * Unlike the reference wrappers, these take a stream, never synchronise,
 * never allocate and never exit: they return 0 or a negative status, and
 * every buffer is owned by the caller.  hipStream_t is passed as void*.
 */
#ifndef CCMPRED_AMD_KERNELS_H
#define CCMPRED_AMD_KERNELS_H

#include <stdint.h>

#include "ccmpred_amd_batch.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef void *ccm_stream_t;

/* reweighting (gpu_compute_weights_simple, cu:804-825) */
int ccm_reweight_words(int L);
int ccm_launch_identity_counts(const uint8_t *d_msa, int L, int N, int idthres, uint32_t *d_planes,
                               int32_t *d_counts, ccm_stream_t s);
int ccm_launch_weights_from_counts(const int32_t *d_counts, int N, int L, int idthres, float *d_w,
                                   ccm_stream_t s);
int ccm_launch_fill_f32(float *d_p, long n, float v, ccm_stream_t s);

/* plain FP32 contractions (gpu_pc cu:560-590; gpu_compute_edge_gradients cu:643-666) */
int ccm_launch_gather_fwd_f32(const uint8_t *d_msap, int Lp, const float *d_x, int L, int N,
                              float *d_z, long ldz, ccm_stream_t s);
int ccm_launch_gather_grad_f32(const float *d_d, long ldd, const uint8_t *d_msap, int Lp, int L,
                               int N, float *d_g, long ldg, int accumulate, ccm_stream_t s);
int ccm_seg_meta_bytes(int L, int N);
int ccm_launch_gather_grad_seg_f32(const float *d_d, long ldd, const uint8_t *d_meta, int L, int N,
                                   float *d_g, long ldg, int accumulate, ccm_stream_t s);
int ccm_launch_neg_weighted_onehot(const uint8_t *d_msap, int Lp, const float *d_w, int L, int N,
                                   float *d_d, long ldd, ccm_stream_t s);

/* log-sum-exp / objective terms (cu:80-103, cu:592-641) */
int ccm_softmax_rows_blocks(int L, int N);
int ccm_softmax_rows_nblk(int N);
int ccm_softmax_cols_blocks(int L, int N);
int ccm_softmax_cols_nblk(int N);
int ccm_softmax_cols_granule(void);
/* sequence chunks must start at multiples of ccm_softmax_cols_granule() */
int ccm_launch_softmax_rows(float *d_zd, long ldz, const uint8_t *d_msap, int Lp, const float *d_w,
                            int L, int Nc, int n0, int mode, double *d_fpart, double *d_g1part,
                            ccm_stream_t s);
int ccm_launch_softmax_cols(const float *d_zt, long ldzt, const float *d_x1, const uint8_t *d_msapt,
                            long ldmt, const float *d_w, int L, int Lp, int Nc, int ncols, int n0,
                            uint16_t *d_dt, long lddt, long lo_off, int passes, int mode,
                            double *d_fpart, double *d_g1part, ccm_stream_t s);

/* BF16 MFMA contractions.  convert_wt: Wt[a*Lp + j][b*Lp + k] = W(b,k,a,j)
 * in BF16 (and the residual half at +21*Lp for passes = 2), zeros for
 * k >= L; rows with j >= L are not written (zero Wt once).  A 16-byte
 * aligned d_x takes 16-byte loads over 128-column blocks, any other d_x
 * 4-byte loads over 64-column blocks; both write the same Wt. */
int ccm_launch_convert_wt(const float *d_x, int L, int Lp, uint16_t *d_wt, long lda, int passes,
                          ccm_stream_t s);
int ccm_gemm_variants(void);
int ccm_gemm_default_variant(int M, int Ncol);
int ccm_launch_gemm_fwd(const uint16_t *d_wt, long lda, int Lp, int passes, const uint8_t *d_msap,
                        int ncols, float *d_zt, long ldc, int variant, ccm_stream_t s);
/* The forward contraction with the column softmax as its epilogue
 * (cu:30-107 and cu:114-141 in one launch; no counterpart kernel): the
 * logits are never stored.  Writes Dt (mode 1, as ccm_launch_softmax_cols),
 * the objective partials fpart[blk][j] and the partials
 * g1part[blk][a][j] = sum_n w_n P_n(a, j) (mode 1; unlike softmax_cols
 * without the data term -sum_n w_n [x_nj = a], which the caller adds as one
 * more block) per block blk = blk0 + n / 128 of 128 sequences;
 * ccm_fwd_softmax_blocks(N) blocks cover N sequences. */
int ccm_fwd_softmax_blocks(int N);
int ccm_fwd_softmax_default_variant(int L, int ncols);
/* whether Wt (lda = passes_fwd * 21 * Lp) and Dt (21 * Lp rows of
 * passes * ncols) fit the 32-bit offsets of the fused forward */
int ccm_fwd_softmax_fits(int Lp, int passes_fwd, int passes, int ncols);
int ccm_launch_gemm_fwd_softmax(const uint16_t *d_wt, long lda, int Lp, int passes_fwd,
                                const uint8_t *d_msap, int ncols, const float *d_x1,
                                const uint8_t *d_msapt, long ldmt, const float *d_w, int L, int Nc,
                                uint16_t *d_dt, long lddt, long lo_off, int passes, int mode,
                                int blk0, double *d_fpart, double *d_g1part, int variant,
                                ccm_stream_t s);
/* The fused forward on 2:4 structured-sparse MFMA (v_smfmac_f32_32x32x32_bf16),
 * same outputs as ccm_launch_gemm_fwd_softmax (cu:30-141 in one launch; no
 * counterpart kernel).  The one-hot operand is the sparse one: every
 * coupling column takes 7 groups of 4 K slots (a zero slot and three
 * residues), and per lane only the slot positions vary, precomputed from the
 * MSA.  convert_ws writes the couplings in that K layout
 * (ccm_sparse_ws_len(Lp, passes) bf16 per row c = a*Lp + j, hi and lo halves
 * per instruction for passes = 2; rows j >= L are not written: zero them
 * once).  sparse_index writes the positions of the Np padded MSA rows:
 * ccm_sparse_steps(Lp, passes) x 2 rows of ccm_sparse_index_ld(Np) words.
 * The forward takes d_idx at the chunk's first sequence (a multiple of 256
 * sequences into the rows); the rows must hold ceil(ncols / 256) * 256
 * sequences from there. */
long ccm_sparse_ws_len(int Lp, int passes);
int ccm_sparse_steps(int Lp, int passes);
long ccm_sparse_index_ld(int Np);
int ccm_sparse_fwd_fits(int Lp, int passes_fwd, int passes, int ncols);
int ccm_launch_convert_ws(const float *d_x, int L, int Lp, uint16_t *d_ws, long lds, int passes,
                          ccm_stream_t s);
/* convert_ws that also writes the a >= 21 padding of the pair gradient
 * (2 lambda_pair x, evaluate_cpu.c:158-161) into d_g and the objective's
 * regularisation (evaluate_cpu.c:151-163) as ccm_convert_ws_reg_parts(L, Lp,
 * passes) partials into d_regpair and one into d_regsingle, so that the
 * assembly (ccm_launch_assemble_fs) leaves both out and forms f itself
 * (the reference's CUDA backend likewise forms the regularisation and
 * starts g before the contractions: evaluate_cuda.c:48-51,
 * gpu_compute_regularization at cu:146-219,691-740) */
int ccm_convert_ws_reg_parts(int L, int Lp, int passes);
int ccm_launch_convert_ws_reg(const float *d_x, float *d_g, int L, int Lp, uint16_t *d_ws, long lds,
                              int passes, float lam1, float lam2, double *d_regpair,
                              double *d_regsingle, int pad, ccm_stream_t s);
/* pad = 0 leaves the padding tiles (their g and partials) to the forward:
 * ccm_launch_sparse_fwd_softmax_pad with d_x, d_g, lam2 and d_regpair runs
 * them in workgroups after its tiles (two per 512-thread workgroup), on the
 * CUs its last, partial round leaves idle */
int ccm_launch_sparse_fwd_softmax_pad(const uint16_t *d_ws, long lds, int Lp, int passes_fwd,
                                      const uint32_t *d_idx, long ldidx, int ncols, const float *d_x1,
                                      const uint8_t *d_msapt, long ldmt, const float *d_w, int L, int Nc,
                                      uint16_t *d_dt, long lddt, long lo_off, int passes, int mode,
                                      int blk0, double *d_fpart, double *d_g1part, const float *d_x,
                                      float *d_g, float lam2, double *d_regpair, ccm_stream_t s);
int ccm_launch_sparse_index(const uint8_t *d_msap, int Lp, int Np, int passes, uint32_t *d_idx,
                            ccm_stream_t s);
/* grouped forms over a batch: one alignment per group entry, descriptors
 * with sparse set (wt = Ws, sidx / ldsidx the positions); the grouped
 * conversion is the ccm_launch_convert_ws_reg form (regpair: the
 * ccm_convert_ws_reg_parts partials, regsingle: one) */
int ccm_convert_ws_blocks(int L, int Lp, int passes);
uint32_t ccm_sparse_fwd_tile_rows(int L);
uint32_t ccm_sparse_fwd_tile_count(int L, int ncols);
int ccm_launch_convert_ws_grouped(const ccm_bprob *d_probs, ccm_group g, int passes, ccm_stream_t s);
int ccm_launch_sparse_fwd_softmax_grouped(const ccm_bprob *d_probs, ccm_group g, ccm_stream_t s);
int ccm_launch_sparse_fwd_softmax(const uint16_t *d_ws, long lds, int Lp, int passes_fwd,
                                  const uint32_t *d_idx, long ldidx, int ncols, const float *d_x1,
                                  const uint8_t *d_msapt, long ldmt, const float *d_w, int L, int Nc,
                                  uint16_t *d_dt, long lddt, long lo_off, int passes, int mode,
                                  int blk0, double *d_fpart, double *d_g1part, ccm_stream_t s);
/* The pair-gradient contraction Gt[c][r] (+)= sum_n Dt[c][n] [x_nk = b]
 * (cu:279-312 without the one-hot term).  Tile variants (ccm_gemm_variants):
 * 0-5 serve both contractions, 6 (192 x 256) the gradient only (the
 * forward launchers run variant 3 for it); every variant sums in the same
 * order, so all give the same bits.  variant < 0: the default, which is
 * one banded launch (256-row tiles over the first rows, 192-row tiles over
 * the rest, one of each per CU) where the CU count favours it, else the
 * cost model's single variant. */
int ccm_launch_gemm_grad(const uint16_t *d_dt, long lda, int Lp, int passes,
                         const uint8_t *d_msapt, long ldmt, int kcols, float *d_gt, long ldc,
                         int accumulate, int variant, ccm_stream_t s);

/* gradient assembly (cu:150-219, cu:390-461) and objective reduction */
int ccm_pair_epilogue_blocks(int L);
int ccm_singles_blocks(int L);
int ccm_launch_pair_epilogue(const float *d_m, long ldm, int Lm, const float *d_hsym,
                             const float *d_x, float *d_g, int L, float lam2, double *d_regpart,
                             ccm_stream_t s);
int ccm_launch_hsym(const float *d_mh, long ldm, float *d_hsym, int L, ccm_stream_t s);
int ccm_launch_singles_finalize(const double *d_g1part, int nblk, const float *d_x, float *d_g,
                                int L, float lam1, double *d_regpart, ccm_stream_t s);
int ccm_reg_only_blocks(void);
int ccm_launch_reg_only(const float *d_x, int L, float lam1, float lam2, double *d_part,
                        ccm_stream_t s);
int ccm_launch_sum3(const double *a, int na, const double *b, int nb, const double *c, int nc,
                    double *out, ccm_stream_t s);
/* pair epilogue and singles in one launch, then the objective's final sum
 * (fixed order) into *d_f */
int ccm_assemble_blocks(int L);
int ccm_launch_assemble(const float *d_m, long ldm, int Lm, const float *d_hsym, const float *d_x,
                        float *d_g, int L, float lam1, float lam2, const double *d_g1part, int n_g1blk,
                        const double *d_fpart, int n_fpart, double *d_regpair, double *d_regsingle,
                        double *d_f, ccm_stream_t s);
/* after ccm_launch_convert_ws_reg: the singles and the a < 21 pair columns
 * of g, and in the same launch *d_f = the sum of d_fpart | d_regpair |
 * d_regsingle in k_sum3's order (the grouped assembly of a sparse
 * descriptor takes this form, ccm_assemble_fs_blocks(L) blocks per
 * alignment, its f into ctl->fx) */
int ccm_assemble_fs_blocks(int L);
int ccm_launch_assemble_fs(const float *d_m, long ldm, int Lm, const float *d_hsym, const float *d_x,
                           float *d_g, int L, float lam1, float lam2, const double *d_g1part,
                           int n_g1blk, const double *d_fpart, int n_fpart, const double *d_regpair,
                           int n_regpair, const double *d_regsingle, int n_regsingle, double *d_f,
                           ccm_stream_t s);

/* conjugate-gradient vector arithmetic */
int ccm_dots_partial_len(void);
int ccm_launch_dots(const float *a, const float *b0, const float *b1, const float *b2, long n,
                    int nd, double *part, double *out, ccm_stream_t s);
int ccm_launch_axpy(float *x, const float *sdir, float alpha, long n, ccm_stream_t s);
int ccm_launch_direction(float *sdir, const float *g, float beta, long n, ccm_stream_t s);
int ccm_launch_negate(float *sdir, const float *g, long n, ccm_stream_t s);
/* The same over the padded layout of an alignment with ncol columns
 * (n = nvar_padded(ncol)), visiting only the float4s that hold singles or
 * pair entries a < 21.  Valid while the a >= 21 padding of the operands is
 * zero, as it stays in a CG whose gradient is evaluate_hip's (2 lambda_pair
 * x there) from an x with zero padding: the results then equal the
 * unpadded launchers' bit for bit (the dots keep their block partition).
 * Any other n falls back to the unpadded launchers.  pad_scan_x sets
 * *d_flag to 1 if x holds a nonzero value in the skipped float4s. */
int ccm_launch_pad_scan_x(const float *x, int ncol, long n, int32_t *d_flag, ccm_stream_t s);
int ccm_launch_dots_padded(const float *a, const float *b0, const float *b1, const float *b2, long n,
                           int nd, int ncol, double *part, double *out, ccm_stream_t s);
int ccm_launch_axpy_padded(float *x, const float *sdir, float alpha, int ncol, long n, ccm_stream_t s);
int ccm_launch_direction_padded(float *sdir, const float *g, float beta, int ncol, long n,
                                ccm_stream_t s);
int ccm_launch_negate_padded(float *sdir, const float *g, int ncol, long n, ccm_stream_t s);

/* contact scores on the device (util.c:16-139): the L x L map of
 * sum_submatrices from the couplings of x, then apc (if apc) and the [0,1]
 * rescaling (if normalize), equal bit for bit to ccm_sum_submatrices /
 * ccm_apc / ccm_normalize on the same x.  d_work holds
 * ccm_contact_map_work_len(L) floats (read only when apc). */
int ccm_contact_map_work_len(int L);
int ccm_launch_contact_map(const float *d_x, int L, int apc, int normalize, float *d_out,
                           float *d_work, ccm_stream_t s);

/* ccm_launch_contact_map without the work buffer, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops).  No
 * counterpart in the reference.  The L-float scratch is a fixed device array
 * of CCM_CONTACT_MAP_NOALLOC_MAX_L floats (32 KiB), one per GPU, held until
 * the process ends; L beyond it returns -5 with a message on stderr.  The
 * stream must belong to the current device.  Calls that overlap on one GPU
 * share that array and race; sequential calls on any streams are safe.
 * Allocates nothing: no hipMalloc, hipFree or allocating library call. */
#define CCM_CONTACT_MAP_NOALLOC_MAX_L 8192
int ccm_contact_map_noalloc_max_l(void);
int ccm_launch_contact_map_noalloc(const float *d_x, int L, int apc, int normalize, float *d_out,
                                   ccm_stream_t s);

/* Grouped forms for a batch of alignments (descriptors in
 * ccmpred_amd_batch.h).  No counterpart in the reference, which runs one
 * alignment per process.  Each step is one launch over every alignment of
 * the group, with the device code of the single-alignment launchers above;
 * d_probs is a device array of ccm_bprob and a group's offs/pidx live in
 * device memory.  Alignments whose control block has eval_on == 0
 * (vec_on == 0 for the vector step) return at once. */
int ccm_convert_wt_blocks(int L, int Lp); /* (21 Lp / 64) x ceil(21 L / 128) */
uint32_t ccm_gemm_tile_count(int variant, int M, int Ncol);
int ccm_gemm_variant_ok(int variant, int mode, int Lp, int kcols);
int ccm_gemm_variant_direct(int variant);
int ccm_gemm_tile_shape(int variant, int *tm, int *tn);
int ccm_gemm_default_variant_for(int mode, int Lp, int ncols);
int ccm_dots_blocks(long n);
int ccm_cg_vec_blocks(long n);
int ccm_launch_convert_wt_grouped(const ccm_bprob *d_probs, ccm_group g, ccm_stream_t s);
int ccm_launch_gemm_fwd_grouped(const ccm_bprob *d_probs, ccm_group g, int variant, ccm_stream_t s);
/* the fused forward over a group (alignments whole, x1 = x): tiles per
 * alignment ccm_fwd_softmax_tile_count, ccm_fwd_softmax_tile_rows row
 * panels of them; register-staged variants map to the direct-to-LDS
 * variant of the same shape */
uint32_t ccm_fwd_softmax_tile_count(int variant, int L, int ncols);
int ccm_fwd_softmax_tile_rows(int variant, int L);
int ccm_launch_gemm_fwd_softmax_grouped(const ccm_bprob *d_probs, ccm_group g, int variant,
                                        ccm_stream_t s);
int ccm_launch_gemm_grad_grouped(const ccm_bprob *d_probs, ccm_group g, int variant, ccm_stream_t s);
/* the gradient contraction of one alignment of a batch (descriptor pi) as
 * the banded launch that ccm_launch_gemm_grad makes by default (256-row
 * tiles, then 192-row tiles; no work once the alignment's solve has
 * ended); returns 1, launching nothing, where its P = 21 Lp takes no
 * banding */
int ccm_launch_gemm_grad_banded_desc(const ccm_bprob *d_probs, int pi, int P, ccm_stream_t s);
int ccm_launch_softmax_cols_grouped(const ccm_bprob *d_probs, ccm_group g, ccm_stream_t s);
int ccm_launch_pair_epilogue_grouped(const ccm_bprob *d_probs, ccm_group g, ccm_stream_t s);
int ccm_launch_assemble_grouped(const ccm_bprob *d_probs, ccm_group g, ccm_stream_t s);
int ccm_launch_singles_finalize_grouped(const ccm_bprob *d_probs, ccm_group g, ccm_stream_t s);
int ccm_launch_sum3_grouped(const ccm_bprob *d_probs, ccm_group g, ccm_stream_t s);
int ccm_launch_dots_grouped(const ccm_bprob *d_probs, ccm_group g, ccm_group gfinal, ccm_stream_t s);
int ccm_launch_cg_vec_grouped(const ccm_bprob *d_probs, ccm_group g, ccm_stream_t s);
/* sets ctl->x_pad_nz of every alignment whose x holds a nonzero value in
 * the a >= 21 padding of the pair block; with the flag clear, the grouped
 * vector step and dots skip that padding (a zero padding stays zero: its
 * gradient is 2 lambda_pair x = 0) */
int ccm_launch_pad_scan(const ccm_bprob *d_probs, int nprob, ccm_stream_t s);
int ccm_launch_cg_ctrl(ccm_bprob *d_probs, int nprob, const ccm_cg_prm *prm, int32_t *d_nactive,
                       ccm_stream_t s);

#ifdef __cplusplus
}
#endif

#endif
