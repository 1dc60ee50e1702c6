// Gradient assembly and the objective's global sum: the assembly kernels
// over the work items of assemble.h (evaluate_cpu.c:131-163 pairs,
// evaluate_cpu.c:102-109,153-156 singles).
#include "assemble.h"

// LDS of an assembly workgroup: the two staged tiles of a pair item (the
// objective's sum block of the FS form reuses them)
struct ep_lds {
  float D1s[EP_TILE_FLOATS];
  float D2s[EP_TILE_FLOATS];
  double red[EP_W];
};
__device__ __forceinline__ ep_lds& ep_shared() {
  __shared__ ep_lds sh;
  return sh;
}

// one pair item per workgroup
template <bool FS = false>
__device__ __forceinline__ void pair_sym_block(const float* __restrict__ M, long ldm, int Lm,
                                               const float* __restrict__ hsym,
                                               const float* __restrict__ x2, float* __restrict__ g2,
                                               int L, float lam2, double* __restrict__ regpart,
                                               uint32_t blk) {
  ep_lds& sh = ep_shared();
  const uint32_t blks[1] = {blk};
  float* const d1[1] = {sh.D1s};
  float* const d2[1] = {sh.D2s};
  pair_sym_items<1, FS>(M, ldm, Lm, hsym, x2, g2, L, lam2, regpart, blks, d1, d2, sh.red);
}

__global__ __launch_bounds__(EP_NTH) void k_pair_epilogue(const float* __restrict__ M, long ldm, int Lm,
                                                       const float* __restrict__ hsym,
                                                       const float* __restrict__ x2,
                                                       float* __restrict__ g2, int L, float lam2,
                                                       double* __restrict__ regpart) {
  pair_sym_block(M, ldm, Lm, hsym, x2, g2, L, lam2, regpart, blockIdx.x);
}

// BF16 layout of every alignment of a group: pair_epilogue_blocks(L) blocks each
__global__ __launch_bounds__(EP_NTH) void k_pair_epilogue_grouped(const ccm_bprob* __restrict__ probs,
                                                               const uint32_t* __restrict__ offs,
                                                               const int32_t* __restrict__ pidx, int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  pair_sym_block(B.gt, B.P, B.Lp, B.hsym, B.x + B.nsp, B.g + B.nsp, B.L, B.lam2, B.regpair,
                 blockIdx.x - offs[gi]);
}

// symmetrised -2 H2 in the x2 layout with zero self-edge blocks and padding:
// hsym(b,k,a,j) = Mh(b,k,a,j) + Mh(a,j,b,k) with Mh = -H2 (FP32 layout)
__global__ __launch_bounds__(256) void k_hsym(const float* __restrict__ Mh, long ldm,
                                              float* __restrict__ hsym, int L, uint32_t tiles_c) {
  __shared__ float T[EP_T][EP_T + 1];
  const int tr = blockIdx.x / tiles_c, tc = blockIdx.x % tiles_c;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int C21 = CCM_NA * L, ld2 = CCM_NAP * L;
  const float invL = 1.0f / (float)L;
  {
    const int R = tr * EP_T + lane;
    for (int cc = wv; cc < EP_T; cc += 4) {
      const int Cp = tc * EP_T + cc;
      T[cc][lane] = (Cp < C21 && R < C21) ? Mh[(long)Cp * ldm + R] : 0.f;
    }
  }
  __syncthreads();
  const int Cp = tc * EP_T + lane;
  const int a = __float2int_rd(((float)Cp + 0.5f) * invL);
  const int j = Cp - a * L;
  for (int rr = wv; rr < EP_T; rr += 4) {
    const int R = tr * EP_T + rr;
    if (R >= C21 || Cp >= ld2) continue;
    const int b = __float2int_rd(((float)R + 0.5f) * invL);
    const int k = R - b * L;
    float v = 0.f;
    if (Cp < C21 && k != j) v = Mh[(long)R * ldm + Cp] + T[lane][rr];
    hsym[(long)R * ld2 + Cp] = v;
  }
}

template <bool FS = false>
__device__ __forceinline__ void singles_block(const double* __restrict__ g1part, int nblk,
                                              const float* __restrict__ x1, float* __restrict__ g1,
                                              int L, long nsp, float lam1, double* __restrict__ regpart,
                                              uint32_t blk) {
  singles_item<FS>(g1part, nblk, x1, g1, L, nsp, lam1, regpart, blk, ep_shared().red);
}

__global__ __launch_bounds__(EP_NTH) void k_singles_finalize(const double* __restrict__ g1part, int nblk,
                                                          const float* __restrict__ x1,
                                                          float* __restrict__ g1, int L, long nsp,
                                                          float lam1, double* __restrict__ regpart) {
  singles_block(g1part, nblk, x1, g1, L, nsp, lam1, regpart, blockIdx.x);
}

// every alignment of a group: singles_blocks(L) blocks each
__global__ __launch_bounds__(EP_NTH) void k_singles_finalize_grouped(const ccm_bprob* __restrict__ probs,
                                                                  const uint32_t* __restrict__ offs,
                                                                  const int32_t* __restrict__ pidx,
                                                                  int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  singles_block(B.g1part, B.n_g1blk, B.x, B.g, B.L, B.nsp, B.lam1, B.regsingle, blockIdx.x - offs[gi]);
}

// regularisation terms alone: lam1 sum V^2 + 0.5 lam2 sum x2^2 (padded block)
__global__ __launch_bounds__(256) void k_reg_only(const float* __restrict__ x, long ns, long nsp,
                                                  long nvp, float lam1, float lam2,
                                                  double* __restrict__ part) {
  __shared__ double red[4];
  double r = 0.0;
  const float half_l = 0.5f * lam2;
  for (long v = (long)blockIdx.x * 256 + threadIdx.x; v < nvp; v += (long)gridDim.x * 256) {
    const float xv = x[v];
    if (v < ns)
      r += (double)(lam1 * xv * xv);
    else if (v >= nsp)
      r += (double)(half_l * xv * xv);
  }
  r = ccm_wave_sum_f64(r);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = r;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Neumaier-compensated FP64 sum of up to three partial arrays, fixed order:
// thread t of SUM_NT sums elements t, t + SUM_NT, ... of the concatenation
// a|b|c (eight loads issued before their adds), then wave 0 folds the
// SUM_NT thread sums, SUM_NT/64 per lane, and a lane tree combines the 64
// (sum, compensation) pairs with exact two-sums.  out[0] = total (double)
// A workgroup of NT threads (a divisor of SUM_NT) computes the same sum
// bit for bit: thread t takes the SUM_NT / NT virtual threads t, t + NT, ..
// in turn.  sh, shc: SUM_NT doubles of LDS each.
#define SUM_NT 1024
template <int NT>
__device__ __forceinline__ void sum3_block_lds(const double* __restrict__ a, int na,
                                               const double* __restrict__ b, int nb,
                                               const double* __restrict__ c, int nc,
                                               double* __restrict__ out, double* sh, double* shc) {
  static_assert(SUM_NT % NT == 0 && NT >= 64, "virtual threads");
  double s = 0.0, comp = 0.0;
  auto add = [&](double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v))
      comp += (s - t) + v;
    else
      comp += (v - t) + s;
    s = t;
  };
  const int n = na + nb + nc;
  for (int vt = threadIdx.x; vt < SUM_NT; vt += NT) {
    s = 0.0;
    comp = 0.0;
    for (int i0 = vt; i0 < n; i0 += SUM_NT * 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int i = i0 + u * SUM_NT;
        v[u] = i < na ? a[i] : i < na + nb ? b[i - na] : i < n ? c[i - na - nb] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (i0 + u * SUM_NT < n) add(v[u]);
    }
    sh[vt] = s;
    shc[vt] = comp;
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    constexpr int PER = SUM_NT / 64;
    const int l = threadIdx.x;
    s = 0.0;
    comp = 0.0;
#pragma unroll
    for (int u = 0; u < PER; u++) {
      add(sh[PER * l + u]);
      comp += shc[PER * l + u];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double s2 = __shfl_down(s, off, 64), c2 = __shfl_down(comp, off, 64);
      const double t = s + s2, bp = t - s;
      const double e = (s - (t - bp)) + (s2 - bp);
      s = t;
      comp = comp + c2 + e;
    }
    if (l == 0) out[0] = s + comp;
  }
}

__device__ __forceinline__ void sum3_block(const double* __restrict__ a, int na,
                                           const double* __restrict__ b, int nb,
                                           const double* __restrict__ c, int nc,
                                           double* __restrict__ out) {
  __shared__ double sh[SUM_NT], shc[SUM_NT];
  sum3_block_lds<SUM_NT>(a, na, b, nb, c, nc, out, sh, shc);
}

__global__ __launch_bounds__(SUM_NT) void k_sum3(const double* __restrict__ a, int na,
                                                 const double* __restrict__ b, int nb,
                                                 const double* __restrict__ c, int nc,
                                                 double* __restrict__ out) {
  sum3_block(a, na, b, nb, c, nc, out);
}

// one block per alignment of a group: f = softmax partials + pair and
// single regularisation partials, into the alignment's control block
__global__ __launch_bounds__(SUM_NT) void k_sum3_grouped(const ccm_bprob* __restrict__ probs,
                                                      const int32_t* __restrict__ pidx) {
  const ccm_bprob& B = probs[pidx[blockIdx.x]];
  if (!B.ctl->eval_on) return;
  sum3_block(B.fpart, B.n_fpart, B.regpair, B.n_regpair, B.regsingle, B.n_regsingle, &B.ctl->fx);
}

// ---- gradient assembly: pair tiles and singles in one launch --------------
//
// Blocks [0, npair) assemble pair tiles, blocks [npair, npair + nsing) the
// singles; each leaves its regularisation partial.  The objective's final
// sum is a separate one-block launch (k_sum3): a last-block reduction here
// would need a device-scope fence in every block, which writes back the L2
// of the block's XCD and costs more than the launch it saves.
extern "C" int ccm_pair_epilogue_blocks(int L);
extern "C" int ccm_singles_blocks(int L);

__device__ __forceinline__ void assemble_block(const float* __restrict__ M, long ldm, int Lm,
                                               const float* __restrict__ hsym,
                                               const float* __restrict__ x, float* __restrict__ g,
                                               int L, long nsp, float lam1, float lam2,
                                               const double* __restrict__ g1part, int n_g1blk,
                                               double* __restrict__ regpair,
                                               double* __restrict__ regsingle, uint32_t blk) {
  // the few singles blocks first, so that their partial sums overlap the
  // pair blocks instead of trailing them
  const uint32_t nsing = (uint32_t)((nsp + EP_NTH - 1) / EP_NTH);  // ccm_singles_blocks(L)
  if (blk < nsing)
    singles_block(g1part, n_g1blk, x, g, L, nsp, lam1, regsingle, blk);
  else
    pair_sym_block(M, ldm, Lm, hsym, x + nsp, g + nsp, L, lam2, regpair, blk - nsing);
}

// FS form (the coupling conversion wrote the regularisation partials and
// the a >= 21 padding of g): singles, then the tile pairs of the a < 21
// columns, no partials
__device__ __forceinline__ void assemble_block_fs(const float* __restrict__ M, long ldm, int Lm,
                                                  const float* __restrict__ hsym,
                                                  const float* __restrict__ x, float* __restrict__ g,
                                                  int L, long nsp, float lam1, float lam2,
                                                  const double* __restrict__ g1part, int n_g1blk,
                                                  uint32_t blk) {
  const uint32_t nsing = (uint32_t)((nsp + EP_NTH - 1) / EP_NTH);
  if (blk < nsing)
    singles_block<true>(g1part, n_g1blk, x, g, L, nsp, lam1, nullptr, blk);
  else
    pair_sym_block<true>(M, ldm, Lm, hsym, x + nsp, g + nsp, L, lam2, nullptr, blk - nsing);
}

// the single-alignment FS launch: block 0 forms the objective (the sum of
// k_sum3, over the softmax partials and the conversion's regularisation
// partials, which earlier launches wrote) while the others assemble g
__global__ __launch_bounds__(EP_NTH) void k_assemble_fs(const float* __restrict__ M, long ldm, int Lm,
                                                     const float* __restrict__ hsym,
                                                     const float* __restrict__ x, float* __restrict__ g,
                                                     int L, long nsp, float lam1, float lam2,
                                                     const double* __restrict__ g1part, int n_g1blk,
                                                     const double* __restrict__ fpart, int n_fpart,
                                                     const double* __restrict__ regpair, int n_regpair,
                                                     const double* __restrict__ regsingle,
                                                     int n_regsingle, double* __restrict__ f_out) {
  if (blockIdx.x == 0) {
    static_assert(EP_TILE_FLOATS * sizeof(float) >= SUM_NT * sizeof(double), "sum in the tile buffers");
    ep_lds& sh = ep_shared();
    sum3_block_lds<EP_NTH>(fpart, n_fpart, regpair, n_regpair, regsingle, n_regsingle, f_out,
                           reinterpret_cast<double*>(sh.D1s), reinterpret_cast<double*>(sh.D2s));
    return;
  }
  assemble_block_fs(M, ldm, Lm, hsym, x, g, L, nsp, lam1, lam2, g1part, n_g1blk, blockIdx.x - 1);
}

__global__ __launch_bounds__(EP_NTH) void k_assemble(const float* __restrict__ M, long ldm, int Lm,
                                                  const float* __restrict__ hsym,
                                                  const float* __restrict__ x, float* __restrict__ g,
                                                  int L, long nsp, float lam1, float lam2,
                                                  const double* __restrict__ g1part, int n_g1blk,
                                                  double* __restrict__ regpair,
                                                  double* __restrict__ regsingle) {
  assemble_block(M, ldm, Lm, hsym, x, g, L, nsp, lam1, lam2, g1part, n_g1blk, regpair, regsingle,
                 blockIdx.x);
}

// BF16 layout of every alignment of a group: assemble_blocks(L) blocks each
__global__ __launch_bounds__(EP_NTH) void k_assemble_grouped(const ccm_bprob* __restrict__ probs,
                                                          const uint32_t* __restrict__ offs,
                                                          const int32_t* __restrict__ pidx, int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  // the sparse form's conversion wrote the regularisation partials and the
  // padding; the alignment's first block forms f as k_assemble_fs does
  if (B.sparse) {
    const uint32_t blk = blockIdx.x - offs[gi];
    if (blk == 0) {
      ep_lds& sh = ep_shared();
      sum3_block_lds<EP_NTH>(B.fpart, B.n_fpart, B.regpair, B.n_regpair, B.regsingle, B.n_regsingle, &B.ctl->fx,
                             reinterpret_cast<double*>(sh.D1s), reinterpret_cast<double*>(sh.D2s));
      return;
    }
    assemble_block_fs(B.gt, B.P, B.Lp, B.hsym, B.x, B.g, B.L, B.nsp, B.lam1, B.lam2, B.g1part, B.n_g1blk,
                      blk - 1);
  } else
    assemble_block(B.gt, B.P, B.Lp, B.hsym, B.x, B.g, B.L, B.nsp, B.lam1, B.lam2, B.g1part, B.n_g1blk,
                   B.regpair, B.regsingle, blockIdx.x - offs[gi]);
}

extern "C" {

// blocks of the fused assembly: pair tiles, then singles
int ccm_assemble_blocks(int L) { return ccm_pair_epilogue_blocks(L) + ccm_singles_blocks(L); }

// blocks of the FS assembly of one alignment: the objective's sum, the
// singles, the tile pairs
int ccm_assemble_fs_blocks(int L) { return 1 + ccm_pair_sym_blocks_fs(L) + ccm_singles_blocks(L); }

// FS form: g from the raw pair gradient, hsym and the node-gradient
// partials for the singles and the a < 21 pair columns (the coupling
// conversion ccm_launch_convert_ws_reg wrote the a >= 21 padding of g and
// the regularisation partials), and in the same launch
// *d_f = sum of d_fpart | d_regpair | d_regsingle (the order of k_sum3)
int ccm_launch_assemble_fs(const float* d_m, long ldm, int Lm, const float* d_hsym, const float* d_x,
                           float* d_g, int L, float lam1, float lam2, const double* d_g1part, int n_g1blk,
                           const double* d_fpart, int n_fpart, const double* d_regpair, int n_regpair,
                           const double* d_regsingle, int n_regsingle, double* d_f, hipStream_t s) {
  if (L <= 0 || !d_m || !d_x || !d_g || !d_f || n_fpart < 0 || n_regpair < 0 || n_regsingle < 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_assemble_fs, dim3(ccm_assemble_fs_blocks(L)), dim3(EP_NTH), 0, s, d_m, ldm, Lm, d_hsym,
                     d_x, d_g, L, ccm_nsp(L), lam1, lam2, d_g1part, n_g1blk, d_fpart, n_fpart, d_regpair,
                     n_regpair, d_regsingle, n_regsingle, d_f);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// g (pairs and singles) from the raw pair gradient M (FP32 layout Lm = L,
// or the padded BF16 layout Lm = Lp with hsym), the node-gradient partials
// and the regularisation, then *d_f = sum of the softmax partials and all
// regularisation partials (second, one-block launch).  d_regpair:
// ccm_pair_epilogue_blocks(L) doubles, d_regsingle: ccm_singles_blocks(L).
int ccm_launch_assemble(const float* d_m, long ldm, int Lm, const float* d_hsym, const float* d_x,
                        float* d_g, int L, float lam1, float lam2, const double* d_g1part, int n_g1blk,
                        const double* d_fpart, int n_fpart, double* d_regpair, double* d_regsingle,
                        double* d_f, hipStream_t s) {
  if (L <= 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_assemble, dim3(ccm_assemble_blocks(L)), dim3(EP_NTH), 0, s, d_m, ldm, Lm, d_hsym, d_x,
                     d_g, L, ccm_nsp(L), lam1, lam2, d_g1part, n_g1blk, d_regpair, d_regsingle);
  CCM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sum3, dim3(1), dim3(SUM_NT), 0, s, d_fpart, n_fpart, d_regpair, ccm_pair_epilogue_blocks(L),
                     d_regsingle, ccm_singles_blocks(L), d_f);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_assemble_grouped(const ccm_bprob* d_probs, ccm_group g, hipStream_t s) {
  if (g.np <= 0 || g.total == 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_assemble_grouped, dim3(g.total), dim3(EP_NTH), 0, s, d_probs, g.offs, g.pidx, g.np);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"

extern "C" {

int ccm_pair_epilogue_blocks(int L) { return ccm_pair_sym_blocks(L); }

int ccm_singles_blocks(int L) { return (int)((ccm_nsp(L) + EP_NTH - 1) / EP_NTH); }

int ccm_launch_pair_epilogue(const float* d_m, long ldm, int Lm, const float* d_hsym, const float* d_x,
                             float* d_g, int L, float lam2, double* d_regpart, hipStream_t s) {
  const long nsp = ccm_nsp(L);
  hipLaunchKernelGGL(k_pair_epilogue, dim3(ccm_pair_sym_blocks(L)), dim3(EP_NTH), 0, s, d_m, ldm, Lm, d_hsym,
                     d_x + nsp, d_g + nsp, L, lam2, d_regpart);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// d_hsym: pair block (21L x 32L floats) in the x2 layout
int ccm_launch_hsym(const float* d_mh, long ldm, float* d_hsym, int L, hipStream_t s) {
  const uint32_t tr = (CCM_NA * L + EP_T - 1) / EP_T, tc = (CCM_NAP * L + EP_T - 1) / EP_T;
  hipLaunchKernelGGL(k_hsym, dim3(tr * tc), dim3(256), 0, s, d_mh, ldm, d_hsym, L, tc);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_singles_finalize(const double* d_g1part, int nblk, const float* d_x, float* d_g, int L,
                                float lam1, double* d_regpart, hipStream_t s) {
  const long nsp = ccm_nsp(L);
  hipLaunchKernelGGL(k_singles_finalize, dim3((unsigned)ccm_singles_blocks(L)), dim3(EP_NTH), 0, s, d_g1part,
                     nblk, d_x, d_g, L, nsp, lam1, d_regpart);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_reg_only_blocks(void) { return 1024; }

int ccm_launch_reg_only(const float* d_x, int L, float lam1, float lam2, double* d_part, hipStream_t s) {
  const long nsp = ccm_nsp(L);
  const long nvp = nsp + (long)L * L * CCM_NA * CCM_NAP;
  hipLaunchKernelGGL(k_reg_only, dim3(1024), dim3(256), 0, s, d_x, 20L * L, nsp, nvp, lam1, lam2, d_part);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_sum3(const double* a, int na, const double* b, int nb, const double* c, int nc,
                    double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_sum3, dim3(1), dim3(SUM_NT), 0, s, a, na, b, nb, c, nc, out);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_pair_epilogue_grouped(const ccm_bprob* d_probs, ccm_group g, hipStream_t s) {
  if (g.np <= 0 || g.total == 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_pair_epilogue_grouped, dim3(g.total), dim3(EP_NTH), 0, s, d_probs, g.offs, g.pidx, g.np);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_singles_finalize_grouped(const ccm_bprob* d_probs, ccm_group g, hipStream_t s) {
  if (g.np <= 0 || g.total == 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_singles_finalize_grouped, dim3(g.total), dim3(EP_NTH), 0, s, d_probs, g.offs, g.pidx,
                     g.np);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// g.total is ignored: one block per group entry
int ccm_launch_sum3_grouped(const ccm_bprob* d_probs, ccm_group g, hipStream_t s) {
  if (g.np <= 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_sum3_grouped, dim3(g.np), dim3(SUM_NT), 0, s, d_probs, g.pidx);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"
