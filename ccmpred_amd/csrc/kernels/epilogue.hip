// Gradient assembly and the objective's global sum.
//
// Pairs (evaluate_cpu.c:131-163): for every padded pair entry (b,k,a,j)
//   g2 = M(b,k,a,j) + M(a,j,b,k) [+ hsym(b,k,a,j)] for a < 21 and k != j,
//   g2 = 0 for self-edges (k == j) and for padding (a >= 21),
//   then g2 += 2 lambda_pair x and reg += 0.5 lambda_pair x^2 over the whole
//   padded block.  M is the raw (unsymmetrised) pair gradient in either the
//   FP32 layout (Lm = L, M(b,k,a,j) = M[b*L + k][a*L + j]) or the padded
//   BF16 layout (Lm = Lp, M(b,k,a,j) = M[a*Lp + j][b*Lp + k]); only the sum
//   of both orientations enters g2, so either works.
// Singles (evaluate_cpu.c:102-109,153-156): g1 = sum of per-block partials
//   + 2 lambda_single V for the 20L real singles, 0 on the padding.
#include "common.h"

#define EP_T 64
// threads of an assembly block (pair tiles and singles): 8 waves, so that a
// thread holds 8 rows of a tile and more waves fit on a SIMD
#define EP_NTH 512
#define EP_W (EP_NTH / 64)

// ---- symmetric form: one block per pair of mirrored tiles ------------------
//
// In the x2 index space (rows R = b*L + k, columns C = a*L + j, both below
// 21L) g2[R][C] = M[phi(C)][phi(R)] + M[phi(R)][phi(C)] + hsym[R][C] + 2 l x[R][C],
// phi(a*L + j) = a*Lm + j, which is symmetric in (R, C) apart from hsym and
// x.  A block takes the tile pair (I, J), I <= J, of 64 x 64 tiles: it
// stages M[phi(I)][phi(J)] + hsym[I][J] and M[phi(J)][phi(I)] in LDS once
// and writes both output tiles, so M is read once and, hsym being
// symmetric, half of hsym.  Blocks past the tile pairs write the padding
// columns C in [64 NT, 32L) (2 l x only).
__device__ __forceinline__ int ccm_phi(int X, int L, int Lm) {
  const int a = X / L;
  return a * Lm + (X - a * L);
}

// tile pairs (ti <= tj) of NT tiles, row by row
__device__ __forceinline__ void ccm_tri_pair(uint32_t idx, int NT, int& ti, int& tj) {
  // ti: largest t with t*NT - t*(t-1)/2 <= idx
  const double n = (double)NT + 0.5;
  int t = (int)(n - sqrt(n * n - 2.0 * (double)idx));
  if (t < 0) t = 0;
  if (t > NT - 1) t = NT - 1;
  while (t > 0 && (uint32_t)(t * NT - t * (t - 1) / 2) > idx) t--;
  while (t + 1 < NT && (uint32_t)((t + 1) * NT - (t + 1) * t / 2) <= idx) t++;
  ti = t;
  tj = t + (int)(idx - (uint32_t)(t * NT - t * (t - 1) / 2));
}

__device__ __forceinline__ int ccm_pair_tiles(int L) { return (CCM_NA * L + EP_T - 1) / EP_T; }

__device__ __forceinline__ void pair_sym_block(const float* __restrict__ M, long ldm, int Lm,
                                               const float* __restrict__ hsym,
                                               const float* __restrict__ x2, float* __restrict__ g2,
                                               int L, float lam2, double* __restrict__ regpart,
                                               uint32_t blk) {
  constexpr int RPT = EP_T / EP_W;  // rows per thread: rows wv, wv + EP_W, ...
  __shared__ float D1[EP_T][EP_T + 1];
  __shared__ float D2[EP_T][EP_T + 1];
  __shared__ double red[EP_W];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int C21 = CCM_NA * L, ld2 = CCM_NAP * L;
  const int NT = ccm_pair_tiles(L);
  const uint32_t npairs = (uint32_t)(NT * (NT + 1) / 2);
  const float two_l = 2.0f * lam2, half_l = 0.5f * lam2;
  double reg = 0.0;
  // one output tile: rows R0 + rr, column C; S(rr) from the staged blocks.
  // All loads of the tile are issued before any result is formed.
  auto out_tile = [&](int R0, int C, bool pair_col, int jC, bool first) {
    float xv[RPT];
#pragma unroll
    for (int i = 0; i < RPT; i++) {
      const int R = R0 + wv + EP_W * i;
      xv[i] = R < C21 ? x2[(long)R * ld2 + C] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < RPT; i++) {
      const int rr = wv + EP_W * i, R = R0 + rr;
      if (R >= C21) continue;
      const int kR = R - (R / L) * L;
      float v = 0.f;
      if (pair_col && kR != jC)
        v = first ? D1[rr][lane] + D2[lane][rr] : D2[rr][lane] + D1[lane][rr];
      g2[(long)R * ld2 + C] = v + two_l * xv[i];
      reg += (double)(half_l * xv[i] * xv[i]);
    }
  };
  if (blk < npairs) {
    int ti, tj;
    ccm_tri_pair(blk, NT, ti, tj);
    // D1[r][c] = M[phi(ti*64 + r)][phi(tj*64 + c)] + hsym[ti*64 + r][tj*64 + c],
    // D2[r][c] = M[phi(tj*64 + r)][phi(ti*64 + c)] (hsym split when ti == tj)
    {
      const int Ci = ti * EP_T + lane, Cj = tj * EP_T + lane;
      const int pci = Ci < C21 ? ccm_phi(Ci, L, Lm) : 0, pcj = Cj < C21 ? ccm_phi(Cj, L, Lm) : 0;
      float d1[RPT], d2[RPT], h1[RPT];
#pragma unroll
      for (int i = 0; i < RPT; i++) {
        const int rr = wv + EP_W * i;
        const int Ri = ti * EP_T + rr, Rj = tj * EP_T + rr;
        const bool ok1 = Ri < C21 && Cj < C21;
        d1[i] = ok1 ? M[(long)ccm_phi(Ri, L, Lm) * ldm + pcj] : 0.f;
        h1[i] = (ok1 && hsym) ? hsym[(long)Ri * ld2 + Cj] : 0.f;
        d2[i] = (Rj < C21 && Ci < C21) ? M[(long)ccm_phi(Rj, L, Lm) * ldm + pci] : 0.f;
      }
      // a diagonal pair (one tile, mirrored onto itself) gives half of hsym
      // to each staged block, so that both orientations add the same two
      // operands and g2 stays exactly symmetric
      const float h1s = ti == tj ? 0.5f : 1.0f, h2s = ti == tj ? 0.5f : 0.0f;
#pragma unroll
      for (int i = 0; i < RPT; i++) {
        D1[wv + EP_W * i][lane] = d1[i] + h1s * h1[i];
        D2[wv + EP_W * i][lane] = d2[i] + h2s * h1[i];
      }
    }
    __syncthreads();
    // tile (ti, tj): rows ti*64 + rr, column tj*64 + lane
    {
      const int C = tj * EP_T + lane;
      if (C < ld2) out_tile(ti * EP_T, C, C < C21, C - (C / L) * L, true);
    }
    // mirrored tile (tj, ti): rows tj*64 + rr, column ti*64 + lane (below C21)
    if (ti != tj) {
      const int C = ti * EP_T + lane;
      out_tile(tj * EP_T, C, true, C - (C / L) * L, false);
    }
  } else {
    // padding columns past the tiles: row tile tr, columns 64 NT + 64 tc + lane
    const int npc = (ld2 - NT * EP_T + EP_T - 1) / EP_T;
    const uint32_t q = blk - npairs;
    const int tr = (int)(q / (uint32_t)npc), tc = (int)(q % (uint32_t)npc);
    const int C = NT * EP_T + tc * EP_T + lane;
    if (C < ld2) out_tile(tr * EP_T, C, false, 0, true);
  }
  reg = ccm_wave_sum_f64(reg);
  if (lane == 0) red[wv] = reg;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < EP_W; w++) t += red[w];
    regpart[blk] = t;
  }
}

// blocks of the symmetric pair assembly: tile pairs, then padding tiles
__host__ __device__ __forceinline__ int ccm_pair_sym_blocks(int L) {
  const int NT = (CCM_NA * L + EP_T - 1) / EP_T;
  const int npc = (CCM_NAP * L - NT * EP_T + EP_T - 1) / EP_T;
  return NT * (NT + 1) / 2 + NT * (npc > 0 ? npc : 0);
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

__device__ __forceinline__ void singles_block(const double* __restrict__ g1part, int nblk,
                                              const float* __restrict__ x1, float* __restrict__ g1,
                                              int L, long nsp, float lam1, double* __restrict__ regpart,
                                              uint32_t blk) {
  __shared__ double red[EP_W];
  const long v = (long)blk * EP_NTH + threadIdx.x;
  const long ns = 20L * L;
  double reg = 0.0;
  if (v < ns) {
    // blocks summed in order; eight loads in flight at a time
    double s = 0.0;
    int b = 0;
    for (; b + 8 <= nblk; b += 8) {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; u++) t[u] = g1part[(long)(b + u) * ns + v];
#pragma unroll
      for (int u = 0; u < 8; u++) s += t[u];
    }
    for (; b < nblk; b++) s += g1part[(long)b * ns + v];
    const float xv = x1[v];
    g1[v] = (float)s + (2.0f * lam1) * xv;
    reg = (double)(lam1 * xv * xv);
  } else if (v < nsp) {
    g1[v] = 0.f;
  }
  reg = ccm_wave_sum_f64(reg);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = reg;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < EP_W; w++) t += red[w];
    regpart[blk] = t;
  }
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
#define SUM_NT 1024
__device__ __forceinline__ void sum3_block(const double* __restrict__ a, int na,
                                           const double* __restrict__ b, int nb,
                                           const double* __restrict__ c, int nc,
                                           double* __restrict__ out) {
  __shared__ double sh[SUM_NT], shc[SUM_NT];
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
  for (int i0 = threadIdx.x; i0 < n; i0 += SUM_NT * 8) {
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
  sh[threadIdx.x] = s;
  shc[threadIdx.x] = comp;
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
  assemble_block(B.gt, B.P, B.Lp, B.hsym, B.x, B.g, B.L, B.nsp, B.lam1, B.lam2, B.g1part, B.n_g1blk,
                 B.regpair, B.regsingle, blockIdx.x - offs[gi]);
}

extern "C" {

// blocks of the fused assembly: pair tiles, then singles
int ccm_assemble_blocks(int L) { return ccm_pair_epilogue_blocks(L) + ccm_singles_blocks(L); }

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
