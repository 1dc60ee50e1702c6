// Work items of the gradient assembly (evaluate_cpu.c:131-163 pairs,
// evaluate_cpu.c:102-109,153-156 singles) for the assembly kernels of the
// single-alignment and the batched form.
//
// Pairs: for every padded pair entry (b,k,a,j)
//   g2 = M(b,k,a,j) + M(a,j,b,k) [+ hsym(b,k,a,j)] for a < 21 and k != j,
//   g2 = 0 for self-edges (k == j) and for padding (a >= 21),
//   then g2 += 2 lambda_pair x and reg += 0.5 lambda_pair x^2 over the whole
//   padded block.  M is the raw (unsymmetrised) pair gradient in either the
//   FP32 layout (Lm = L, M(b,k,a,j) = M[b*L + k][a*L + j]) or the padded
//   BF16 layout (Lm = Lp, M(b,k,a,j) = M[a*Lp + j][b*Lp + k]); only the sum
//   of both orientations enters g2, so either works.
// Singles: g1 = sum of per-block partials + 2 lambda_single V for the 20L
//   real singles, 0 on the padding.
// The FS form of the items (the sparse BF16 path) leaves out the a >= 21
// padding columns and the regularisation: its coupling conversion and
// forward write those (kernels/sparse_fwd.hip).
#pragma once

#include "common.h"

#define EP_T 64
// threads of an assembly workgroup (pair tiles and singles): 8 waves, so
// that a thread holds 8 rows of a tile and more waves fit on a SIMD
#define EP_NTH 512
#define EP_W (EP_NTH / 64)
// floats of one staged 64 x 64 tile (rows padded to 65 against bank conflicts)
#define EP_TILE_FLOATS (EP_T * (EP_T + 1))

// ---- symmetric form: one item per pair of mirrored tiles -------------------
//
// In the x2 index space (rows R = b*L + k, columns C = a*L + j, both below
// 21L) g2[R][C] = M[phi(C)][phi(R)] + M[phi(R)][phi(C)] + hsym[R][C] + 2 l x[R][C],
// phi(a*L + j) = a*Lm + j, which is symmetric in (R, C) apart from hsym and
// x.  Item blk < NT (NT + 1) / 2 takes the tile pair (I, J), I <= J, of
// 64 x 64 tiles: it stages M[phi(I)][phi(J)] + hsym[I][J] and
// M[phi(J)][phi(I)] in LDS once and writes both output tiles, so M is read
// once and, hsym being symmetric, half of hsym.  The items past the tile
// pairs write the padding columns C in [64 NT, 32L) (2 l x only).
// Division of 0 <= n < 33 L by L through a multiply-high: with
// m = ceil(2^32 / L) and e = m L - 2^32 < L, floor(n m / 2^32) = floor(n / L)
// while n e < 2^32, which holds for n < 33 L when L <= 8192 (larger L
// divide).  Integer division is a long VALU sequence; the assembly items
// take several per row.
struct ccm_div_l {
  uint32_t m;
  int L;
  __device__ __forceinline__ explicit ccm_div_l(int l)
      : m(l <= 8192 ? (uint32_t)((0x100000000ull + (uint64_t)l - 1) / (uint64_t)l) : 0u), L(l) {}
  __device__ __forceinline__ int div(int n) const {
    return m ? (int)__umulhi((uint32_t)n, m) : n / L;
  }
  __device__ __forceinline__ int mod(int n) const { return n - div(n) * L; }
  // phi(X) = (X / L) Lm + X % L
  __device__ __forceinline__ int phi(int X, int Lm) const {
    const int a = div(X);
    return a * Lm + (X - a * L);
  }
};

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

// index of the tile pair (ti <= tj) in ccm_tri_pair's order
__device__ __forceinline__ uint32_t ccm_tri_index(int ti, int tj, int NT) {
  return (uint32_t)(ti * NT - ti * (ti - 1) / 2 + (tj - ti));
}

__device__ __forceinline__ int ccm_pair_tiles(int L) { return (CCM_NA * L + EP_T - 1) / EP_T; }

// K pair items at once (blks[k]; D1[k], D2[k]: EP_TILE_FLOATS each; red:
// K * EP_W doubles): every load of the K items is issued before the first
// result is formed, and each item's arithmetic, its partial sum's order
// included, is that of the item alone, so that K changes no bit.
// regpart[blks[k]] receives the item's regularisation partial.
// FS (the form whose regularisation and padding gradient the coupling
// conversion wrote): items cover the a < 21 columns only and leave no
// regularisation partial.
template <int K, bool FS = false>
__device__ __forceinline__ void pair_sym_items(const float* __restrict__ M, long ldm, int Lm,
                                               const float* __restrict__ hsym,
                                               const float* __restrict__ x2, float* __restrict__ g2,
                                               int L, float lam2, double* __restrict__ regpart,
                                               const uint32_t (&blks)[K], float* const (&D1)[K],
                                               float* const (&D2)[K], double* red) {
  constexpr int RPT = EP_T / EP_W;  // rows per thread: rows wv, wv + EP_W, ...
  constexpr int LD = EP_T + 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int C21 = CCM_NA * L, ld2 = CCM_NAP * L;
  const int NT = ccm_pair_tiles(L);
  const uint32_t npairs = (uint32_t)(NT * (NT + 1) / 2);
  const float two_l = 2.0f * lam2, half_l = 0.5f * lam2;
  const ccm_div_l dl(L);
  // per item: output tile a (rows Ra0 + rr, column Ca) and, for an
  // off-diagonal pair, the mirrored tile b (rows Rb0 + rr, column Cb)
  bool pair[K], has_a[K], has_b[K], pc_a[K];
  int Ra0[K], Ca[K], ja[K], Rb0[K], Cb[K], jb[K];
  float h1s[K], h2s[K];
  float d1[K][RPT], d2[K][RPT], h1[K][RPT], xa[K][RPT], xb[K][RPT];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const uint32_t blk = blks[k];
    pair[k] = blk < npairs;
    if (pair[k]) {
      int ti, tj;
      ccm_tri_pair(blk, NT, ti, tj);
      Ra0[k] = ti * EP_T;
      Ca[k] = tj * EP_T + lane;
      has_a[k] = Ca[k] < (FS ? C21 : ld2);
      pc_a[k] = Ca[k] < C21;
      ja[k] = dl.mod(Ca[k]);
      has_b[k] = ti != tj;
      Rb0[k] = tj * EP_T;
      Cb[k] = ti * EP_T + lane;
      jb[k] = dl.mod(Cb[k]);
      // d1[r] = M[phi(ti*64 + r)][phi(tj*64 + lane)], h1 = hsym there,
      // d2[r] = M[phi(tj*64 + r)][phi(ti*64 + lane)]
      const int Ci = ti * EP_T + lane, Cj = tj * EP_T + lane;
      const int pci = Ci < C21 ? dl.phi(Ci, Lm) : 0, pcj = Cj < C21 ? dl.phi(Cj, Lm) : 0;
#pragma unroll
      for (int i = 0; i < RPT; i++) {
        const int rr = wv + EP_W * i;
        const int Ri = ti * EP_T + rr, Rj = tj * EP_T + rr;
        const bool ok1 = Ri < C21 && Cj < C21;
        d1[k][i] = ok1 ? M[(long)dl.phi(Ri, Lm) * ldm + pcj] : 0.f;
        h1[k][i] = (ok1 && hsym) ? hsym[(long)Ri * ld2 + Cj] : 0.f;
        d2[k][i] = (Rj < C21 && Ci < C21) ? M[(long)dl.phi(Rj, Lm) * ldm + pci] : 0.f;
      }
      // a diagonal pair (one tile, mirrored onto itself) gives half of hsym
      // to each staged block, so that both orientations add the same two
      // operands and g2 stays exactly symmetric
      h1s[k] = ti == tj ? 0.5f : 1.0f;
      h2s[k] = ti == tj ? 0.5f : 0.0f;
    } else {
      // padding columns past the tiles: row tile tr, columns 64 NT + 64 tc + lane
      const int npc = (ld2 - NT * EP_T + EP_T - 1) / EP_T;
      const uint32_t q = blk - npairs;
      const int tr = (int)(q / (uint32_t)npc), tc = (int)(q % (uint32_t)npc);
      Ra0[k] = tr * EP_T;
      Ca[k] = NT * EP_T + tc * EP_T + lane;
      has_a[k] = Ca[k] < ld2;
      pc_a[k] = false;
      ja[k] = 0;
      has_b[k] = false;
      Rb0[k] = 0;
      Cb[k] = 0;
      jb[k] = 0;
      h1s[k] = h2s[k] = 0.f;
    }
#pragma unroll
    for (int i = 0; i < RPT; i++) {
      const int Ra = Ra0[k] + wv + EP_W * i, Rb = Rb0[k] + wv + EP_W * i;
      xa[k][i] = (has_a[k] && Ra < C21) ? x2[(long)Ra * ld2 + Ca[k]] : 0.f;
      xb[k][i] = (has_b[k] && Rb < C21) ? x2[(long)Rb * ld2 + Cb[k]] : 0.f;
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++)
    if (pair[k]) {
#pragma unroll
      for (int i = 0; i < RPT; i++) {
        const int rr = wv + EP_W * i;
        D1[k][rr * LD + lane] = d1[k][i] + h1s[k] * h1[k][i];
        D2[k][rr * LD + lane] = d2[k][i] + h2s[k] * h1[k][i];
      }
    }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    double reg = 0.0;
    if (has_a[k]) {
#pragma unroll
      for (int i = 0; i < RPT; i++) {
        const int rr = wv + EP_W * i, R = Ra0[k] + rr;
        if (R >= C21) continue;
        const int kR = dl.mod(R);
        float v = 0.f;
        if (pc_a[k] && kR != ja[k]) v = D1[k][rr * LD + lane] + D2[k][lane * LD + rr];
        g2[(long)R * ld2 + Ca[k]] = v + two_l * xa[k][i];
        if (!FS) reg += (double)(half_l * xa[k][i] * xa[k][i]);
      }
    }
    if (has_b[k]) {
#pragma unroll
      for (int i = 0; i < RPT; i++) {
        const int rr = wv + EP_W * i, R = Rb0[k] + rr;
        if (R >= C21) continue;
        const int kR = dl.mod(R);
        float v = 0.f;
        if (kR != jb[k]) v = D2[k][rr * LD + lane] + D1[k][lane * LD + rr];
        g2[(long)R * ld2 + Cb[k]] = v + two_l * xb[k][i];
        if (!FS) reg += (double)(half_l * xb[k][i] * xb[k][i]);
      }
    }
    if (FS) continue;
    reg = ccm_wave_sum_f64(reg);
    if (lane == 0) red[k * EP_W + wv] = reg;
  }
  if (FS) return;
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < EP_W; w++) t += red[k * EP_W + w];
      regpart[blks[k]] = t;
    }
  }
}

// items of the symmetric pair assembly: tile pairs, then padding tiles
__host__ __device__ __forceinline__ int ccm_pair_sym_blocks(int L) {
  const int NT = (CCM_NA * L + EP_T - 1) / EP_T;
  const int npc = (CCM_NAP * L - NT * EP_T + EP_T - 1) / EP_T;
  return NT * (NT + 1) / 2 + NT * (npc > 0 ? npc : 0);
}

// items of the FS form: tile pairs only
__host__ __device__ __forceinline__ int ccm_pair_sym_blocks_fs(int L) {
  const int NT = (CCM_NA * L + EP_T - 1) / EP_T;
  return NT * (NT + 1) / 2;
}

// singles item blk: EP_NTH singles (red: EP_W doubles); FS: no
// regularisation partial
template <bool FS = false>
__device__ __forceinline__ void singles_item(const double* __restrict__ g1part, int nblk,
                                             const float* __restrict__ x1, float* __restrict__ g1, int L,
                                             long nsp, float lam1, double* __restrict__ regpart,
                                             uint32_t blk, double* red) {
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
  if (FS) return;
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
