// BF16 MFMA form of the two PLL contractions on gfx950.
//
// Index spaces use a padded column count Lp (multiple of 32): the
// coupling index r = b*Lp + k and the output index c = a*Lp + j, so that
// P = 21*Lp.  Both contractions are C[m][col] = sum_kk A[m][kk] * B[kk][col]
// with a dense BF16 operand A (rows m = c, K contiguous, LDS-staged) and a
// one-hot operand B that is never stored: each lane builds its MFMA
// fragment from MSA bytes in registers.
//   forward : A = Wt[c][r] = W(b,k,a,j), B[r][n] = [x_nk = b]  -> Zt[c][n]
//             (evaluate_cpu.c:55-75; Z = A_onehot . W transposed)
//   gradient: A = Dt[c][n] = w_n P_n(a,j), B[n][r] = [x_nk = b] -> Gt[c][r]
//             (evaluate_cpu.c:112-126 without the -w onehot term, which is
//              the exact init-time histogram)
// With two passes, A carries a hi and a lo BF16 half along K (K = 2*Kb), the
// one-hot operand repeats, and the FP32 accumulator sums both.
#include "common.h"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// coupling matrix -> Wt (bf16, P x lda), zero outside the real 21L x 21L block
__device__ __forceinline__ void convert_wt_block(const float* __restrict__ x2, int L, int Lp,
                                                 uint16_t* __restrict__ wt, long lda, int passes,
                                                 uint32_t tiles_r, uint32_t blk) {
  __shared__ float T[64][65];
  const int P = CCM_NA * Lp;
  const int tc = blk / tiles_r, tr = blk % tiles_r;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long ld2 = (long)CCM_NAP * L;
  {
    const int c = tc * 64 + lane;
    const int a = c / Lp, j = c % Lp;
    const bool cok = c < P && a < CCM_NA && j < L;
    for (int rr = wv; rr < 64; rr += 4) {
      const int r = tr * 64 + rr;
      const int b = r / Lp, k = r % Lp;
      float v = 0.f;
      if (cok && r < P && k < L) v = x2[(long)(b * L + k) * ld2 + a * L + j];
      T[rr][lane] = v;
    }
  }
  __syncthreads();
  const int r = tr * 64 + lane;
  if (r < P) {
    for (int cc = wv; cc < 64; cc += 4) {
      const int c = tc * 64 + cc;
      if (c >= P) break;
      const float v = T[lane][cc];
      const uint16_t hi = ccm_f2bf(v);
      wt[(long)c * lda + r] = hi;
      if (passes > 1) wt[(long)c * lda + P + r] = ccm_f2bf(v - ccm_bf2f(hi));
    }
  }
}

__global__ __launch_bounds__(256) void k_convert_wt(const float* __restrict__ x2, int L, int Lp,
                                                    uint16_t* __restrict__ wt, long lda, int passes,
                                                    uint32_t tiles_r) {
  convert_wt_block(x2, L, Lp, wt, lda, passes, tiles_r, blockIdx.x);
}

// every alignment of a group: (P/64)^2 blocks each
__global__ __launch_bounds__(256) void k_convert_wt_grouped(const ccm_bprob* __restrict__ probs,
                                                            const uint32_t* __restrict__ offs,
                                                            const int32_t* __restrict__ pidx, int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  convert_wt_block(B.x + B.nsp, B.L, B.Lp, B.wt, B.ldwt, B.passes_fwd, (uint32_t)((B.P + 63) / 64),
                   blockIdx.x - offs[gi]);
}

__device__ __forceinline__ uint32_t oh_pair(uint32_t lo_byte, uint32_t hi_byte, uint32_t b) {
  return (lo_byte == b ? 0x3F80u : 0u) | (hi_byte == b ? 0x3F800000u : 0u);
}

// 8 MSA bytes -> 8 bf16 one-hot values ([byte == b] ? 1 : 0), element i in
// the i-th bf16 slot of the fragment
__device__ __forceinline__ bf16x8_t onehot8(uint2 v, uint32_t b) {
  uint4 o;
  o.x = oh_pair(v.x & 0xffu, (v.x >> 8) & 0xffu, b);
  o.y = oh_pair((v.x >> 16) & 0xffu, v.x >> 24, b);
  o.z = oh_pair(v.y & 0xffu, (v.y >> 8) & 0xffu, b);
  o.w = oh_pair((v.y >> 16) & 0xffu, v.y >> 24, b);
  return __builtin_bit_cast(bf16x8_t, o);
}

// One workgroup computes a TM x TN tile of C with WM x WN waves; each wave
// owns RB x CB blocks of 32 x 32 and issues v_mfma_f32_32x32x16_bf16.  The
// dense operand moves through LDS in K steps of TK (two LDS stages, the next
// step prefetched through registers while the current one is used).  Within
// a step, lane half q of a wave takes the K range [q TK/2, (q+1) TK/2) and
// sub-step s the 8 indices s*8.. of it (the order of K inside an MFMA is
// free as long as both operands agree), so that the one-hot bytes a lane
// needs for a step are TK/2 contiguous bytes, fetched with 16-byte loads.
//
// MODE 0: forward.  cols = sequences n; K index (pass, bb, k) -> A column
//   pass*P + bb*Lp + k, B = [msap[n][k] == bb], Kb = P.  The K loop takes
//   blocks of TK columns k outermost and (pass, bb) inside, so that a lane's
//   MSA bytes are loaded once per block and serve all 21*passes steps.
// MODE 1: gradient. cols = coupling rows r = bb*Lp + k; K = sequences n of
//   the chunk (Kb of them): B = [msapt[k][n] == bb], bytes loaded per step.
// Requires Lp % TK == 0 (forward) and Kb % TK == 0.
template <int MODE, int WM, int WN, int RB, int CB, int TK>
__device__ __forceinline__ void gemm_onehot_tile(const uint16_t* __restrict__ A, long lda, int M, int Kb,
                                                 int passes, const uint8_t* __restrict__ bytes,
                                                 long ldbytes, int Lp, int Ncol, float* __restrict__ C,
                                                 long ldc, int accumulate, int mt, int nt) {
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = WM * RB * 32, TN = WN * CB * 32;
  constexpr int LDA = TK + 8;  // LDS row stride in bf16
  constexpr int KCH = TK / 8;  // 16-byte chunks per row and step
  constexpr int CPT = TM * KCH / NT;
  constexpr int HALF = TK / 2, NB16 = HALF / 16, SUBS = TK / 16;
  static_assert((TM * KCH) % NT == 0, "loader mapping");
  __shared__ __attribute__((aligned(16))) uint16_t As[2][TM * LDA];

  const int m0 = mt * TM, n0 = nt * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int l32 = lane & 31, q = lane >> 5;

  int colv[CB];
  uint32_t bval[CB];
  const uint8_t* bptr[CB];
#pragma unroll
  for (int cb = 0; cb < CB; cb++) {
    const int col = n0 + (wn * CB + cb) * 32 + l32;
    colv[cb] = col;
    if (MODE == 0) {
      bval[cb] = 0;
      // msap row n; columns past Ncol read row Ncol-1 and are never stored
      bptr[cb] = bytes + (long)(col < Ncol ? col : Ncol - 1) * ldbytes + q * HALF;
    } else {
      const int bb = col / Lp, kk = col - bb * Lp;
      bval[cb] = (uint32_t)bb;                           // >= 21 past the real rows: never matches
      bptr[cb] = bytes + (long)kk * ldbytes + q * HALF;  // msapt row k
    }
  }

  floatx16 acc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; i++)
#pragma unroll
    for (int j = 0; j < CB; j++) acc[i][j] = (floatx16)0.0f;

  // steps: MODE 0 (kc, pass, bb) with kc outermost; MODE 1 (pass, kc)
  const int kblocks = (MODE == 0 ? Lp : Kb) / TK;
  const int inner = MODE == 0 ? passes * CCM_NA : 1;
  const int nsteps = MODE == 0 ? kblocks * inner : passes * kblocks;
  auto acol = [&](int s) -> long {
    if (MODE == 0) {
      const int kc = s / inner, r = s - kc * inner, pass = r / CCM_NA, bb = r - pass * CCM_NA;
      return (long)pass * Kb + (long)bb * Lp + (long)kc * TK;
    } else {
      const int pass = s / kblocks, kc = s - pass * kblocks;
      return (long)pass * Kb + (long)kc * TK;
    }
  };

  // loader: chunk c = tid + i*NT -> row c / KCH, 16-byte column c % KCH; the
  // rows of one thread are NT/KCH apart
  constexpr int RSTEP = NT / KCH;
  const int arow0 = tid / KCH, acol8 = tid % KCH;
  const uint16_t* abase = A + (long)(m0 + arow0) * lda + acol8 * 8;
  const int adst0 = arow0 * LDA + acol8 * 8;
  auto load = [&](uint4(&r)[CPT], int step) {
    const long c = acol(step);
#pragma unroll
    for (int i = 0; i < CPT; i++)
      r[i] = (m0 + arow0 + i * RSTEP < M)
                 ? *reinterpret_cast<const uint4*>(abase + (long)i * RSTEP * lda + c)
                 : make_uint4(0, 0, 0, 0);
  };
  auto store = [&](const uint4(&r)[CPT], int buf) {
#pragma unroll
    for (int i = 0; i < CPT; i++) *reinterpret_cast<uint4*>(&As[buf][adst0 + i * RSTEP * LDA]) = r[i];
  };
  // a lane's one-hot bytes of K block kc
  auto loadb = [&](uint4(&bv)[CB][NB16], int kc) {
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
      for (int i = 0; i < NB16; i++)
        bv[cb][i] = *reinterpret_cast<const uint4*>(bptr[cb] + (long)kc * TK + 16 * i);
  };
  // with 16 accumulator blocks per wave the K sub-steps are not unrolled, so
  // that the scheduler does not hoist fragments of several sub-steps at once
  constexpr int SUB_UNROLL = RB * CB >= 16 ? 1 : SUBS;
  auto compute = [&](int buf, const uint4(&bv)[CB][NB16], uint32_t bsel) {
    const uint16_t* as = As[buf];
#pragma unroll SUB_UNROLL
    for (int sub = 0; sub < SUBS; sub++) {
      bf16x8_t bfrag[CB];
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        const uint4 w = bv[cb][sub >> 1];
        const uint2 v = (sub & 1) ? make_uint2(w.z, w.w) : make_uint2(w.x, w.y);
        bfrag[cb] = onehot8(v, MODE == 0 ? bsel : bval[cb]);
      }
#pragma unroll
      for (int rb = 0; rb < RB; rb++) {
        const uint4 av = *reinterpret_cast<const uint4*>(
            &as[((wm * RB + rb) * 32 + l32) * LDA + q * HALF + sub * 8]);
        const bf16x8_t afrag = __builtin_bit_cast(bf16x8_t, av);
#pragma unroll
        for (int cb = 0; cb < CB; cb++)
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afrag, bfrag[cb], acc[rb][cb], 0, 0, 0);
      }
    }
  };

  // MODE 1 prefetches the next step's bytes; MODE 0 reloads them once per
  // K block (every 21*passes steps), after the block's last step
  constexpr int NBX = MODE == 1 ? NB16 : 1;
  uint4 pre[CPT];
  uint4 bcur[CB][NB16], bnxt[CB][NBX];
  load(pre, 0);
  store(pre, 0);
  loadb(bcur, 0);
  __syncthreads();
  for (int s = 0; s < nsteps; s++) {
    const int cur = s & 1;
    if (s + 1 < nsteps) {
      load(pre, s + 1);
      if constexpr (MODE == 1) loadb(bnxt, (s + 1) % kblocks);
    }
    uint32_t bsel = 0;
    if (MODE == 0) {
      const int r = s % inner;
      bsel = (uint32_t)(r - (r / CCM_NA) * CCM_NA);
    }
    compute(cur, bcur, bsel);
    if (s + 1 < nsteps) {
      store(pre, cur ^ 1);
      if constexpr (MODE == 1) {
#pragma unroll
        for (int cb = 0; cb < CB; cb++)
#pragma unroll
          for (int i = 0; i < NB16; i++) bcur[cb][i] = bnxt[cb][i];
      } else {
        if ((s + 1) % inner == 0) loadb(bcur, (s + 1) / inner);
      }
    }
    __syncthreads();
  }

  // 32x32 accumulator layout: lane -> column l32; register i -> row
  // (i/4)*8 + 4*q + (i%4)
#pragma unroll
  for (int rb = 0; rb < RB; rb++) {
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      const int col = colv[cb];
      if (col >= Ncol) continue;
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int row = m0 + (wm * RB + rb) * 32 + (i >> 2) * 8 + 4 * q + (i & 3);
        if (row < M) {
          float* dst = C + (long)row * ldc + col;
          *dst = accumulate ? *dst + acc[rb][cb][i] : acc[rb][cb][i];
        }
      }
    }
  }
}

template <int MODE, int WM, int WN, int RB, int CB, int TK, int OCC>
__global__ __launch_bounds__(64 * WM * WN, OCC) void k_gemm_onehot(
    const uint16_t* __restrict__ A, long lda, int M, int Kb, int passes,
    const uint8_t* __restrict__ bytes, long ldbytes, int Lp, int Ncol, float* __restrict__ C,
    long ldc, int accumulate, uint32_t tiles_n, uint32_t tiles, uint32_t total) {
  const uint32_t qid = ccm_xcd_remap(blockIdx.x, total);
  if (qid >= tiles) return;
  gemm_onehot_tile<MODE, WM, WN, RB, CB, TK>(A, lda, M, Kb, passes, bytes, ldbytes, Lp, Ncol, C, ldc,
                                                 accumulate, qid / tiles_n, qid % tiles_n);
}

// The same tiles over a group of alignments, all with this tile variant:
// the XCD remap runs over the whole grid, so that each XCD takes a run of
// consecutive tiles (mostly of one alignment).
//   MODE 0: Zt = Wt . onehot(msap)      (M = P, K = P, cols = ncols)
//   MODE 1: Gt = Dt . onehot(msapt)^T   (M = P, K = ncols, cols = P)
template <int MODE, int WM, int WN, int RB, int CB, int TK, int OCC>
__global__ __launch_bounds__(64 * WM * WN, OCC) void k_gemm_onehot_grouped(
    const ccm_bprob* __restrict__ probs, const uint32_t* __restrict__ offs,
    const int32_t* __restrict__ pidx, int np, uint32_t total) {
  constexpr int TN = WN * CB * 32;
  const uint32_t qid = ccm_xcd_remap(blockIdx.x, total);
  if (qid >= offs[np]) return;
  const int gi = ccm_group_find(offs, np, qid);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  const uint32_t t = qid - offs[gi];
  if (MODE == 0) {
    const uint32_t tiles_n = (uint32_t)((B.ncols + TN - 1) / TN);
    gemm_onehot_tile<0, WM, WN, RB, CB, TK>(B.wt, B.ldwt, B.P, B.P, B.passes_fwd, B.msap, B.Lp, B.Lp,
                                                B.ncols, B.zt, B.ncols, 0, t / tiles_n, t % tiles_n);
  } else {
    const uint32_t tiles_n = (uint32_t)((B.P + TN - 1) / TN);
    gemm_onehot_tile<1, WM, WN, RB, CB, TK>(B.dt, B.lddt, B.P, B.ncols, B.passes, B.msapt, B.ldmt,
                                                B.Lp, B.P, B.gt, B.P, 0, t / tiles_n, t % tiles_n);
  }
}

static inline uint32_t round8(uint32_t v) { return (v + 7u) & ~7u; }

extern "C" {

int ccm_launch_convert_wt(const float* d_x, int L, int Lp, uint16_t* d_wt, long lda, int passes,
                          hipStream_t s) {
  const int P = CCM_NA * Lp;
  if (Lp % 32 || Lp < L || lda < (long)passes * P) return CCM_ERR_ARG;
  const uint32_t tiles = (uint32_t)((P + 63) / 64);
  hipLaunchKernelGGL(k_convert_wt, dim3(tiles * tiles), dim3(256), 0, s, d_x + ccm_nsp(L), L, Lp,
                     d_wt, lda, passes, tiles);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"

// Tile variants (the K step is not part of a variant: it follows from the
// contraction and the shape, see gemm_tk, so that every variant sums each
// output in the same order and gives bit-identical results):
//   {waves M x N, 32 x 32 blocks per wave M x N, workgroups per CU (one
//   with the K step 128, whose LDS stages take 139 KB for 256 rows)}
//   0: 2x4, 4x2, 1 -> 256 x 256, 8 waves
//   1: 2x2, 4x2, 2 -> 256 x 128, 4 waves
//   2: 2x2, 2x2, 2 -> 128 x 128, 4 waves
#define GM_VARIANTS(X) \
  X(0, 2, 4, 4, 2, 1)  \
  X(1, 2, 2, 4, 2, 2)  \
  X(2, 2, 2, 2, 2, 2)
#define GM_NVAR 3
static const int gm_tm[GM_NVAR] = {256, 256, 128};
static const int gm_tn[GM_NVAR] = {256, 128, 128};

// K step: the forward takes 128 where the padded column count allows (its
// one-hot bytes are reused over 21 steps, so longer steps only cut barrier
// and loop costs); the gradient loads its bytes every step and runs best at 64
static int gemm_tk(int mode, int Lp) { return mode == 0 && Lp % 128 == 0 ? 128 : 64; }

template <int MODE>
static int launch_gemm(int variant, const uint16_t* A, long lda, int M, int Kb, int passes,
                       const uint8_t* bytes, long ldbytes, int Lp, int Ncol, float* C, long ldc,
                       int accumulate, hipStream_t s) {
  if (variant < 0 || variant >= GM_NVAR) return CCM_ERR_ARG;
  const int tk = gemm_tk(MODE, Lp);
  if ((MODE == 0 && Lp % tk) || Kb % tk) return CCM_ERR_ARG;
  const uint32_t tiles_m = (M + gm_tm[variant] - 1) / gm_tm[variant];
  const uint32_t tiles_n = (Ncol + gm_tn[variant] - 1) / gm_tn[variant];
  const uint32_t tiles = tiles_m * tiles_n, total = round8(tiles);
#define GM_LAUNCH(TK, WM, WN, RB, CB, OCC)                                                             \
  hipLaunchKernelGGL((k_gemm_onehot<MODE, WM, WN, RB, CB, TK, OCC>), dim3(total), dim3(64 * WM * WN), 0, s, \
                     A, lda, M, Kb, passes, bytes, ldbytes, Lp, Ncol, C, ldc, accumulate, tiles_n, tiles, total)
#define GM_CASE(V, WM, WN, RB, CB, OCC)                 \
  case V:                                               \
    if (MODE == 0 && tk == 128)                         \
      GM_LAUNCH((MODE == 0 ? 128 : 64), WM, WN, RB, CB, 1); \
    else                                                \
      GM_LAUNCH(64, WM, WN, RB, CB, OCC);               \
    break;
  switch (variant) { GM_VARIANTS(GM_CASE) }
#undef GM_CASE
#undef GM_LAUNCH
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// grouped: every alignment of the group shares the K step tk
template <int MODE>
static int launch_gemm_grouped(int variant, int tk, const ccm_bprob* probs, const ccm_group& g,
                               hipStream_t s) {
  if (variant < 0 || variant >= GM_NVAR || g.np <= 0 || (tk != 64 && tk != 128)) return CCM_ERR_ARG;
  const uint32_t total = round8(g.total);
#define GG_LAUNCH(TK, WM, WN, RB, CB, OCC)                                                              \
  hipLaunchKernelGGL((k_gemm_onehot_grouped<MODE, WM, WN, RB, CB, TK, OCC>), dim3(total), dim3(64 * WM * WN), \
                     0, s, probs, g.offs, g.pidx, g.np, total)
#define GG_CASE(V, WM, WN, RB, CB, OCC)                 \
  case V:                                               \
    if (MODE == 0 && tk == 128)                         \
      GG_LAUNCH((MODE == 0 ? 128 : 64), WM, WN, RB, CB, 1); \
    else                                                \
      GG_LAUNCH(64, WM, WN, RB, CB, OCC);               \
    break;
  switch (variant) { GM_VARIANTS(GG_CASE) }
#undef GG_CASE
#undef GG_LAUNCH
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// workgroups of a launch for a tile variant
static uint32_t gemm_tiles(int variant, int M, int Ncol) {
  return (uint32_t)(((M + gm_tm[variant] - 1) / gm_tm[variant]) * ((Ncol + gm_tn[variant] - 1) / gm_tn[variant]));
}

// default variant: the largest tile that still gives most CUs a workgroup
// (forward: 256 x 256; gradient, measured faster with 4-wave workgroups:
// 256 x 128), then smaller tiles
static int gemm_default_variant(int mode, int M, int Ncol) {
  if (mode == 0 && gemm_tiles(0, M, Ncol) >= 192) return 0;
  if (gemm_tiles(1, M, Ncol) >= 192) return 1;
  return 2;
}

extern "C" {

int ccm_gemm_variants(void) { return GM_NVAR; }

// tiles of a variant over an M x Ncol output (0 for an unknown variant)
uint32_t ccm_gemm_tile_count(int variant, int M, int Ncol) {
  if (variant < 0 || variant >= GM_NVAR) return 0;
  return gemm_tiles(variant, M, Ncol);
}

// whether a variant fits a contraction: the forward (mode 0) needs Lp a
// multiple of 64, the gradient (mode 1) kcols a multiple of 64
int ccm_gemm_variant_ok(int variant, int mode, int Lp, int kcols) {
  if (variant < 0 || variant >= GM_NVAR) return 0;
  return mode == 0 ? Lp % gemm_tk(0, Lp) == 0 : kcols % gemm_tk(1, Lp) == 0;
}

// default variant of the forward (mode 0) or gradient (mode 1) contraction
int ccm_gemm_default_variant_for(int mode, int Lp, int ncols) {
  const int P = CCM_NA * Lp;
  return gemm_default_variant(mode, P, mode == 0 ? ncols : P);
}

int ccm_convert_wt_blocks(int Lp) {
  const int t = (CCM_NA * Lp + 63) / 64;
  return t * t;
}

// grouped forms: g.offs holds each alignment's block range; every alignment
// of a GEMM group must accept the variant (ccm_gemm_variant_ok)
int ccm_launch_convert_wt_grouped(const ccm_bprob* d_probs, ccm_group g, hipStream_t s) {
  if (g.np <= 0 || g.total == 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_convert_wt_grouped, dim3(g.total), dim3(256), 0, s, d_probs, g.offs, g.pidx, g.np);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// tk: the K step of every alignment of the group (ccm_gemm_tk)
int ccm_launch_gemm_fwd_grouped(const ccm_bprob* d_probs, ccm_group g, int variant, int tk, hipStream_t s) {
  return launch_gemm_grouped<0>(variant, tk, d_probs, g, s);
}

int ccm_launch_gemm_grad_grouped(const ccm_bprob* d_probs, ccm_group g, int variant, int tk, hipStream_t s) {
  return launch_gemm_grouped<1>(variant, tk, d_probs, g, s);
}

int ccm_gemm_tk(int mode, int Lp) { return gemm_tk(mode, Lp); }

int ccm_gemm_default_variant(int M, int Ncol) { return gemm_default_variant(0, M, Ncol); }

// Zt[c][n] (P x ldc floats, ldc >= ncols) = sum_r Wt[c][r] [x_nk = b] for the
// ncols (multiple of 128) sequences whose padded MSA rows start at d_msap;
// variant < 0 picks the default tile
int ccm_launch_gemm_fwd(const uint16_t* d_wt, long lda, int Lp, int passes, const uint8_t* d_msap,
                        int ncols, float* d_zt, long ldc, int variant, hipStream_t s) {
  const int P = CCM_NA * Lp;
  if (Lp % 64 || ncols % 128 || ncols <= 0 || lda < (long)passes * P || ldc < ncols || (lda % 8))
    return CCM_ERR_ARG;
  if (variant < 0) variant = gemm_default_variant(0, P, ncols);
  return launch_gemm<0>(variant, d_wt, lda, P, P, passes, d_msap, (long)Lp, Lp, ncols, d_zt, ldc, 0, s);
}

// Gt[c][r] (P x ldc floats, ldc >= P) (+)= sum_n Dt[c][n] [x_nk = b] over the
// kcols (multiple of 64) sequences of Dt; d_msapt points at the chunk's first
// column of the transposed MSA, whose rows are ldmt bytes apart
int ccm_launch_gemm_grad(const uint16_t* d_dt, long lda, int Lp, int passes, const uint8_t* d_msapt,
                         long ldmt, int kcols, float* d_gt, long ldc, int accumulate, int variant,
                         hipStream_t s) {
  const int P = CCM_NA * Lp;
  if (Lp % 64 || kcols % 64 || kcols <= 0 || lda < (long)passes * kcols || ldc < P || (lda % 8))
    return CCM_ERR_ARG;
  if (variant < 0) variant = gemm_default_variant(1, P, P);
  return launch_gemm<1>(variant, d_dt, lda, P, kcols, passes, d_msapt, ldmt, Lp, P, d_gt, ldc,
                        accumulate, s);
}

}  // extern "C"
