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
// dense operand moves through LDS in K steps of TK (two stages, the next one
// prefetched through registers while the current one is consumed); the
// one-hot operand is synthesised per lane from 8 MSA bytes per fragment.
//
// MODE 0: forward.  cols = sequences n;     B[kk][n] = [msap[n][kk%Lp] == kk/Lp], Kb = P
// MODE 1: gradient. cols = coupling rows r; B[kk][r] = [msapt[r%Lp][kk] == r/Lp], Kb = chunk
// Requires Lp % TK == 0 (forward) and Kb % TK == 0.
template <int MODE, int WM, int WN, int RB, int CB, int TK, int PF>
__device__ __forceinline__ void gemm_onehot_tile(const uint16_t* __restrict__ A, long lda, int M, int Kb,
                                                 int passes, const uint8_t* __restrict__ bytes,
                                                 long ldbytes, int Lp, int Ncol, float* __restrict__ C,
                                                 long ldc, int accumulate, int mt, int nt) {
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = WM * RB * 32, TN = WN * CB * 32;
  constexpr int LDA = TK + 8;  // LDS row stride in bf16
  constexpr int KCH = TK / 8;  // 16-byte chunks per row and step
  constexpr int CPT = TM * KCH / NT;
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
      bptr[cb] = bytes + (long)(col < Ncol ? col : Ncol - 1) * ldbytes;
    } else {
      const int bb = col / Lp, kk = col - bb * Lp;
      bval[cb] = (uint32_t)bb;                // >= 21 past the real rows: never matches
      bptr[cb] = bytes + (long)kk * ldbytes;  // msapt row k
    }
  }

  floatx16 acc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; i++)
#pragma unroll
    for (int j = 0; j < CB; j++) acc[i][j] = (floatx16)0.0f;

  const int nsteps = passes * Kb / TK;
  // loader: chunk c = tid + i*NT -> row c / KCH, 16-byte column c % KCH; the
  // rows of one thread are NT/KCH apart
  constexpr int RSTEP = NT / KCH;
  const int arow0 = tid / KCH, acol8 = tid % KCH;
  const uint16_t* abase = A + (long)(m0 + arow0) * lda + acol8 * 8;
  const int adst0 = arow0 * LDA + acol8 * 8;
  // uniform position of the current K step inside one pass: for the forward
  // the coupling row block bb and the column offset kin; for the gradient
  // the sequence offset kin
  int bb = 0, kin = 0;
  auto load = [&](uint4(&r)[CPT], int step) {
#pragma unroll
    for (int i = 0; i < CPT; i++)
      r[i] = (m0 + arow0 + i * RSTEP < M)
                 ? *reinterpret_cast<const uint4*>(abase + (long)i * RSTEP * lda + (long)step * TK)
                 : make_uint4(0, 0, 0, 0);
  };
  auto store = [&](const uint4(&r)[CPT], int buf) {
#pragma unroll
    for (int i = 0; i < CPT; i++) *reinterpret_cast<uint4*>(&As[buf][adst0 + i * RSTEP * LDA]) = r[i];
  };
  auto advance = [&]() {
    kin += TK;
    if (MODE == 0) {
      if (kin == Lp) {
        kin = 0;
        if (++bb == CCM_NA) bb = 0;
      }
    } else {
      if (kin == Kb) kin = 0;
    }
  };
  // with 16 accumulator blocks per wave the K sub-steps are not unrolled, so
  // that the scheduler does not hoist fragments of several sub-steps at once
  constexpr int SUB_UNROLL = RB * CB >= 16 ? 1 : TK / 16;
  auto compute = [&](int buf) {
    const uint16_t* as = As[buf];
#pragma unroll SUB_UNROLL
    for (int sub = 0; sub < TK / 16; sub++) {
      const int koff = kin + sub * 16 + 8 * q;  // first of this lane's 8 K indices
      bf16x8_t bfrag[CB];
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        const uint2 v = *reinterpret_cast<const uint2*>(bptr[cb] + koff);
        bfrag[cb] = onehot8(v, MODE == 0 ? (uint32_t)bb : bval[cb]);
      }
#pragma unroll
      for (int rb = 0; rb < RB; rb++) {
        const uint4 av = *reinterpret_cast<const uint4*>(
            &as[((wm * RB + rb) * 32 + l32) * LDA + sub * 16 + 8 * q]);
        const bf16x8_t afrag = __builtin_bit_cast(bf16x8_t, av);
#pragma unroll
        for (int cb = 0; cb < CB; cb++)
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afrag, bfrag[cb], acc[rb][cb], 0, 0, 0);
      }
    }
  };

  if (PF == 1) {
    // one register stage: the next step's tile is loaded while this one is used
    uint4 pre[CPT];
    load(pre, 0);
    store(pre, 0);
    __syncthreads();
    for (int s = 0; s < nsteps; s++) {
      const int cur = s & 1;
      if (s + 1 < nsteps) load(pre, s + 1);
      compute(cur);
      advance();
      if (s + 1 < nsteps) store(pre, cur ^ 1);
      __syncthreads();
    }
  } else {
    // two register stages: a tile is loaded two steps before it is used
    uint4 ra[CPT], rb2[CPT];
    load(ra, 0);
    store(ra, 0);
    if (nsteps > 1) load(rb2, 1);
    __syncthreads();
    for (int s = 0; s < nsteps; s += 2) {
      if (s + 2 < nsteps) load(ra, s + 2);
      compute(0);
      advance();
      if (s + 1 < nsteps) store(rb2, 1);
      __syncthreads();
      if (s + 1 >= nsteps) break;
      if (s + 3 < nsteps) load(rb2, s + 3);
      compute(1);
      advance();
      if (s + 2 < nsteps) store(ra, 0);
      __syncthreads();
    }
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

template <int MODE, int WM, int WN, int RB, int CB, int TK, int PF, int OCC>
__global__ __launch_bounds__(64 * WM * WN, OCC) void k_gemm_onehot(
    const uint16_t* __restrict__ A, long lda, int M, int Kb, int passes,
    const uint8_t* __restrict__ bytes, long ldbytes, int Lp, int Ncol, float* __restrict__ C,
    long ldc, int accumulate, uint32_t tiles_n, uint32_t tiles, uint32_t total) {
  const uint32_t qid = ccm_xcd_remap(blockIdx.x, total);
  if (qid >= tiles) return;
  gemm_onehot_tile<MODE, WM, WN, RB, CB, TK, PF>(A, lda, M, Kb, passes, bytes, ldbytes, Lp, Ncol, C, ldc,
                                                 accumulate, qid / tiles_n, qid % tiles_n);
}

// The same tiles over a group of alignments, all with this tile variant:
// the XCD remap runs over the whole grid, so that each XCD takes a run of
// consecutive tiles (mostly of one alignment).
//   MODE 0: Zt = Wt . onehot(msap)      (M = P, K = P, cols = ncols)
//   MODE 1: Gt = Dt . onehot(msapt)^T   (M = P, K = ncols, cols = P)
template <int MODE, int WM, int WN, int RB, int CB, int TK, int PF, int OCC>
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
    gemm_onehot_tile<0, WM, WN, RB, CB, TK, PF>(B.wt, B.ldwt, B.P, B.P, B.passes_fwd, B.msap, B.Lp, B.Lp,
                                                B.ncols, B.zt, B.ncols, 0, t / tiles_n, t % tiles_n);
  } else {
    const uint32_t tiles_n = (uint32_t)((B.P + TN - 1) / TN);
    gemm_onehot_tile<1, WM, WN, RB, CB, TK, PF>(B.dt, B.lddt, B.P, B.ncols, B.passes, B.msapt, B.ldmt,
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

// Tile variants: {waves M x N, row blocks x col blocks per wave, K step,
// prefetch stages, workgroups per CU}
//   0: 2x4, 4x2, 64, 1, 1 -> 256 x 256, 8 waves, 2 waves/SIMD
//   1: 2x2, 4x2, 64, 1, 2 -> 256 x 128
//   2: 2x2, 2x2, 64, 2, 2 -> 128 x 128
//   3: 2x2, 4x4, 64, 2, 1 -> 256 x 256, 4 waves of 128 x 128, 1 wave/SIMD
//   4: 2x4, 4x2, 128, 1, 1 -> 256 x 256, K step 128
//   5: 2x2, 4x4, 64, 1, 1 -> as 3 with one prefetch stage
#define GM_NVAR 6
static const int gm_tm[GM_NVAR] = {256, 256, 128, 256, 256, 256};
static const int gm_tn[GM_NVAR] = {256, 128, 128, 256, 256, 256};
static const int gm_tk[GM_NVAR] = {64, 64, 64, 64, 128, 64};

template <int MODE>
static int launch_gemm(int variant, const uint16_t* A, long lda, int M, int Kb, int passes,
                       const uint8_t* bytes, long ldbytes, int Lp, int Ncol, float* C, long ldc,
                       int accumulate, hipStream_t s) {
  if (variant < 0 || variant >= GM_NVAR) return CCM_ERR_ARG;
  const int tk = gm_tk[variant];
  if ((MODE == 0 && Lp % tk) || Kb % tk) return CCM_ERR_ARG;
  const uint32_t tiles_m = (M + gm_tm[variant] - 1) / gm_tm[variant];
  const uint32_t tiles_n = (Ncol + gm_tn[variant] - 1) / gm_tn[variant];
  const uint32_t tiles = tiles_m * tiles_n, total = round8(tiles);
#define GM_ARGS A, lda, M, Kb, passes, bytes, ldbytes, Lp, Ncol, C, ldc, accumulate, tiles_n, tiles, total
  switch (variant) {
    case 0:
      hipLaunchKernelGGL((k_gemm_onehot<MODE, 2, 4, 4, 2, 64, 1, 1>), dim3(total), dim3(512), 0, s, GM_ARGS);
      break;
    case 1:
      hipLaunchKernelGGL((k_gemm_onehot<MODE, 2, 2, 4, 2, 64, 1, 2>), dim3(total), dim3(256), 0, s, GM_ARGS);
      break;
    case 2:
      hipLaunchKernelGGL((k_gemm_onehot<MODE, 2, 2, 2, 2, 64, 2, 2>), dim3(total), dim3(256), 0, s, GM_ARGS);
      break;
    case 3:
      hipLaunchKernelGGL((k_gemm_onehot<MODE, 2, 2, 4, 4, 64, 2, 1>), dim3(total), dim3(256), 0, s, GM_ARGS);
      break;
    case 4:
      hipLaunchKernelGGL((k_gemm_onehot<MODE, 2, 4, 4, 2, 128, 1, 1>), dim3(total), dim3(512), 0, s, GM_ARGS);
      break;
    default:
      hipLaunchKernelGGL((k_gemm_onehot<MODE, 2, 2, 4, 4, 64, 1, 1>), dim3(total), dim3(256), 0, s, GM_ARGS);
      break;
  }
#undef GM_ARGS
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

template <int MODE>
static int launch_gemm_grouped(int variant, const ccm_bprob* probs, const ccm_group& g, hipStream_t s) {
  if (variant < 0 || variant >= GM_NVAR || g.np <= 0) return CCM_ERR_ARG;
  const uint32_t total = round8(g.total);
#define GG_ARGS probs, g.offs, g.pidx, g.np, total
  switch (variant) {
    case 0:
      hipLaunchKernelGGL((k_gemm_onehot_grouped<MODE, 2, 4, 4, 2, 64, 1, 1>), dim3(total), dim3(512), 0, s, GG_ARGS);
      break;
    case 1:
      hipLaunchKernelGGL((k_gemm_onehot_grouped<MODE, 2, 2, 4, 2, 64, 1, 2>), dim3(total), dim3(256), 0, s, GG_ARGS);
      break;
    case 2:
      hipLaunchKernelGGL((k_gemm_onehot_grouped<MODE, 2, 2, 2, 2, 64, 2, 2>), dim3(total), dim3(256), 0, s, GG_ARGS);
      break;
    case 3:
      hipLaunchKernelGGL((k_gemm_onehot_grouped<MODE, 2, 2, 4, 4, 64, 2, 1>), dim3(total), dim3(256), 0, s, GG_ARGS);
      break;
    case 4:
      hipLaunchKernelGGL((k_gemm_onehot_grouped<MODE, 2, 4, 4, 2, 128, 1, 1>), dim3(total), dim3(512), 0, s, GG_ARGS);
      break;
    default:
      hipLaunchKernelGGL((k_gemm_onehot_grouped<MODE, 2, 2, 4, 4, 64, 1, 1>), dim3(total), dim3(256), 0, s, GG_ARGS);
      break;
  }
#undef GG_ARGS
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// workgroups of a launch for a tile variant
static uint32_t gemm_tiles(int variant, int M, int Ncol) {
  return (uint32_t)(((M + gm_tm[variant] - 1) / gm_tm[variant]) * ((Ncol + gm_tn[variant] - 1) / gm_tn[variant]));
}

// default variant: the 256 x 256 tile with the longest K step the shape
// allows while it still gives most CUs a workgroup, then smaller tiles
static int gemm_default_variant(int M, int Ncol, int kstep_ok128) {
  if (kstep_ok128 && gemm_tiles(4, M, Ncol) >= 192) return 4;
  if (gemm_tiles(0, M, Ncol) >= 192) return 0;
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

// whether a variant's K step divides the contraction: the forward needs
// Lp % TK == 0 (mode 0), the gradient kcols % TK == 0 (mode 1)
int ccm_gemm_variant_ok(int variant, int mode, int Lp, int kcols) {
  if (variant < 0 || variant >= GM_NVAR) return 0;
  return mode == 0 ? Lp % gm_tk[variant] == 0 : kcols % gm_tk[variant] == 0;
}

// default variant of the forward (mode 0) or gradient (mode 1) contraction
int ccm_gemm_default_variant_for(int mode, int Lp, int ncols) {
  const int P = CCM_NA * Lp;
  return mode == 0 ? gemm_default_variant(P, ncols, Lp % 128 == 0)
                   : gemm_default_variant(P, P, ncols % 128 == 0);
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

int ccm_launch_gemm_fwd_grouped(const ccm_bprob* d_probs, ccm_group g, int variant, hipStream_t s) {
  return launch_gemm_grouped<0>(variant, d_probs, g, s);
}

int ccm_launch_gemm_grad_grouped(const ccm_bprob* d_probs, ccm_group g, int variant, hipStream_t s) {
  return launch_gemm_grouped<1>(variant, d_probs, g, s);
}

int ccm_gemm_default_variant(int M, int Ncol) { return gemm_default_variant(M, Ncol, 1); }

// Zt[c][n] (P x ldc floats, ldc >= ncols) = sum_r Wt[c][r] [x_nk = b] for the
// ncols (multiple of 128) sequences whose padded MSA rows start at d_msap;
// variant < 0 picks the default tile
int ccm_launch_gemm_fwd(const uint16_t* d_wt, long lda, int Lp, int passes, const uint8_t* d_msap,
                        int ncols, float* d_zt, long ldc, int variant, hipStream_t s) {
  const int P = CCM_NA * Lp;
  if (Lp % 64 || ncols % 128 || ncols <= 0 || lda < (long)passes * P || ldc < ncols || (lda % 8))
    return CCM_ERR_ARG;
  if (variant < 0) variant = gemm_default_variant(P, ncols, Lp % 128 == 0);
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
  if (variant < 0) variant = gemm_default_variant(P, P, kcols % 128 == 0);
  return launch_gemm<1>(variant, d_dt, lda, P, kcols, passes, d_msapt, ldmt, Lp, P, d_gt, ldc,
                        accumulate, s);
}

}  // extern "C"
