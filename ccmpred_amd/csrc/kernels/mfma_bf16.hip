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
// The forward also exists with the column softmax as its epilogue (MODE 2):
// its tiles gather whole softmax groups (all 21 residues of 12 or 6
// alignment columns), and the logits stay in LDS; Dt and the objective and
// node-gradient partials are its outputs (evaluate_cpu.c:76-109).
#include "dtl.h"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// coupling matrix -> Wt (bf16, P x lda).  A block takes 64 rows r = b*Lp + k
// (one residue block b: Lp is a multiple of 64) and 64 columns C = a*L + j
// of the x layout starting at a multiple of 64, so that every row segment it
// reads is 256 contiguous, aligned bytes; column C goes to Wt row
// c = a*Lp + j.  Columns k >= L of Wt are written as zeros; rows j >= L
// (a*Lp + j) are never written: the caller zeroes Wt once.
__device__ __forceinline__ void convert_wt_block(const float* __restrict__ x2, int L, int Lp,
                                                 uint16_t* __restrict__ wt, long lda, int passes,
                                                 uint32_t tiles_r, uint32_t blk) {
  __shared__ float T[64][65];
  const int P = CCM_NA * Lp;  // a multiple of 64: the row tiles cover it exactly
  const int C21 = CCM_NA * L;
  const int tc = blk / tiles_r, tr = blk % tiles_r;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long ld2 = (long)CCM_NAP * L;
  {
    // T[rr][cc] = W at coupling row r = tr*64 + rr, x column C = tc*64 + cc;
    // the 16 loads of a thread are issued before they are stored
    const int b = (tr * 64) / Lp, k0 = tr * 64 - b * Lp;
    const int C = tc * 64 + lane;
    const bool cok = C < C21;
    const float* src = x2 + (long)(b * L + k0) * ld2 + C;
    float v[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int rr = wv + 4 * i;
      v[i] = (cok && k0 + rr < L) ? src[(long)rr * ld2] : 0.f;
    }
#pragma unroll
    for (int i = 0; i < 16; i++) T[wv + 4 * i][lane] = v[i];
  }
  __syncthreads();
  // each thread converts 16 consecutive r of one output row c and stores
  // them as two 16-byte vectors (and the residual half likewise)
  const int cl = threadIdx.x >> 2, qr = threadIdx.x & 3;
  const int C = tc * 64 + cl;
  if (C >= C21) return;
  const int a = C / L;
  const int c = a * Lp + (C - a * L), r0 = tr * 64 + qr * 16;
  uint32_t hi[8], lo[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const float v0 = T[qr * 16 + 2 * i][cl], v1 = T[qr * 16 + 2 * i + 1][cl];
    hi[i] = ccm_f2bf2(v0, v1);
    lo[i] = ccm_f2bf2(v0 - ccm_bf2f((uint16_t)hi[i]), v1 - ccm_bf2f((uint16_t)(hi[i] >> 16)));
  }
  uint4* dst = reinterpret_cast<uint4*>(wt + (long)c * lda + r0);
  dst[0] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
  dst[1] = make_uint4(hi[4], hi[5], hi[6], hi[7]);
  if (passes > 1) {
    uint4* dlo = reinterpret_cast<uint4*>(wt + (long)c * lda + P + r0);
    dlo[0] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
    dlo[1] = make_uint4(lo[4], lo[5], lo[6], lo[7]);
  }
}

// The same conversion over 64 rows and CW_TC = 128 columns with 16-byte
// loads (each row segment read is 512 contiguous bytes): for x addresses
// that are 16-byte aligned.  A float4 that straddles column 21 L reads a few
// padding columns of the same row, which are not written.
#define CW_TC 128
__device__ __forceinline__ void convert_wt_block128(const float* __restrict__ x2, int L, int Lp,
                                                    uint16_t* __restrict__ wt, long lda, int passes,
                                                    uint32_t tiles_r, uint32_t blk) {
  __shared__ float T[64][CW_TC + 1];
  const int P = CCM_NA * Lp;
  const int C21 = CCM_NA * L;
  const int tc = blk / tiles_r, tr = blk % tiles_r;
  const long ld2 = (long)CCM_NAP * L;
  {
    const int b = (tr * 64) / Lp, k0 = tr * 64 - b * Lp;
    const float* src = x2 + (long)(b * L + k0) * ld2 + tc * CW_TC;
    float4 v[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int f = threadIdx.x + 256 * i, rr = f >> 5, c4 = f & 31;
      v[i] = (tc * CW_TC + c4 * 4 < C21 && k0 + rr < L)
                 ? *reinterpret_cast<const float4*>(src + (long)rr * ld2 + c4 * 4)
                 : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int f = threadIdx.x + 256 * i, rr = f >> 5, c4 = f & 31;
      T[rr][c4 * 4 + 0] = v[i].x;
      T[rr][c4 * 4 + 1] = v[i].y;
      T[rr][c4 * 4 + 2] = v[i].z;
      T[rr][c4 * 4 + 3] = v[i].w;
    }
  }
  __syncthreads();
  // each thread converts 32 consecutive r of one output row c (four 16-byte
  // stores, and the residual half likewise)
  const int cl = threadIdx.x >> 1, half = threadIdx.x & 1;
  const int C = tc * CW_TC + cl;
  if (C >= C21) return;
  const int a = C / L;
  const int c = a * Lp + (C - a * L), r0 = tr * 64 + half * 32;
  uint32_t hi[16], lo[16];
#pragma unroll
  for (int i = 0; i < 16; i++) {
    const float v0 = T[half * 32 + 2 * i][cl], v1 = T[half * 32 + 2 * i + 1][cl];
    hi[i] = ccm_f2bf2(v0, v1);
    lo[i] = ccm_f2bf2(v0 - ccm_bf2f((uint16_t)hi[i]), v1 - ccm_bf2f((uint16_t)(hi[i] >> 16)));
  }
  uint4* dst = reinterpret_cast<uint4*>(wt + (long)c * lda + r0);
#pragma unroll
  for (int q = 0; q < 4; q++) dst[q] = make_uint4(hi[4 * q], hi[4 * q + 1], hi[4 * q + 2], hi[4 * q + 3]);
  if (passes > 1) {
    uint4* dlo = reinterpret_cast<uint4*>(wt + (long)c * lda + P + r0);
#pragma unroll
    for (int q = 0; q < 4; q++) dlo[q] = make_uint4(lo[4 * q], lo[4 * q + 1], lo[4 * q + 2], lo[4 * q + 3]);
  }
}

__global__ __launch_bounds__(256) void k_convert_wt128(const float* __restrict__ x2, int L, int Lp,
                                                       uint16_t* __restrict__ wt, long lda, int passes,
                                                       uint32_t tiles_r) {
  convert_wt_block128(x2, L, Lp, wt, lda, passes, tiles_r, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_convert_wt(const float* __restrict__ x2, int L, int Lp,
                                                    uint16_t* __restrict__ wt, long lda, int passes,
                                                    uint32_t tiles_r) {
  convert_wt_block(x2, L, Lp, wt, lda, passes, tiles_r, blockIdx.x);
}

// every alignment of a group: ccm_convert_wt_blocks(L, Lp) blocks each (x
// buffers of a batch are allocated, so 16-byte aligned)
__global__ __launch_bounds__(256) void k_convert_wt_grouped(const ccm_bprob* __restrict__ probs,
                                                            const uint32_t* __restrict__ offs,
                                                            const int32_t* __restrict__ pidx, int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  convert_wt_block128(B.x + B.nsp, B.L, B.Lp, B.wt, B.ldwt, B.passes_fwd, (uint32_t)((B.P + 63) / 64),
                      blockIdx.x - offs[gi]);
}

// One-hot fragments from MSA bytes with packed 16-bit integer arithmetic.
// A "spread" dword holds two bytes in the low bytes of its 16-bit halves;
// after XOR with the residue (in both halves) a half is zero exactly where
// the byte matched, and min(t, 1) * 0xC080 + 0x3F80 (mod 2^16) maps 0 to
// bf16 1.0 (0x3F80) and 1 to 0.  The min is written as the instruction
// because the compiler otherwise turns the sequence back into per-half
// compares and selects; the multiply-add becomes one v_pk_mad_u16.
typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t oh_halves(uint32_t t) {
  uint32_t m;
  asm("v_pk_min_u16 %0, %1, %2" : "=v"(m) : "v"(t), "v"(0x00010001u));
  const u16x2_t r = __builtin_bit_cast(u16x2_t, m) * (u16x2_t)(0xC080u) + (u16x2_t)(0x3F80u);
  return __builtin_bit_cast(uint32_t, r);
}

// bytes 0,1 (sel 0) or 2,3 (sel 1) of v into the low bytes of the halves
__device__ __forceinline__ uint32_t spread_lo(uint32_t v) { return __builtin_amdgcn_perm(0u, v, 0x0c010c00u); }
__device__ __forceinline__ uint32_t spread_hi(uint32_t v) { return __builtin_amdgcn_perm(0u, v, 0x0c030c02u); }

// 8 MSA bytes -> 8 bf16 one-hot values ([byte == b] ? 1 : 0), element i in
// the i-th bf16 slot of the fragment
__device__ __forceinline__ bf16x8_t onehot8(uint2 v, uint32_t b) {
  const uint32_t p = b * 0x01010101u;
  const uint32_t t0 = v.x ^ p, t1 = v.y ^ p;
  uint4 o;
  o.x = oh_halves(spread_lo(t0));
  o.y = oh_halves(spread_hi(t0));
  o.z = oh_halves(spread_lo(t1));
  o.w = oh_halves(spread_hi(t1));
  return __builtin_bit_cast(bf16x8_t, o);
}

// the same from 8 bytes already spread into four dwords; p2 = b * 0x00010001
__device__ __forceinline__ bf16x8_t onehot8_spread(uint32_t s0, uint32_t s1, uint32_t s2, uint32_t s3,
                                                   uint32_t p2) {
  uint4 o;
  o.x = oh_halves(s0 ^ p2);
  o.y = oh_halves(s1 ^ p2);
  o.z = oh_halves(s2 ^ p2);
  o.w = oh_halves(s3 ^ p2);
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

// ---- the same contraction with direct-to-LDS staging ----------------------
//
// The dense operand (and, for the gradient, the one-hot bytes) go from global
// memory straight into LDS with global_load_lds_dwordx4: each wave
// instruction fills 1 KB of LDS, lane i taking the 16 bytes at offset 16 i.
// S stages are in flight, so a step's loads are issued S-1 steps before they
// are used and no registers hold them.  LDS rows are unpadded; bank conflicts
// are avoided by an XOR swizzle of the 16-byte granules of a row (the loading
// lane fetches the granule that belongs at its position, the reading lane
// looks it up through the same XOR).  Waits are explicit vmcnt counts and the
// step barrier carries no memory fence, so later stages stay in flight across
// it.  Same K order as gemm_onehot_tile: the results are bit-identical.
// SB > 0 sizes the stages (at least the tile's own), so that tiles of two
// shapes in one kernel share the same three LDS arrays.

template <int MODE, int WM, int WN, int RB, int CB, int TK, int S, int SB = 0>
__device__ __forceinline__ void gemm_onehot_tile_dtl(const uint16_t* __restrict__ A, long lda, int M,
                                                     int Kb, int passes,
                                                     const uint8_t* __restrict__ bytes, long ldbytes,
                                                     int Lp, int Ncol, float* __restrict__ C, long ldc,
                                                     int accumulate, int mt, int nt,
                                                     const ccm_fwd_epi& epi) {
  constexpr bool FWD = MODE != 1;  // forward contraction (MODE 0: logits out; 2: softmax fused)
  constexpr int NW = WM * WN;
  constexpr int TM = WM * RB * 32, TN = WN * CB * 32;
  constexpr int G = TK / 8;                   // 16-byte granules per dense row
  constexpr int A_INS = TM * G / (64 * NW);   // DTL instructions per wave and stage
  constexpr int GB = TK / 16;                 // 16-byte granules per byte column
  constexpr int B_INS = MODE == 1 ? TN * GB / (64 * NW) : 0;
  constexpr int TNB = MODE == 1 ? TN * TK : 0;  // bytes of a byte stage
  constexpr int HALF = TK / 2, NB16 = HALF / 16, SUBS = TK / 16;
  constexpr int PER_STEP = A_INS + B_INS;     // vector-memory instructions per step
  static_assert((TM * G) % (64 * NW) == 0, "dense loader mapping");
  static_assert(FWD || (TN * GB) % (64 * NW) == 0, "byte loader mapping");
  static_assert(FWD || TK == 64, "gradient K step");
  static_assert(S == 3, "stages");
  static_assert((S - 2) * PER_STEP < 64, "vmcnt range");

  const int m0 = mt * TM, n0 = nt * TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int l32 = lane & 31, q = lane >> 5;
  // the LDS operand of a DTL load must be wave-uniform
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);

  // The direct-to-LDS loads use the buffer form (buffer_load ... lds): the
  // compiler counts the FLAT form (global_load_lds) against the LDS counter
  // as well and then waits for every LDS read in flight (lgkmcnt(0)) where a
  // count would do.  The dense operand's resource starts at the tile's first
  // row, so the per-lane byte offsets (row, granule) fit 32 bits and stay
  // fixed; a step's column offset goes in the scalar offset.
  // MODE 2 gathers its rows: tile mt holds alignment columns j = mt*JG ..
  // mt*JG + JG-1 for all 21 residues, local row a*JG + jj <- row a*Lp + j of
  // the dense operand (rows past 21*JG, and columns j >= L, repeat a row
  // whose results are dropped), so that every softmax of the tile is local.
  constexpr int JG = TM / CCM_NA;
  const ccm_rsrc_t arsrc = ccm_make_rsrc(MODE == 2 ? A : A + (long)m0 * lda);
  uint32_t aoff[A_INS];
#pragma unroll
  for (int i = 0; i < A_INS; i++) {
    const int p = (wave * A_INS + i) * 64 + lane;
    const int r = p / G, g = swz_a<G>(r, p % G);
    long row;
    if constexpr (MODE == 2) {
      const int a = r / JG < CCM_NA ? r / JG : 0;
      const int j = mt * JG + (r - (r / JG) * JG);
      row = (long)a * Lp + (j < epi.L ? j : epi.L - 1);
    } else {
      row = m0 + r < M ? r : M - 1 - m0;  // rows past M are read but never stored
    }
    aoff[i] = (uint32_t)((row * lda + g * 8) * 2);
  }
  const ccm_rsrc_t brsrc = ccm_make_rsrc(bytes);
  uint32_t boff[B_INS > 0 ? B_INS : 1];
#pragma unroll
  for (int i = 0; i < B_INS; i++) {
    const int p = (wave * B_INS + i) * 64 + lane;
    const int c = p / GB, g = swz_b<GB>(c, p % GB);
    const int kk = (n0 + c) % Lp;  // msapt row of the column; bb >= 21 past the real rows
    boff[i] = (uint32_t)((long)kk * ldbytes + g * 16);
  }

  // per lane and column block: the one-hot selector
  int colv[CB];
  uint32_t bval[CB];
  const uint8_t* bptr[CB];
#pragma unroll
  for (int cb = 0; cb < CB; cb++) {
    const int col = n0 + (wn * CB + cb) * 32 + l32;
    colv[cb] = col;
    if (FWD) {
      bval[cb] = 0;
      bptr[cb] = bytes + (long)(col < Ncol ? col : Ncol - 1) * ldbytes + q * HALF;
    } else {
      bval[cb] = (uint32_t)(col / Lp);
      bptr[cb] = nullptr;
    }
  }

  floatx16 acc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; i++)
#pragma unroll
    for (int j = 0; j < CB; j++) acc[i][j] = (floatx16)0.0f;

  const int kblocks = (FWD ? Lp : Kb) / TK;
  const int inner = FWD ? passes * CCM_NA : 1;
  const int nsteps = FWD ? kblocks * inner : passes * kblocks;
  auto acol = [&](int st) -> long {
    if (FWD) {
      const int kc = st / inner, r = st - kc * inner, pass = r / CCM_NA, bb = r - pass * CCM_NA;
      return (long)pass * Kb + (long)bb * Lp + (long)kc * TK;
    } else {
      const int pass = st / kblocks, kc = st - pass * kblocks;
      return (long)pass * Kb + (long)kc * TK;
    }
  };
  constexpr int A_BYTES = TM * TK * 2;
  // DTL loads of step st into the stage at sp (dense tile, then bytes)
  auto issue = [&](uint8_t* __restrict__ sp, int st) {
    uint16_t* as = reinterpret_cast<uint16_t*>(sp);
    uint8_t* bs = sp + A_BYTES;
    const uint32_t c = (uint32_t)(acol(st) * 2);
#pragma unroll
    for (int i = 0; i < A_INS; i++) ccm_buffer_load_lds16(arsrc, as + (wave_u * A_INS + i) * 512, aoff[i], c);
    if constexpr (MODE == 1) {
      const uint32_t nb = (uint32_t)((st % kblocks) * TK);
#pragma unroll
      for (int i = 0; i < B_INS; i++) ccm_buffer_load_lds16(brsrc, bs + (wave_u * B_INS + i) * 1024, boff[i], nb);
    }
  };
  auto loadb = [&](uint4(&bv)[CB][NB16], int kc) {
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
      for (int i = 0; i < NB16; i++)
        bv[cb][i] = *reinterpret_cast<const uint4*>(bptr[cb] + (long)kc * TK + 16 * i);
  };
  uint4 bcur[CB][FWD ? NB16 : 1];
  // forward: the block's bytes spread for onehot8_spread, once per K block
  // (at its first step, so that the byte loads overlap the previous steps)
  uint32_t bsp[CB][FWD ? NB16 * 8 : 1];
  auto spread_bytes = [&]() {
#pragma unroll
    for (int cb = 0; cb < CB; cb++)
#pragma unroll
      for (int i = 0; i < (FWD ? NB16 : 0); i++) {
        const uint4 w = bcur[cb][i];
        bsp[cb][i * 8 + 0] = spread_lo(w.x);
        bsp[cb][i * 8 + 1] = spread_hi(w.x);
        bsp[cb][i * 8 + 2] = spread_lo(w.y);
        bsp[cb][i * 8 + 3] = spread_hi(w.y);
        bsp[cb][i * 8 + 4] = spread_lo(w.z);
        bsp[cb][i * 8 + 5] = spread_hi(w.z);
        bsp[cb][i * 8 + 6] = spread_lo(w.w);
        bsp[cb][i * 8 + 7] = spread_hi(w.w);
      }
  };
  // LDS fragments of the dense operand (and MODE 1's bytes) in two register
  // buffers: sub-step s uses af[s & 1] and bw[(s >> 1) & 1]
  uint4 af[2][RB];
  uint4 bw[2][CB];
  auto lda_frags = [&](const uint8_t* __restrict__ sp, uint4(&a)[RB], int sub) {
    const uint16_t* as = reinterpret_cast<const uint16_t*>(sp);
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
      const int r = (wm * RB + rb) * 32 + l32;
      const int g = q * SUBS + sub;
      a[rb] = *reinterpret_cast<const uint4*>(&as[r * TK + swz_a<G>(r, g) * 8]);
    }
  };
  // MODE 1: the 16 staged bytes that serve sub-steps 2*pair and 2*pair + 1
  // (the forward reads the spread bytes held in registers)
  auto ld_bytes = [&](const uint8_t* __restrict__ sp, uint4(&w)[CB], int pair) {
    if constexpr (MODE == 1) {
      const uint8_t* bs = sp + A_BYTES;
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        const int c = (wn * CB + cb) * 32 + l32;
        const int g = q * (GB / 2) + pair;
        w[cb] = *reinterpret_cast<const uint4*>(&bs[c * TK + swz_b<GB>(c, g) * 16]);
      }
    }
  };
  // One step on the stage at cur; nxt is the next step's stage.  The
  // sub-steps are software-pipelined: the fragments of sub-step s+1 are
  // requested before the MFMAs of sub-step s.  Before the MFMAs of the last
  // sub-step, whose operands are then in registers, the step hands over:
  // wait for the next stage's loads, barrier (after which no wave reads cur
  // any more), refill cur with step st+3 and request the next step's first
  // fragments from nxt.  Neither the barrier nor the next step's LDS latency
  // then leaves the matrix pipe without queued work.  The gradient issues a
  // sub-step's MFMAs column block by column block and builds the next
  // sub-step's column-block-0 one-hot fragment (bnext0) between them, so
  // that its one-hot arithmetic overlaps MFMAs instead of preceding them
  // (every accumulator still sums its K steps in the same order).
  bf16x8_t bnext0;
  auto step = [&](uint8_t* cur, const uint8_t* nxt, int st) {
    uint32_t bsel = 0;
    if constexpr (FWD) {
      const int r = st % inner;
      bsel = (uint32_t)(r - (r / CCM_NA) * CCM_NA);
    }
#pragma unroll
    for (int sub = 0; sub < SUBS; sub++) {
      if (sub + 1 < SUBS) {
        lda_frags(cur, af[(sub + 1) & 1], sub + 1);
        if (((sub + 1) & 1) == 0) ld_bytes(cur, bw[((sub + 1) >> 1) & 1], (sub + 1) >> 1);
      }
      bf16x8_t bfrag[CB];
      if constexpr (FWD) {
#pragma unroll
        for (int cb = 0; cb < CB; cb++) {
          const int d = (sub >> 1) * 8 + (sub & 1) * 4;
          bfrag[cb] = onehot8_spread(bsp[cb][d], bsp[cb][d + 1], bsp[cb][d + 2], bsp[cb][d + 3],
                                     bsel * 0x00010001u);
        }
      } else {
        bfrag[0] = bnext0;
        const uint4 w = bw[(sub >> 1) & 1][1];
        bfrag[1] = onehot8((sub & 1) ? make_uint2(w.z, w.w) : make_uint2(w.x, w.y), bval[1]);
      }
      if (sub == SUBS - 1) {
        // loads in flight: those of steps st+1 and st+2 (and, in the forward, a
        // block's bytes issued before the latter).  The refill and the LDS
        // reads below are unconditional: past the last step the refill
        // repeats the last step's loads into a stage nobody reads and the
        // reads go unused, so that no path without them makes the compiler
        // wait for all loads before the LDS reads or this sub-step's MFMAs.
        ccm_wait_vmcnt<PER_STEP>();
        ccm_lds_barrier();
        if constexpr (FWD) {
          // the next step opens a K block: spread its bytes (loaded a block
          // earlier) and request the following block's
          const int kn = (st + 1) / inner;
          if (st + 1 < nsteps && st + 1 == kn * inner) {
            spread_bytes();
            if (kn + 1 < kblocks) loadb(bcur, kn + 1);
          }
        }
        lda_frags(nxt, af[0], 0);
        ld_bytes(nxt, bw[0], 0);
        // keep the reads ahead of the refill: with the refill's loads in
        // flight the compiler cannot tell the two stages apart and waits
        __builtin_amdgcn_sched_barrier(0);
        issue(cur, st + 3 < nsteps ? st + 3 : nsteps - 1);
      }
      if constexpr (FWD) {
#pragma unroll
        for (int rb = 0; rb < RB; rb++) {
          const bf16x8_t afrag = __builtin_bit_cast(bf16x8_t, af[sub & 1][rb]);
#pragma unroll
          for (int cb = 0; cb < CB; cb++)
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afrag, bfrag[cb], acc[rb][cb], 0, 0, 0);
        }
      } else {
        // column block 0's MFMAs, the next sub-step's column-block-0
        // fragment (into the registers bfrag[0] leaves: past the step's
        // last sub-step from the next stage's bytes, read at the hand-over),
        // then column block 1's MFMAs, which the one-hot arithmetic overlaps
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
          acc[rb][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, af[sub & 1][rb]),
                                                               bfrag[0], acc[rb][0], 0, 0, 0);
        {
          const int ns = sub + 1 < SUBS ? sub + 1 : 0;
          const uint4 w = bw[(ns >> 1) & 1][0];
          bnext0 = onehot8((ns & 1) ? make_uint2(w.z, w.w) : make_uint2(w.x, w.y), bval[0]);
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
          acc[rb][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8_t, af[sub & 1][rb]),
                                                               bfrag[1], acc[rb][1], 0, 0, 0);
      }
    }
  };

  static_assert(SB == 0 || SB >= A_BYTES + TNB, "stage size");
  static_assert(FWD || CB == 2, "gradient: two column blocks per wave");
  constexpr int STAGE_BYTES = SB > 0 ? SB : A_BYTES + TNB;
  uint8_t* s0 = dtl_stage<STAGE_BYTES, 0>();
  uint8_t* s1 = dtl_stage<STAGE_BYTES, 1>();
  uint8_t* s2 = dtl_stage<STAGE_BYTES, 2>();
  // prologue: all three stages requested, drained, the first fragments read
  issue(s0, 0);
  if (nsteps > 1) issue(s1, 1);
  if (nsteps > 2) issue(s2, 2);
  if constexpr (FWD) loadb(bcur, 0);
  ccm_wait_vmcnt<0>();
  ccm_lds_barrier();
  if constexpr (FWD) {
    spread_bytes();
    if (kblocks > 1) loadb(bcur, 1);
  }
  lda_frags(s0, af[0], 0);
  ld_bytes(s0, bw[0], 0);
  if constexpr (!FWD) bnext0 = onehot8(make_uint2(bw[0][0].x, bw[0][0].y), bval[0]);
  // one flat loop over the steps, three (one per stage) per iteration: a
  // nested loop over K blocks made the compiler lose track of which stage
  // the loads in flight at the inner loop's head write, and wait for all of
  // them before the first step's LDS reads
  if constexpr (MODE == 1) {
    // complete triples without conditions, so that each iteration is one
    // scheduling region and a step's one-hot build for the next step
    // interleaves with its last MFMAs; the last one or two steps after it
    int st = 0;
    for (; st + 3 <= nsteps; st += 3) {
      step(s0, s1, st);
      step(s1, s2, st + 1);
      step(s2, s0, st + 2);
    }
    if (st < nsteps) {
      step(s0, s1, st);
      if (st + 1 < nsteps) step(s1, s2, st + 1);
    }
  } else {
    // (the same form for the forward, whose K blocks open at every 21st step,
    // measured slower)
    for (int st = 0; st < nsteps; st += 3) {
      step(s0, s1, st);
      if (st + 1 < nsteps) step(s1, s2, st + 1);
      if (st + 2 < nsteps) step(s2, s0, st + 2);
    }
  }
  // the refills past the last step complete before the workgroup ends
  ccm_wait_vmcnt<0>();

  if constexpr (MODE == 2) {
    static_assert(CB == 2 && WN % 2 == 0, "fused epilogue: 64 sequences per wave column");
    // Softmax of the tile's logits in rounds of 128 sequences (the wave
    // columns wn = 2r, 2r + 1): those waves put their 32-column blocks into
    // four LDS buffers of TM rows x 32 floats (the three stages and one more;
    // the writes meet two-way bank conflicts, a swizzle that avoids them made
    // the compiler keep every write address in a register).  Item it of the
    // round is alignment column jj = it / 2 for the 64 sequences of wave
    // column 2r + it % 2, one sequence per lane; waves take items it = w,
    // w + NW, .. (three each with eight waves).  The objective and
    // node-gradient terms are summed over each row of 16 lanes and the eight
    // row sums of a 128-sequence block combined in a fixed order, the same
    // for every tile shape.
    ccm_lds_barrier();  // no wave reads a stage any more
    float* const b0 = reinterpret_cast<float*>(s0);
    float* const b1 = reinterpret_cast<float*>(s1);
    float* const b2 = reinterpret_cast<float*>(s2);
    float* const b3 = reinterpret_cast<float*>(dtl_stage<A_BYTES, 3>());
    // per round, wave column, row of lanes and column jj: the 20
    // node-gradient sums and the objective; the singles V(j, a) of the tile
    constexpr int NSLOT = (WN / 2) * 2 * 4 * JG;
    __shared__ float gsum[NSLOT * 20];
    __shared__ double fsum[NSLOT];
    __shared__ float xs[20 * JG];
    for (int e = tid; e < 20 * JG; e += NW * 64) {
      const int a = e / JG, j = mt * JG + (e - a * JG);
      xs[e] = j < epi.L ? epi.x1[(long)a * epi.L + j] : 0.f;
    }
    const ccm_rsrc_t drsrc = ccm_make_rsrc(epi.dt);
    const uint32_t dstep = (uint32_t)((long)Lp * epi.lddt * 2), dlo = (uint32_t)(epi.lo_off * 2);
    for (int r = 0; r < WN / 2; r++) {
      if ((wn >> 1) == r) {
        float* const w0 = (wn & 1) ? b2 : b0;
        float* const w1 = (wn & 1) ? b3 : b1;
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
          for (int cb = 0; cb < CB; cb++)
#pragma unroll
            for (int i = 0; i < 16; i++) {
              const int R = (wm * RB + rb) * 32 + (i >> 2) * 8 + 4 * q + (i & 3);
              (cb ? w1 : w0)[R * 32 + l32] = acc[rb][cb][i];
            }
      }
      __syncthreads();
      fwd_softmax_round<JG, NW, 32>(b0, b1, b2, b3, xs, gsum, fsum, epi, mt, n0, Ncol, r, wave_u, lane,
                                    drsrc, dstep, dlo);
      __syncthreads();
    }
    fwd_softmax_partials<JG, NW * 64, WN / 2>(gsum, fsum, epi, mt, n0, Ncol, tid);
  } else {
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
}

template <int MODE, int WM, int WN, int RB, int CB, int TK, int S, int OCC>
__global__ __launch_bounds__(64 * WM * WN, OCC) void k_gemm_onehot_dtl(
    const uint16_t* __restrict__ A, long lda, int M, int Kb, int passes,
    const uint8_t* __restrict__ bytes, long ldbytes, int Lp, int Ncol, float* __restrict__ C,
    long ldc, int accumulate, uint32_t tiles_n, uint32_t tiles, uint32_t total, ccm_fwd_epi epi) {
  const uint32_t qid = ccm_xcd_remap(blockIdx.x, total);
  if (qid >= tiles) return;
  gemm_onehot_tile_dtl<MODE, WM, WN, RB, CB, TK, S>(A, lda, M, Kb, passes, bytes, ldbytes, Lp, Ncol, C,
                                                    ldc, accumulate, qid / tiles_n, qid % tiles_n, epi);
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

// The banded gradient (MODE 1) in one launch: tiles_a tiles of 256 x 256
// over rows [0, rows_a), then tiles_b tiles of 192 x 256 over rows
// [rows_a, M).  Each XCD takes a contiguous run of the 256-row tiles (the
// remainder on the low XCDs) and then a contiguous run of the 192-row tiles
// (the remainder on the high XCDs), so that every XCD runs about as many of
// each; within an XCD the 256-row tiles are dispatched first, and a CU that
// finishes one takes a 192-row tile without waiting for a kernel boundary.
// Both shapes use the 256-row tile's three LDS stages.
template <int TK, int S>
__device__ __forceinline__ void grad_banded_tile(const uint16_t* __restrict__ A, long lda, int M, int Kb,
                                                 int passes, const uint8_t* __restrict__ bytes, long ldbytes,
                                                 int Lp, int Ncol, float* __restrict__ C, long ldc,
                                                 int accumulate, uint32_t tiles_n, int rows_a,
                                                 uint32_t tiles_a, uint32_t tiles_b) {
  constexpr int SB = 256 * TK * 2 + 256 * TK;
  const uint32_t x = blockIdx.x & 7u, i = blockIdx.x >> 3;
  const uint32_t qa = tiles_a >> 3, ra = tiles_a & 7u, qb = tiles_b >> 3, rb = tiles_b & 7u;
  const uint32_t cnt_a = qa + (x < ra ? 1u : 0u), start_a = x * qa + (x < ra ? x : ra);
  const uint32_t cnt_b = qb + (x + rb >= 8u ? 1u : 0u), start_b = x * qb + (x + rb > 8u ? x + rb - 8u : 0u);
  const ccm_fwd_epi epi{};
  if (i < cnt_a) {
    const uint32_t t = start_a + i;
    gemm_onehot_tile_dtl<1, 2, 4, 4, 2, TK, S, SB>(A, lda, rows_a, Kb, passes, bytes, ldbytes, Lp, Ncol, C,
                                                   ldc, accumulate, t / tiles_n, t % tiles_n, epi);
  } else if (i - cnt_a < cnt_b) {
    const uint32_t t = start_b + (i - cnt_a);
    gemm_onehot_tile_dtl<1, 2, 4, 3, 2, TK, S, SB>(A + (long)rows_a * lda, lda, M - rows_a, Kb, passes, bytes,
                                                   ldbytes, Lp, Ncol, C + (long)rows_a * ldc, ldc, accumulate,
                                                   t / tiles_n, t % tiles_n, epi);
  }
}

template <int TK, int S>
__global__ __launch_bounds__(512, 1) void k_gemm_grad_banded(
    const uint16_t* __restrict__ A, long lda, int M, int Kb, int passes,
    const uint8_t* __restrict__ bytes, long ldbytes, int Lp, int Ncol, float* __restrict__ C,
    long ldc, int accumulate, uint32_t tiles_n, int rows_a, uint32_t tiles_a, uint32_t tiles_b) {
  grad_banded_tile<TK, S>(A, lda, M, Kb, passes, bytes, ldbytes, Lp, Ncol, C, ldc, accumulate, tiles_n, rows_a,
                          tiles_a, tiles_b);
}

// The same for alignment pi of a batch, its buffers from its descriptor
// (the whole alignment as one chunk, as the grouped gradient takes it); no
// work once the alignment's solve has ended.
template <int TK, int S>
__global__ __launch_bounds__(512, 1) void k_gemm_grad_banded_desc(const ccm_bprob* __restrict__ probs, int pi,
                                                                  uint32_t tiles_n, int rows_a,
                                                                  uint32_t tiles_a, uint32_t tiles_b) {
  const ccm_bprob& B = probs[pi];
  if (!B.ctl->eval_on) return;
  grad_banded_tile<TK, S>(B.dt, B.lddt, B.P, B.ncols, B.passes, B.msapt, B.ldmt, B.Lp, B.P, B.gt, B.P, 0,
                          tiles_n, rows_a, tiles_a, tiles_b);
}

// The same tiles over a group of alignments, all with this tile variant:
// the XCD remap runs over the whole grid, so that each XCD takes a run of
// consecutive tiles (mostly of one alignment).
//   MODE 0: Zt = Wt . onehot(msap)      (M = P, K = P, cols = ncols)
//   MODE 1: Gt = Dt . onehot(msapt)^T   (M = P, K = ncols, cols = P)
// logical tile of a grouped contraction's workgroup: from the launch's tile
// map when it has one, else the XCD-contiguous order
__device__ __forceinline__ uint32_t ccm_group_tile(const uint32_t* __restrict__ qmap, uint32_t total) {
  return qmap ? qmap[blockIdx.x] : ccm_xcd_remap(blockIdx.x, total);
}

template <int MODE, int WM, int WN, int RB, int CB, int TK, int OCC>
__global__ __launch_bounds__(64 * WM * WN, OCC) void k_gemm_onehot_grouped(
    const ccm_bprob* __restrict__ probs, const uint32_t* __restrict__ offs,
    const int32_t* __restrict__ pidx, int np, uint32_t total, const uint32_t* __restrict__ qmap) {
  constexpr int TN = WN * CB * 32;
  const uint32_t qid = ccm_group_tile(qmap, total);
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

template <int MODE, int WM, int WN, int RB, int CB, int TK, int S, int OCC>
__global__ __launch_bounds__(64 * WM * WN, OCC) void k_gemm_onehot_dtl_grouped(
    const ccm_bprob* __restrict__ probs, const uint32_t* __restrict__ offs,
    const int32_t* __restrict__ pidx, int np, uint32_t total, const uint32_t* __restrict__ qmap) {
  constexpr int TN = WN * CB * 32;
  const uint32_t qid = ccm_group_tile(qmap, total);
  if (qid >= offs[np]) return;
  const int gi = ccm_group_find(offs, np, qid);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  const uint32_t t = qid - offs[gi];
  if constexpr (MODE == 0) {
    const uint32_t tiles_n = (uint32_t)((B.ncols + TN - 1) / TN);
    gemm_onehot_tile_dtl<0, WM, WN, RB, CB, TK, S>(B.wt, B.ldwt, B.P, B.P, B.passes_fwd, B.msap, B.Lp, B.Lp,
                                                   B.ncols, B.zt, B.ncols, 0, t / tiles_n, t % tiles_n,
                                                   ccm_fwd_epi{});
  } else if constexpr (MODE == 2) {
    // the softmax of softmax_cols_grouped as the epilogue (whole alignment
    // as one chunk)
    const uint32_t tiles_n = (uint32_t)((B.ncols + TN - 1) / TN);
    const ccm_fwd_epi epi{B.x,   B.msapt, B.ldmt,          B.w,      B.L, B.N, B.dt, B.lddt,
                          (long)B.ncols, B.passes, 1, 0, B.fpart, B.g1part};
    gemm_onehot_tile_dtl<2, WM, WN, RB, CB, TK, S>(B.wt, B.ldwt, B.P, B.P, B.passes_fwd, B.msap, B.Lp, B.Lp,
                                                   B.ncols, nullptr, 0, 0, t / tiles_n, t % tiles_n, epi);
  } else {
    const uint32_t tiles_n = (uint32_t)((B.P + TN - 1) / TN);
    gemm_onehot_tile_dtl<1, WM, WN, RB, CB, TK, S>(B.dt, B.lddt, B.P, B.ncols, B.passes, B.msapt, B.ldmt,
                                                   B.Lp, B.P, B.gt, B.P, 0, t / tiles_n, t % tiles_n,
                                                   ccm_fwd_epi{});
  }
}

static inline uint32_t round8(uint32_t v) { return (v + 7u) & ~7u; }

extern "C" {

int ccm_launch_convert_wt(const float* d_x, int L, int Lp, uint16_t* d_wt, long lda, int passes,
                          hipStream_t s) {
  const int P = CCM_NA * Lp;
  if (Lp % 64 || Lp < L || lda < (long)passes * P || lda % 8) return CCM_ERR_ARG;
  const uint32_t tiles_r = (uint32_t)((P + 63) / 64);
  if (((uintptr_t)d_x & 15u) == 0) {
    const uint32_t tiles_c = (uint32_t)((CCM_NA * L + CW_TC - 1) / CW_TC);
    hipLaunchKernelGGL(k_convert_wt128, dim3(tiles_c * tiles_r), dim3(256), 0, s, d_x + ccm_nsp(L), L, Lp,
                       d_wt, lda, passes, tiles_r);
  } else {
    const uint32_t tiles_c = (uint32_t)((CCM_NA * L + 63) / 64);
    hipLaunchKernelGGL(k_convert_wt, dim3(tiles_c * tiles_r), dim3(256), 0, s, d_x + ccm_nsp(L), L, Lp,
                       d_wt, lda, passes, tiles_r);
  }
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"

// Tile variants.  Every variant runs the contraction in K steps of 64 in
// the same order (for the forward: blocks of 64 coupling columns outermost,
// then pass and residue), so all variants give bit-identical results.
//   {waves M x N, 32 x 32 blocks per wave M x N, workgroups per CU, staging}
//   0: 2x4, 4x2, 1, registers -> 256 x 256, 8 waves
//   1: 2x2, 4x2, 2, registers -> 256 x 128, 4 waves
//   2: 2x2, 2x2, 2, registers -> 128 x 128, 4 waves
//   3: 2x4, 4x2, 1, direct    -> 256 x 256, 8 waves
//   4: 2x2, 4x2, 1, direct    -> 256 x 128, 4 waves
//   5: 2x2, 2x2, 2, direct    -> 128 x 128, 4 waves
//   6: 2x4, 3x2, 1, direct    -> 192 x 256, 8 waves (gradient only)
// "registers": next step prefetched through registers, two LDS stages;
// "direct": direct-to-LDS loads, three stages in flight.
// Variant 6 is the short row band of the gradient's banded launch (a band
// of 256-row tiles, then one of 192-row tiles, so that each CU runs one
// tile of each instead of two rounds of 256-row tiles); the forward
// launchers take variant 3 in its place.
#define GM_VARIANTS(X)   \
  X(0, 2, 4, 4, 2, 1, 0) \
  X(1, 2, 2, 4, 2, 2, 0) \
  X(2, 2, 2, 2, 2, 2, 0) \
  X(3, 2, 4, 4, 2, 1, 1) \
  X(4, 2, 2, 4, 2, 1, 1) \
  X(5, 2, 2, 2, 2, 2, 1) \
  X(6, 2, 4, 3, 2, 1, 1)
#define GM_NVAR 7
#define GM_GRAD_ONLY 6
#define GM_TK 64
static const int gm_tm[GM_NVAR] = {256, 256, 128, 256, 256, 128, 192};
static const int gm_tn[GM_NVAR] = {256, 128, 128, 256, 128, 128, 256};
static const int gm_wgpc[GM_NVAR] = {1, 2, 2, 1, 1, 2, 1};
static const int gm_direct[GM_NVAR] = {0, 0, 0, 1, 1, 1, 1};
// throughput of a fully loaded CU relative to variant 3 (measured at
// L = 250, N = 5000 and L = 300, N = 3000 on MI355X)
static const float gm_eff[GM_NVAR] = {0.66f, 0.70f, 0.70f, 1.0f, 0.65f, 0.88f, 0.90f};

// the variant a forward launch runs for a requested one
static int fwd_variant(int variant) { return variant == GM_GRAD_ONLY ? 3 : variant; }

// kernel of a variant (only the staging form in use is instantiated)
template <int MODE, int WM, int WN, int RB, int CB, int OCC, int DTL, bool GROUPED>
static const void* gm_kernel() {
  if constexpr (MODE == 2 && !DTL) {
    return nullptr;  // the fused forward exists in the direct-to-LDS form only
  } else if constexpr (MODE != 1 && RB == 3) {
    return nullptr;  // the 192-row tile is the gradient's
  } else if constexpr (GROUPED) {
    if constexpr (DTL)
      return (const void*)k_gemm_onehot_dtl_grouped<MODE, WM, WN, RB, CB, GM_TK, 3, OCC>;
    else
      return (const void*)k_gemm_onehot_grouped<MODE, WM, WN, RB, CB, GM_TK, OCC>;
  } else {
    if constexpr (DTL)
      return (const void*)k_gemm_onehot_dtl<MODE, WM, WN, RB, CB, GM_TK, 3, OCC>;
    else
      return (const void*)k_gemm_onehot<MODE, WM, WN, RB, CB, GM_TK, OCC>;
  }
}

template <int MODE>
static int launch_gemm(int variant, const uint16_t* A, long lda, int M, int Kb, int passes,
                       const uint8_t* bytes, long ldbytes, int Lp, int Ncol, float* C, long ldc,
                       int accumulate, hipStream_t s) {
  if (variant < 0 || variant >= GM_NVAR) return CCM_ERR_ARG;
  if ((MODE == 0 && Lp % GM_TK) || Kb % GM_TK) return CCM_ERR_ARG;
  const uint32_t tiles_m = (M + gm_tm[variant] - 1) / gm_tm[variant];
  const uint32_t tiles_n = (Ncol + gm_tn[variant] - 1) / gm_tn[variant];
  const uint32_t tiles = tiles_m * tiles_n, total = round8(tiles);
  const void* fn = nullptr;
  int nthreads = 0;
#define GM_CASE(V, WM, WN, RB, CB, OCC, DTL)                                   \
  case V:                                                                      \
    fn = gm_kernel<MODE, WM, WN, RB, CB, OCC, DTL, false>();                  \
    nthreads = 64 * WM * WN;                                                   \
    break;
  switch (variant) { GM_VARIANTS(GM_CASE) }
#undef GM_CASE
  if (!fn) return CCM_ERR_ARG;
  ccm_fwd_epi epi{};  // (the register-staged kernels take no epilogue argument)
  void* args[] = {(void*)&A,     (void*)&lda,   (void*)&M,  (void*)&Kb,   (void*)&passes,
                  (void*)&bytes, (void*)&ldbytes, (void*)&Lp, (void*)&Ncol, (void*)&C,
                  (void*)&ldc,   (void*)&accumulate, (void*)&tiles_n, (void*)&tiles, (void*)&total,
                  (void*)&epi};
  if (hipLaunchKernel(fn, dim3(total), dim3(nthreads), args, 0, s) != hipSuccess) return CCM_ERR_LAUNCH;
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

template <int MODE>
static int launch_gemm_grouped(int variant, const ccm_bprob* probs, const ccm_group& g, hipStream_t s) {
  if (variant < 0 || variant >= GM_NVAR || g.np <= 0) return CCM_ERR_ARG;
  // with a tile map the grid is the map; else padded to whole XCD rounds
  const uint32_t total = g.qmap ? g.total : round8(g.total);
  const void* fn = nullptr;
  int nthreads = 0;
#define GG_CASE(V, WM, WN, RB, CB, OCC, DTL)                                                    \
  case V:                                                                                       \
    fn = gm_kernel<MODE, WM, WN, RB, CB, OCC, DTL, true>();                                     \
    nthreads = 64 * WM * WN;                                                                    \
    break;
  switch (variant) { GM_VARIANTS(GG_CASE) }
#undef GG_CASE
  if (!fn) return CCM_ERR_ARG;
  const uint32_t* offs = g.offs;
  const int32_t* pidx = g.pidx;
  int np = g.np;
  const uint32_t* qmap = g.qmap;
  void* args[] = {(void*)&probs, (void*)&offs, (void*)&pidx, (void*)&np, (void*)&total, (void*)&qmap};
  if (hipLaunchKernel(fn, dim3(total), dim3(nthreads), args, 0, s) != hipSuccess) return CCM_ERR_LAUNCH;
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// workgroups of a launch for a tile variant
static uint32_t gemm_tiles(int variant, int M, int Ncol) {
  return (uint32_t)(((M + gm_tm[variant] - 1) / gm_tm[variant]) * ((Ncol + gm_tn[variant] - 1) / gm_tn[variant]));
}

static int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      cus = v;
    else
      cus = 256;
  }
  return cus;
}

// default variant: the direct-to-LDS tile with the least estimated time,
// waves of workgroups over the CUs times the work of one wave per CU over
// the variant's relative throughput (the last, partial wave costs a full one)
static int gemm_default_variant(int M, int Ncol) {
  const int cus = device_cus();
  int best = 3;
  float best_t = 0.f;
  const int cand[2] = {3, 5};
  for (int i = 0; i < 2; i++) {
    const int v = cand[i];
    const uint32_t tiles = gemm_tiles(v, M, Ncol);
    const uint32_t slots = (uint32_t)(cus * gm_wgpc[v]);
    const float waves = (float)((tiles + slots - 1) / slots);
    const float t = waves * (float)(gm_tm[v] * gm_tn[v] * gm_wgpc[v]) / gm_eff[v];
    if (i == 0 || t < best_t) {
      best_t = t;
      best = v;
    }
  }
  return best;
}

// Banded gradient launch: rows [0, rows_a) in 256-row tiles (variant 3),
// the rest in 192-row tiles (variant 6), two launches, when the estimated
// time (rounds of tiles over the CUs, a 192-row tile costing 0.75 of a
// 256-row one over its relative throughput) beats the best single variant
// by more than 2 %.  At P = 5376 (L = 250): 252 tiles of each shape, one
// of each per CU, against 441 tiles of 256 x 256 in two rounds.  Returns
// rows_a, or 0 for a single-variant launch.
static int grad_band_rows(int P) {
  const int cus = device_cus();
  const int cols = (P + 255) / 256;
  const int v = gemm_default_variant(P, P);
  const uint32_t slots = (uint32_t)(cus * gm_wgpc[v]);
  const float plain = (float)((gemm_tiles(v, P, P) + slots - 1) / slots) *
                      (float)(gm_tm[v] * gm_tn[v] * gm_wgpc[v]) / (gm_eff[v] * 65536.f);
  const float rel_b = 0.75f / gm_eff[GM_GRAD_ONLY];
  float best = plain * 0.98f;
  int rows_a = 0;
  for (int a = 1; 256 * a < P; a++) {
    const int b = (P - 256 * a + 191) / 192;
    const float t = (float)((a * cols + cus - 1) / cus) + (float)((b * cols + cus - 1) / cus) * rel_b;
    if (t < best) {
      best = t;
      rows_a = 256 * a;
    }
  }
  return rows_a;
}

// The banded launch of a P x P gradient, where it applies (returns 0 where
// not): the rows of 256-row tiles, the column tiles, the tiles of each band
// and the grid (8 XCDs times the most workgroups one XCD takes: its
// 256-row tiles, remainder on the low XCDs, and its 192-row tiles,
// remainder on the high XCDs).
static int grad_band_plan(int P, int* rows_a, uint32_t* tiles_n, uint32_t* tiles_a, uint32_t* tiles_b,
                          uint32_t* grid) {
  *rows_a = grad_band_rows(P);
  if (*rows_a <= 0) return 0;
  *tiles_n = (uint32_t)((P + 255) / 256);
  *tiles_a = (uint32_t)(*rows_a / 256) * *tiles_n;
  *tiles_b = (uint32_t)((P - *rows_a + 191) / 192) * *tiles_n;
  uint32_t per = 0;
  for (uint32_t x = 0; x < 8; x++) {
    const uint32_t c = *tiles_a / 8 + (x < *tiles_a % 8 ? 1u : 0u) + *tiles_b / 8 + (x + *tiles_b % 8 >= 8 ? 1u : 0u);
    per = c > per ? c : per;
  }
  *grid = 8 * per;
  return 1;
}

extern "C" {

// The banded gradient of alignment pi of a batch (descriptors at d_probs,
// P its coupling rows); 1 (nothing launched) where P takes no banding
int ccm_launch_gemm_grad_banded_desc(const ccm_bprob* d_probs, int pi, int P, hipStream_t s) {
  int rows_a;
  uint32_t tiles_n, tiles_a, tiles_b, grid;
  if (!d_probs || pi < 0 || P <= 0) return CCM_ERR_ARG;
  if (!grad_band_plan(P, &rows_a, &tiles_n, &tiles_a, &tiles_b, &grid)) return 1;
  hipLaunchKernelGGL((k_gemm_grad_banded_desc<GM_TK, 3>), dim3(grid), dim3(512), 0, s, d_probs, pi, tiles_n, rows_a,
                     tiles_a, tiles_b);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_gemm_variants(void) { return GM_NVAR; }

// output tile of a variant (rows x columns); CCM_ERR_ARG for an unknown variant
int ccm_gemm_tile_shape(int variant, int* tm, int* tn) {
  if (variant < 0 || variant >= GM_NVAR) return CCM_ERR_ARG;
  *tm = gm_tm[variant];
  *tn = gm_tn[variant];
  return CCM_OK;
}

// whether a variant stages through direct-to-LDS loads (the register-staged
// ones are slower at every measured shape; the batched autotuner skips them)
int ccm_gemm_variant_direct(int variant) {
  return variant >= 0 && variant < GM_NVAR ? gm_direct[variant] : 0;
}

// tiles of a variant over an M x Ncol output (0 for an unknown variant)
uint32_t ccm_gemm_tile_count(int variant, int M, int Ncol) {
  if (variant < 0 || variant >= GM_NVAR) return 0;
  return gemm_tiles(variant, M, Ncol);
}

// whether a variant fits a contraction: the forward (mode 0) needs Lp, the
// gradient (mode 1) kcols, a multiple of the K step 64
int ccm_gemm_variant_ok(int variant, int mode, int Lp, int kcols) {
  if (variant < 0 || variant >= GM_NVAR) return 0;
  if (mode == 0 && variant == GM_GRAD_ONLY) return 0;
  return mode == 0 ? Lp % GM_TK == 0 : kcols % GM_TK == 0;
}

// default variant of the forward (mode 0) or gradient (mode 1) contraction
int ccm_gemm_default_variant_for(int mode, int Lp, int ncols) {
  const int P = CCM_NA * Lp;
  return gemm_default_variant(P, mode == 0 ? ncols : P);
}

int ccm_convert_wt_blocks(int L, int Lp) {
  return ((CCM_NA * Lp + 63) / 64) * ((CCM_NA * L + CW_TC - 1) / CW_TC);
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

int ccm_gemm_default_variant(int M, int Ncol) { return gemm_default_variant(M, Ncol); }

// ---- forward contraction with the column softmax fused (MODE 2) --------
// Tiles of the direct-to-LDS variants only; a register-staged variant maps
// to the direct one of the same shape.  Tile rows cover TM / 21 alignment
// columns, so the row tiles number ceil(L / (TM / 21)).
static int fused_variant(int variant) {
  variant = fwd_variant(variant);
  return variant >= 0 && variant < 3 ? variant + 3 : variant;
}

static uint32_t fused_tiles(int v, int L, int ncols) {
  const int jg = gm_tm[v] / CCM_NA;
  return (uint32_t)(((L + jg - 1) / jg) * ((ncols + gm_tn[v] - 1) / gm_tn[v]));
}

int ccm_fwd_softmax_blocks(int N) { return N <= 0 ? 0 : (N + 127) / 128; }

uint32_t ccm_fwd_softmax_tile_count(int variant, int L, int ncols) {
  variant = fused_variant(variant);
  if (variant < 3 || variant >= GM_GRAD_ONLY || L < 1 || ncols < 1) return 0;
  return fused_tiles(variant, L, ncols);
}

int ccm_fwd_softmax_tile_rows(int variant, int L) {
  variant = fused_variant(variant);
  if (variant < 3 || variant >= GM_GRAD_ONLY || L < 1) return 0;
  const int jg = gm_tm[variant] / CCM_NA;
  return (L + jg - 1) / jg;
}

int ccm_fwd_softmax_fits(int Lp, int passes_fwd, int passes, int ncols) {
  const long P = (long)CCM_NA * Lp;
  return (P + 1) * (passes_fwd * P) * 2 <= 0x7FFFFFFFL && P * passes * (long)ncols * 2 <= 0x7FFFFFFFL;
}

int ccm_launch_gemm_fwd_softmax_grouped(const ccm_bprob* d_probs, ccm_group g, int variant, hipStream_t s) {
  variant = fused_variant(variant);
  if (variant < 3 || variant >= GM_GRAD_ONLY) return CCM_ERR_ARG;
  return launch_gemm_grouped<2>(variant, d_probs, g, s);
}

int ccm_fwd_softmax_default_variant(int L, int ncols) {
  const int cus = device_cus();
  int best = 3;
  float best_t = 0.f;
  const int cand[2] = {3, 5};
  for (int i = 0; i < 2; i++) {
    const int v = cand[i];
    const uint32_t tiles = fused_tiles(v, L, ncols);
    const uint32_t slots = (uint32_t)(cus * gm_wgpc[v]);
    const float waves = (float)((tiles + slots - 1) / slots);
    const float t = waves * (float)(gm_tm[v] * gm_tn[v] * gm_wgpc[v]) / gm_eff[v];
    if (i == 0 || t < best_t) {
      best_t = t;
      best = v;
    }
  }
  return best;
}

// The logits of ccm_launch_gemm_fwd for the chunk's ncols sequences (never
// stored) with softmax_cols as the epilogue: Dt (mode 1), the objective
// partials fpart[blk][j] and the partials g1part[blk][a][j] = sum w P (the
// node gradient without its constant data term)
// (mode 1) per block blk = blk0 + n / 128 of 128 sequences.
int ccm_launch_gemm_fwd_softmax(const uint16_t* d_wt, long lda, int Lp, int passes_fwd,
                                const uint8_t* d_msap, int ncols, const float* d_x1,
                                const uint8_t* d_msapt, long ldmt, const float* d_w, int L, int Nc,
                                uint16_t* d_dt, long lddt, long lo_off, int passes, int mode, int blk0,
                                double* d_fpart, double* d_g1part, int variant, hipStream_t s) {
  const long P = (long)CCM_NA * Lp;
  if (Lp % 64 || Lp < L || L < 1 || ncols % 128 || ncols <= 0 || Nc < 1 || Nc > ncols ||
      lda < passes_fwd * P || lda % 8 || !d_fpart || (mode && (!d_dt || !d_g1part)) || blk0 < 0)
    return CCM_ERR_ARG;
  if (mode && (passes < 1 || passes > 2 || lddt < (long)passes * ncols || (passes > 1 && lo_off < ncols)))
    return CCM_ERR_ARG;
  // the gathered rows address the whole dense operand, and the stores the
  // whole of Dt, through 32-bit offsets
  if ((P + 1) * lda * 2 > 0x7FFFFFFFL || (mode && P * lddt * 2 > 0x7FFFFFFFL)) return CCM_ERR_ARG;
  variant = variant < 0 ? ccm_fwd_softmax_default_variant(L, ncols) : fused_variant(variant);
  if (variant < 3 || variant >= GM_GRAD_ONLY) return CCM_ERR_ARG;
  ccm_fwd_epi epi{d_x1, d_msapt, ldmt, d_w, L, Nc, d_dt, lddt, lo_off, passes, mode, blk0, d_fpart, d_g1part};
  const uint32_t tiles_n = (uint32_t)((ncols + gm_tn[variant] - 1) / gm_tn[variant]);
  const uint32_t tiles = fused_tiles(variant, L, ncols), total = round8(tiles);
  const int M = (int)P, Kb = (int)P, accumulate = 0;
  const long ldbytes = Lp, ldc = 0;
  float* C = nullptr;
  const void* fn = nullptr;
  int nthreads = 0;
  switch (variant) {
    case 3: fn = (const void*)k_gemm_onehot_dtl<2, 2, 4, 4, 2, GM_TK, 3, 1>; nthreads = 512; break;
    case 4: fn = (const void*)k_gemm_onehot_dtl<2, 2, 2, 4, 2, GM_TK, 3, 1>; nthreads = 256; break;
    default: fn = (const void*)k_gemm_onehot_dtl<2, 2, 2, 2, 2, GM_TK, 3, 2>; nthreads = 256; break;
  }
  void* args[] = {(void*)&d_wt,   (void*)&lda,     (void*)&M,  (void*)&Kb,    (void*)&passes_fwd,
                  (void*)&d_msap, (void*)&ldbytes, (void*)&Lp, (void*)&ncols, (void*)&C,
                  (void*)&ldc,    (void*)&accumulate, (void*)&tiles_n, (void*)&tiles, (void*)&total,
                  (void*)&epi};
  if (hipLaunchKernel(fn, dim3(total), dim3(nthreads), args, 0, s) != hipSuccess) return CCM_ERR_LAUNCH;
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// Zt[c][n] (P x ldc floats, ldc >= ncols) = sum_r Wt[c][r] [x_nk = b] for the
// ncols (multiple of 128) sequences whose padded MSA rows start at d_msap;
// variant < 0 picks the default tile
int ccm_launch_gemm_fwd(const uint16_t* d_wt, long lda, int Lp, int passes, const uint8_t* d_msap,
                        int ncols, float* d_zt, long ldc, int variant, hipStream_t s) {
  const int P = CCM_NA * Lp;
  if (Lp % 64 || ncols % 128 || ncols <= 0 || lda < (long)passes * P || ldc < ncols || (lda % 8))
    return CCM_ERR_ARG;
  variant = variant < 0 ? gemm_default_variant(P, ncols) : fwd_variant(variant);
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
  if (variant < 0) {
    int rows_a;
    uint32_t tiles_n, tiles_a, tiles_b, grid;
    if (grad_band_plan(P, &rows_a, &tiles_n, &tiles_a, &tiles_b, &grid)) {
      const int M = P, Ncol = P;
      void* args[] = {(void*)&d_dt,  (void*)&lda,    (void*)&M,  (void*)&kcols, (void*)&passes,
                      (void*)&d_msapt, (void*)&ldmt, (void*)&Lp, (void*)&Ncol,  (void*)&d_gt,
                      (void*)&ldc,   (void*)&accumulate, (void*)&tiles_n, (void*)&rows_a,
                      (void*)&tiles_a, (void*)&tiles_b};
      if (hipLaunchKernel((const void*)k_gemm_grad_banded<GM_TK, 3>, dim3(grid), dim3(512), args, 0, s) !=
          hipSuccess)
        return CCM_ERR_LAUNCH;
      CCM_LAUNCH_CHECK();
      return CCM_OK;
    }
    variant = gemm_default_variant(P, P);
  }
  return launch_gemm<1>(variant, d_dt, lda, P, kcols, passes, d_msapt, ldmt, Lp, P, d_gt, ldc,
                        accumulate, s);
}

}  // extern "C"
