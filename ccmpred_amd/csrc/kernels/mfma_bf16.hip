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

#define GM_TM 256
#define GM_TN 128
#define GM_TK 32
#define GM_LDA (GM_TK + 8)  // LDS row stride in bf16 (80 bytes)

// coupling matrix -> Wt (bf16, P x lda), zero outside the real 21L x 21L block
__global__ __launch_bounds__(256) void k_convert_wt(const float* __restrict__ x2, int L, int Lp,
                                                    uint16_t* __restrict__ wt, long lda, int passes,
                                                    uint32_t tiles_r) {
  __shared__ float T[64][65];
  const int P = CCM_NA * Lp;
  const int tc = blockIdx.x / tiles_r, tr = blockIdx.x % tiles_r;
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

// MODE 0: forward.  cols = sequences n;      B[kk][n] = [msap[n][kk%Lp] == kk/Lp], Kb = P
// MODE 1: gradient. cols = coupling rows r;  B[kk][r] = [msapt[r%Lp][kk] == r/Lp], Kb = Np
template <int MODE>
__global__ __launch_bounds__(256, 2) void k_gemm_onehot(const uint16_t* __restrict__ A, long lda,
                                                        int M, int Kb, int passes,
                                                        const uint8_t* __restrict__ bytes,
                                                        long ldbytes, int Lp, int Ncol,
                                                        float* __restrict__ C, long ldc,
                                                        int accumulate, uint32_t tiles_n,
                                                        uint32_t tiles, uint32_t total) {
  __shared__ __attribute__((aligned(16))) uint16_t As[2][GM_TM * GM_LDA];
  const uint32_t qid = ccm_xcd_remap(blockIdx.x, total);
  if (qid >= tiles) return;
  const int mt = qid / tiles_n, nt = qid % tiles_n;
  const int m0 = mt * GM_TM, n0 = nt * GM_TN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int l32 = lane & 31, q = lane >> 5;

  // per-lane one-hot column information for the 2 column blocks
  int colv[2];
  uint32_t bval[2];
  const uint8_t* bptr[2];
#pragma unroll
  for (int cb = 0; cb < 2; cb++) {
    const int col = n0 + wn * 64 + cb * 32 + l32;
    colv[cb] = col;
    if (MODE == 0) {
      bval[cb] = 0;
      bptr[cb] = bytes + (long)col * ldbytes;  // msap row n (Np rows, always allocated)
    } else {
      const int bb = col / Lp, kk = col % Lp;
      bval[cb] = (uint32_t)bb;                 // >= 21 past the real rows: never matches
      bptr[cb] = bytes + (long)kk * ldbytes;   // msapt row k
    }
  }

  floatx16 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = (floatx16)0.0f;

  const int K = passes * Kb;
  const int nsteps = K / GM_TK;
  // A tile loader: thread t loads row m0+t, 32 bf16 = 4 x 16 bytes
  const int arow = m0 + tid;
  const bool arow_ok = arow < M;
  const uint4* asrc = reinterpret_cast<const uint4*>(A + (long)(arow_ok ? arow : 0) * lda);
  uint4 pre[4];
#pragma unroll
  for (int i = 0; i < 4; i++) pre[i] = arow_ok ? asrc[i] : make_uint4(0, 0, 0, 0);
  {
    uint4* dst = reinterpret_cast<uint4*>(&As[0][tid * GM_LDA]);
#pragma unroll
    for (int i = 0; i < 4; i++) dst[i] = pre[i];
  }
  __syncthreads();

  // uniform position of the current K step inside one pass: for the forward
  // the coupling row block bb and the column offset kin (kin + 32 <= Lp since
  // Lp is a multiple of 32); for the gradient the sequence offset kin
  int bb = 0, kin = 0;
  for (int s = 0; s < nsteps; s++) {
    const int cur = s & 1;
    if (s + 1 < nsteps) {
      const uint4* src = asrc + (long)(s + 1) * (GM_TK / 8);
#pragma unroll
      for (int i = 0; i < 4; i++) pre[i] = arow_ok ? src[i] : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int sub = 0; sub < 2; sub++) {
      const int koff = kin + sub * 16 + 8 * q;  // first of this lane's 8 K indices
      bf16x8_t bfrag[2];
#pragma unroll
      for (int cb = 0; cb < 2; cb++) {
        const uint2 v = *reinterpret_cast<const uint2*>(bptr[cb] + koff);
        bfrag[cb] = onehot8(v, MODE == 0 ? (uint32_t)bb : bval[cb]);
      }
#pragma unroll
      for (int rb = 0; rb < 4; rb++) {
        const uint4 av =
            *reinterpret_cast<const uint4*>(&As[cur][(wm * 128 + rb * 32 + l32) * GM_LDA + sub * 16 + 8 * q]);
        const bf16x8_t afrag = __builtin_bit_cast(bf16x8_t, av);
#pragma unroll
        for (int cb = 0; cb < 2; cb++)
          acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afrag, bfrag[cb], acc[rb][cb], 0, 0, 0);
      }
    }
    if (s + 1 < nsteps) {
      uint4* dst = reinterpret_cast<uint4*>(&As[cur ^ 1][tid * GM_LDA]);
#pragma unroll
      for (int i = 0; i < 4; i++) dst[i] = pre[i];
    }
    __syncthreads();
    kin += GM_TK;
    if (MODE == 0) {
      if (kin == Lp) {
        kin = 0;
        if (++bb == CCM_NA) bb = 0;
      }
    } else {
      if (kin == Kb) kin = 0;
    }
  }

  // 32x32 accumulator layout: lane -> column l32; register i -> row
  // (i/4)*8 + 4*q + (i%4)
#pragma unroll
  for (int rb = 0; rb < 4; rb++) {
#pragma unroll
    for (int cb = 0; cb < 2; cb++) {
      const int col = colv[cb];
      if (col >= Ncol) continue;
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const int row = m0 + wm * 128 + rb * 32 + (i >> 2) * 8 + 4 * q + (i & 3);
        if (row < M) {
          float* dst = C + (long)row * ldc + col;
          *dst = accumulate ? *dst + acc[rb][cb][i] : acc[rb][cb][i];
        }
      }
    }
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

// Zt[c][n] (P x ldc floats, ldc >= ncols) = sum_r Wt[c][r] [x_nk = b] for the
// ncols (multiple of 128) sequences whose padded MSA rows start at d_msap
int ccm_launch_gemm_fwd(const uint16_t* d_wt, long lda, int Lp, int passes, const uint8_t* d_msap,
                        int ncols, float* d_zt, long ldc, hipStream_t s) {
  const int P = CCM_NA * Lp;
  if (Lp % 32 || ncols % GM_TN || ncols <= 0 || lda < (long)passes * P || ldc < ncols || (lda % 8))
    return CCM_ERR_ARG;
  const uint32_t tm = (P + GM_TM - 1) / GM_TM, tn = ncols / GM_TN;
  const uint32_t tiles = tm * tn, total = round8(tiles);
  hipLaunchKernelGGL(k_gemm_onehot<0>, dim3(total), dim3(256), 0, s, d_wt, lda, P, P, passes, d_msap,
                     (long)Lp, Lp, ncols, d_zt, ldc, 0, tn, tiles, total);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// Gt[c][r] (P x ldc floats, ldc >= P) (+)= sum_n Dt[c][n] [x_nk = b] over the
// kcols (multiple of 32) sequences of Dt; d_msapt points at the chunk's first
// column of the transposed MSA, whose rows are ldmt bytes apart
int ccm_launch_gemm_grad(const uint16_t* d_dt, long lda, int Lp, int passes, const uint8_t* d_msapt,
                         long ldmt, int kcols, float* d_gt, long ldc, int accumulate, hipStream_t s) {
  const int P = CCM_NA * Lp;
  if (Lp % 32 || kcols % GM_TK || kcols <= 0 || lda < (long)passes * kcols || ldc < P || (lda % 8))
    return CCM_ERR_ARG;
  const uint32_t tm = (P + GM_TM - 1) / GM_TM, tn = (P + GM_TN - 1) / GM_TN;
  const uint32_t tiles = tm * tn, total = round8(tiles);
  hipLaunchKernelGGL(k_gemm_onehot<1>, dim3(total), dim3(256), 0, s, d_dt, lda, P, kcols, passes,
                     d_msapt, ldmt, Lp, P, d_gt, ldc, accumulate, tn, tiles, total);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"
