// Per-(sequence, column) log-sum-exp over the 21 residue logits, the
// objective terms and the derivative operand of the pair gradient.
//
// For each (n, j) with logits z_a (a = 0..20, singles already included):
//   lse = log sum_a exp(z_a)                     (evaluate_cpu.c:78-87)
//   f  += w_n (lse - z_{x_nj})                   (evaluate_cpu.c:96-101)
//   P_a = exp(z_a - lse)                         (evaluate_cpu.c:89-93)
//   g1(j, a) += w_n (P_a - [x_nj = a]), a < 20   (evaluate_cpu.c:102-109)
// The reference sums exp(z_a) without a shift; here the maximum logit is
// subtracted first (same value, no overflow for large logits).
//
// Two layouts:
//  * rows (plain FP32 path): Z[n][a*L+j] in place -> D[n][a*L+j] = w_n (P - onehot)
//  * cols (BF16 path): Zt[a*Lp+j][n] (+ singles) -> Dt[a*Lp+j][n] = bf16(w_n P)
//    and, for two passes, the residual bf16(w_n P - hi) at column offset lo_off.
// Both work on a chunk of sequences; block partials of f and g1 are written
// at the chunk's block offset blk0 so that chunks fill one partial array.
// mode 0 computes f only (no D, no g1); mode 1 computes everything.
#include "common.h"

__global__ __launch_bounds__(256) void k_softmax_rows(float* __restrict__ ZD, long ldz,
                                                      const uint8_t* __restrict__ msap, int Lp,
                                                      const float* __restrict__ w, int L, int Nc,
                                                      int mode, int blk0, int nblk_j,
                                                      double* __restrict__ fpart,
                                                      double* __restrict__ g1part) {
  __shared__ float red[4][20][64];
  __shared__ double fred[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + lane;
  const int nb = blockIdx.y;
  double fl = 0.0;
  float g1[20];
#pragma unroll
  for (int a = 0; a < 20; a++) g1[a] = 0.f;
  if (j < L) {
    for (int i = 0; i < 16; i++) {
      const int n = nb * 64 + wv * 16 + i;
      if (n >= Nc) break;
      float* row = ZD + (long)n * ldz + j;
      float z[CCM_NA];
      float m = -INFINITY;
#pragma unroll
      for (int a = 0; a < CCM_NA; a++) {
        z[a] = row[(long)a * L];
        m = fmaxf(m, z[a]);
      }
      float s = 0.f;
#pragma unroll
      for (int a = 0; a < CCM_NA; a++) s += expf(z[a] - m);
      const float lse = m + logf(s);
      const int x = msap[(long)n * Lp + j];
      const float wn = w[n];
      float zx = 0.f;
#pragma unroll
      for (int a = 0; a < CCM_NA; a++) zx = (a == x) ? z[a] : zx;
      fl += (double)(wn * (lse - zx));
      if (mode) {
#pragma unroll
        for (int a = 0; a < CCM_NA; a++) {
          const float p = expf(z[a] - lse);
          const float d = wn * p - ((a == x) ? wn : 0.f);
          row[(long)a * L] = d;
          if (a < 20) g1[a] += d;
        }
      }
    }
  }
  if (mode) {
#pragma unroll
    for (int a = 0; a < 20; a++) red[wv][a][lane] = g1[a];
  }
  fl = ccm_wave_sum_f64(fl);
  if (lane == 0) fred[wv] = fl;
  __syncthreads();
  if (wv == 0) {
    const long gb = (long)blk0 + nb;
    if (mode && j < L) {
#pragma unroll
      for (int a = 0; a < 20; a++) {
        double t = (double)red[0][a][lane] + (double)red[1][a][lane] + (double)red[2][a][lane] +
                   (double)red[3][a][lane];
        g1part[(gb * 20 + a) * L + j] = t;
      }
    }
    if (lane == 0) fpart[gb * nblk_j + blockIdx.x] = fred[0] + fred[1] + fred[2] + fred[3];
  }
}

#define SC_SEQ 512  // sequences per block in the transposed variant

// One block: column j (blockIdx.y) for 512 consecutive sequences; thread t
// takes the adjacent pair n, n+1 (n = 512 bx + 2t), so that its logits are
// 8-byte loads and its BF16 results 4-byte stores, and accumulates its g1
// terms before a single wave reduction per residue.
__device__ __forceinline__ void softmax_cols_block(const float* __restrict__ Zt, long ldzt,
                                                   const float* __restrict__ x1,
                                                   const uint8_t* __restrict__ msapt, long ldmt,
                                                   const float* __restrict__ w, int L, int Lp, int Nc,
                                                   int ncols, uint16_t* __restrict__ Dt, long lddt,
                                                   long lo_off, int passes, int mode, int blk0,
                                                   double* __restrict__ fpart,
                                                   double* __restrict__ g1part, int bx, int j) {
  __shared__ float g1red[4][20];
  __shared__ double fred[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = bx * SC_SEQ + 2 * threadIdx.x;
  float g1[20];
#pragma unroll
  for (int a = 0; a < 20; a++) g1[a] = 0.f;
  double fl = 0.0;
  if (n < ncols) {
    // both sequences' logits (singles added), loads first
    float2 z[CCM_NA];
    const float2* zc = reinterpret_cast<const float2*>(Zt + (long)j * ldzt + n);
    const long zs = (long)Lp * ldzt / 2;
#pragma unroll
    for (int a = 0; a < CCM_NA; a++) z[a] = zc[a * zs];
    const bool ok0 = n < Nc, ok1 = n + 1 < Nc;
    const int x0 = ok0 ? msapt[(long)j * ldmt + n] : 255;
    const int x1v = ok1 ? msapt[(long)j * ldmt + n + 1] : 255;
    const float w0 = ok0 ? w[n] : 0.f, w1 = ok1 ? w[n + 1] : 0.f;
#pragma unroll
    for (int a = 0; a < 20; a++) {
      const float v = x1[(long)a * L + j];
      z[a].x += v;
      z[a].y += v;
    }
    // padding sequences past Nc: finite placeholder logits, weight 0
#pragma unroll
    for (int a = 0; a < CCM_NA; a++) {
      z[a].x = ok0 ? z[a].x : 0.f;
      z[a].y = ok1 ? z[a].y : 0.f;
    }
    float m0 = -INFINITY, m1 = -INFINITY, zx0 = 0.f, zx1 = 0.f;
#pragma unroll
    for (int a = 0; a < CCM_NA; a++) {
      m0 = fmaxf(m0, z[a].x);
      m1 = fmaxf(m1, z[a].y);
      zx0 = (a == x0) ? z[a].x : zx0;
      zx1 = (a == x1v) ? z[a].y : zx1;
    }
    // e_a = exp(z_a - m) once; P_a = e_a / sum.  The hardware exponential
    // and logarithm (relative error ~1e-6 over the arguments that matter)
    // are far inside the BF16 rounding of P; the FP32 path keeps the
    // library functions.
    float s0 = 0.f, s1 = 0.f;
#pragma unroll
    for (int a = 0; a < CCM_NA; a++) {
      z[a].x = __expf(z[a].x - m0);
      z[a].y = __expf(z[a].y - m1);
      s0 += z[a].x;
      s1 += z[a].y;
    }
    if (ok0) fl += (double)(w0 * (m0 + __logf(s0) - zx0));
    if (ok1) fl += (double)(w1 * (m1 + __logf(s1) - zx1));
    if (mode) {
      // padded sequences (weight 0) store zeros
      const float ws0 = w0 / s0, ws1 = w1 / s1;
      uint32_t* dcol = reinterpret_cast<uint32_t*>(Dt + (long)j * lddt + n);
      const long ds = (long)Lp * lddt / 2;
#pragma unroll
      for (int a = 0; a < CCM_NA; a++) {
        const float q0 = z[a].x * ws0, q1 = z[a].y * ws1;
        const uint32_t h = ccm_f2bf2(q0, q1);
        dcol[a * ds] = h;
        if (passes > 1)
          // residual of the exact product z * (w / s): an explicit fused
          // multiply-add, so that no contraction choice of the compiler
          // changes it (the fused forward's epilogue forms it the same way)
          dcol[a * ds + lo_off / 2] = ccm_f2bf2(__builtin_fmaf(z[a].x, ws0, -ccm_bf2f((uint16_t)h)),
                                                __builtin_fmaf(z[a].y, ws1, -ccm_bf2f((uint16_t)(h >> 16))));
        if (a < 20) g1[a] += (q0 - ((a == x0) ? w0 : 0.f)) + (q1 - ((a == x1v) ? w1 : 0.f));
      }
    }
  }
  if (mode) {
#pragma unroll
    for (int a = 0; a < 20; a++) {
      const float t = ccm_wave_sum_f32(g1[a]);
      if (lane == 0) g1red[wv][a] = t;
    }
  }
  fl = ccm_wave_sum_f64(fl);
  if (lane == 0) fred[wv] = fl;
  __syncthreads();
  const long gb = (long)blk0 + bx;
  if (mode && threadIdx.x < 20) {
    const int a = threadIdx.x;
    g1part[(gb * 20 + a) * L + j] = (double)g1red[0][a] + (double)g1red[1][a] + (double)g1red[2][a] +
                                    (double)g1red[3][a];
  }
  if (threadIdx.x == 0) fpart[gb * L + j] = fred[0] + fred[1] + fred[2] + fred[3];
}

__global__ __launch_bounds__(256) void k_softmax_cols(const float* __restrict__ Zt, long ldzt,
                                                      const float* __restrict__ x1,
                                                      const uint8_t* __restrict__ msapt, long ldmt,
                                                      const float* __restrict__ w, int L, int Lp,
                                                      int Nc, int ncols, uint16_t* __restrict__ Dt,
                                                      long lddt, long lo_off, int passes, int mode,
                                                      int blk0, double* __restrict__ fpart,
                                                      double* __restrict__ g1part) {
  softmax_cols_block(Zt, ldzt, x1, msapt, ldmt, w, L, Lp, Nc, ncols, Dt, lddt, lo_off, passes, mode, blk0,
                     fpart, g1part, blockIdx.x, blockIdx.y);
}

// every alignment of a group, whole alignment as one chunk: nblk(N) * L blocks
__global__ __launch_bounds__(256) void k_softmax_cols_grouped(const ccm_bprob* __restrict__ probs,
                                                              const uint32_t* __restrict__ offs,
                                                              const int32_t* __restrict__ pidx, int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  const int nb = (B.ncols + SC_SEQ - 1) / SC_SEQ;
  const int t = (int)(blockIdx.x - offs[gi]);
  softmax_cols_block(B.zt, B.ncols, B.x, B.msapt, B.ldmt, B.w, B.L, B.Lp, B.N, B.ncols, B.dt, B.lddt,
                     B.ncols, B.passes, 1, 0, B.fpart, B.g1part, t % nb, t / nb);
}

extern "C" {

// partial-array geometry over the whole alignment (chunks are multiples of 256)
int ccm_softmax_rows_blocks(int L, int N) { return ((L + 63) / 64) * ((N + 63) / 64); }
int ccm_softmax_rows_nblk(int N) { return (N + 63) / 64; }
int ccm_softmax_cols_blocks(int L, int N) { return ((N + SC_SEQ - 1) / SC_SEQ) * L; }
int ccm_softmax_cols_nblk(int N) { return (N + SC_SEQ - 1) / SC_SEQ; }
int ccm_softmax_cols_granule(void) { return SC_SEQ; }

// rows layout on sequences [n0, n0+Nc) of the alignment; pointers are
// already offset to the chunk, n0 must be a multiple of 64
int ccm_launch_softmax_rows(float* d_zd, long ldz, const uint8_t* d_msap, int Lp, const float* d_w,
                            int L, int Nc, int n0, int mode, double* d_fpart, double* d_g1part,
                            hipStream_t s) {
  if (L <= 0 || Nc <= 0 || (n0 % 64)) return CCM_ERR_ARG;
  const int nbj = (L + 63) / 64;
  hipLaunchKernelGGL(k_softmax_rows, dim3(nbj, (Nc + 63) / 64), dim3(256), 0, s, d_zd, ldz, d_msap, Lp,
                     d_w, L, Nc, mode, n0 / 64, nbj, d_fpart, d_g1part);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// cols layout on a chunk of Nc sequences starting at n0 (a multiple of
// ccm_softmax_cols_granule()); ncols: padded chunk width of Zt/Dt; ldmt: row
// stride of the transposed MSA
int ccm_launch_softmax_cols(const float* d_zt, long ldzt, const float* d_x1, const uint8_t* d_msapt,
                            long ldmt, const float* d_w, int L, int Lp, int Nc, int ncols, int n0,
                            uint16_t* d_dt, long lddt, long lo_off, int passes, int mode,
                            double* d_fpart, double* d_g1part, hipStream_t s) {
  if (L <= 0 || Nc <= 0 || ncols < Nc || (n0 % SC_SEQ) || (ncols & 1) || (ldzt & 1) || (lddt & 1) ||
      (lo_off & 1))
    return CCM_ERR_ARG;
  const int nb = (ncols + SC_SEQ - 1) / SC_SEQ;
  hipLaunchKernelGGL(k_softmax_cols, dim3(nb, L), dim3(256), 0, s, d_zt, ldzt, d_x1, d_msapt, ldmt, d_w, L,
                     Lp, Nc, ncols, d_dt, lddt, lo_off, passes, mode, n0 / SC_SEQ, d_fpart,
                     d_g1part);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_softmax_cols_grouped(const ccm_bprob* d_probs, ccm_group g, hipStream_t s) {
  if (g.np <= 0 || g.total == 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_softmax_cols_grouped, dim3(g.total), dim3(256), 0, s, d_probs, g.offs, g.pidx, g.np);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"
