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
//    and, for two passes, the residual bf16(w_n P - hi) at column offset Np.
#include "common.h"

__global__ __launch_bounds__(256) void k_softmax_rows(float* __restrict__ ZD, long ldz,
                                                      const uint8_t* __restrict__ msap, int Lp,
                                                      const float* __restrict__ w, int L, int N,
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
      if (n >= N) break;
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
#pragma unroll
      for (int a = 0; a < CCM_NA; a++) {
        const float p = expf(z[a] - lse);
        const float d = wn * p - ((a == x) ? wn : 0.f);
        row[(long)a * L] = d;
        if (a < 20) g1[a] += d;
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 20; a++) red[wv][a][lane] = g1[a];
  fl = ccm_wave_sum_f64(fl);
  if (lane == 0) fred[wv] = fl;
  __syncthreads();
  if (wv == 0) {
    if (j < L) {
#pragma unroll
      for (int a = 0; a < 20; a++) {
        double t = (double)red[0][a][lane] + (double)red[1][a][lane] + (double)red[2][a][lane] +
                   (double)red[3][a][lane];
        g1part[((long)nb * 20 + a) * L + j] = t;
      }
    }
    if (lane == 0) fpart[(long)nb * gridDim.x + blockIdx.x] = fred[0] + fred[1] + fred[2] + fred[3];
  }
}

#define SC_JT 8  // columns j per block in the transposed variant

__global__ __launch_bounds__(256) void k_softmax_cols(const float* __restrict__ Zt, long ldzt,
                                                      const float* __restrict__ x1,
                                                      const uint8_t* __restrict__ msapt, int Np,
                                                      const float* __restrict__ w, int L, int Lp, int N,
                                                      uint16_t* __restrict__ Dt, long lddt, int passes,
                                                      double* __restrict__ fpart,
                                                      double* __restrict__ g1part) {
  __shared__ double g1red[4][SC_JT][20];
  __shared__ double fred[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int n = blockIdx.x * 256 + threadIdx.x;
  const bool live = n < N;
  const float wn = live ? w[n] : 0.f;
  double fl = 0.0;
  for (int jj = 0; jj < SC_JT; jj++) {
    const int j = blockIdx.y * SC_JT + jj;
    if (j >= L) break;
    float z[CCM_NA];
    float m = -INFINITY;
#pragma unroll
    for (int a = 0; a < CCM_NA; a++) {
      float v = live ? Zt[(long)(a * Lp + j) * ldzt + n] : 0.f;
      if (a < 20) v += x1[(long)a * L + j];
      z[a] = v;
      m = fmaxf(m, v);
    }
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < CCM_NA; a++) s += expf(z[a] - m);
    const float lse = m + logf(s);
    const int x = live ? msapt[(long)j * Np + n] : 255;
    float zx = 0.f;
#pragma unroll
    for (int a = 0; a < CCM_NA; a++) zx = (a == x) ? z[a] : zx;
    if (live) fl += (double)(wn * (lse - zx));
#pragma unroll
    for (int a = 0; a < CCM_NA; a++) {
      const float p = expf(z[a] - lse);
      const float qv = wn * p;
      if (live) {
        const uint16_t hi = ccm_f2bf(qv);
        uint16_t* dst = Dt + (long)(a * Lp + j) * lddt + n;
        dst[0] = hi;
        if (passes > 1) dst[Np] = ccm_f2bf(qv - ccm_bf2f(hi));
      }
      if (a < 20) {
        const float d = live ? (qv - ((a == x) ? wn : 0.f)) : 0.f;
        const double t = ccm_wave_sum_f64((double)d);
        if (lane == 0) g1red[wv][jj][a] = t;
      }
    }
  }
  fl = ccm_wave_sum_f64(fl);
  if (lane == 0) fred[wv] = fl;
  __syncthreads();
  const int t = threadIdx.x;
  if (t < SC_JT * 20) {
    const int jj = t / 20, a = t % 20;
    const int j = blockIdx.y * SC_JT + jj;
    if (j < L) {
      const double v = g1red[0][jj][a] + g1red[1][jj][a] + g1red[2][jj][a] + g1red[3][jj][a];
      g1part[((long)blockIdx.x * 20 + a) * L + j] = v;
    }
  }
  if (t == 0) fpart[(long)blockIdx.x * gridDim.y + blockIdx.y] = fred[0] + fred[1] + fred[2] + fred[3];
}

extern "C" {

int ccm_softmax_rows_blocks(int L, int N) { return ((L + 63) / 64) * ((N + 63) / 64); }
int ccm_softmax_rows_nblk(int N) { return (N + 63) / 64; }
int ccm_softmax_cols_blocks(int L, int N) { return ((N + 255) / 256) * ((L + SC_JT - 1) / SC_JT); }
int ccm_softmax_cols_nblk(int N) { return (N + 255) / 256; }

int ccm_launch_softmax_rows(float* d_zd, long ldz, const uint8_t* d_msap, int Lp, const float* d_w,
                            int L, int N, double* d_fpart, double* d_g1part, hipStream_t s) {
  if (L <= 0 || N <= 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_softmax_rows, dim3((L + 63) / 64, (N + 63) / 64), dim3(256), 0, s, d_zd, ldz,
                     d_msap, Lp, d_w, L, N, d_fpart, d_g1part);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_softmax_cols(const float* d_zt, long ldzt, const float* d_x1, const uint8_t* d_msapt,
                            int Np, const float* d_w, int L, int Lp, int N, uint16_t* d_dt, long lddt,
                            int passes, double* d_fpart, double* d_g1part, hipStream_t s) {
  if (L <= 0 || N <= 0 || Np < N) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_softmax_cols, dim3((N + 255) / 256, (L + SC_JT - 1) / SC_JT), dim3(256), 0, s,
                     d_zt, ldzt, d_x1, d_msapt, Np, d_w, L, Lp, N, d_dt, lddt, passes, d_fpart,
                     d_g1part);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"
