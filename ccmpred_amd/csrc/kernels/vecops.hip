// Device vector arithmetic of the conjugate-gradient loop.  All dot
// products accumulate FP32 products in FP64 with a fixed reduction order
// (per-block partials, then one block sums the partials), so repeated runs
// give bitwise-identical scalars.
#include "common.h"

#define VO_BLOCKS 1024

// up to three dot products sharing the first operand: out[i] = <a, b_i>
__global__ __launch_bounds__(256) void k_dots_partial(const float* __restrict__ a,
                                                      const float* __restrict__ b0,
                                                      const float* __restrict__ b1,
                                                      const float* __restrict__ b2, long n, int nd,
                                                      double* __restrict__ part) {
  __shared__ double red[3][4];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  const long stride = (long)gridDim.x * 256;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    const double av = a[i];
    s0 += av * (double)b0[i];
    if (nd > 1) s1 += av * (double)b1[i];
    if (nd > 2) s2 += av * (double)b2[i];
  }
  s0 = ccm_wave_sum_f64(s0);
  s1 = ccm_wave_sum_f64(s1);
  s2 = ccm_wave_sum_f64(s2);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wv] = s0;
    red[1][wv] = s1;
    red[2][wv] = s2;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int d = threadIdx.x;
    part[(long)d * gridDim.x + blockIdx.x] = red[d][0] + red[d][1] + red[d][2] + red[d][3];
  }
}

__global__ __launch_bounds__(256) void k_dots_final(const double* __restrict__ part, int nblk, int nd,
                                                    double* __restrict__ out) {
  __shared__ double sh[256];
  for (int d = 0; d < nd; d++) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nblk; i += 256) s += part[(long)d * nblk + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int i = 0; i < 256; i++) t += sh[i];
      out[d] = t;
    }
    __syncthreads();
  }
}

// x += alpha * s
__global__ void k_axpy(float* __restrict__ x, const float* __restrict__ s, float alpha, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += alpha * s[i];
}

// s = -g + beta * s
__global__ void k_direction(float* __restrict__ s, const float* __restrict__ g, float beta, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = -g[i] + beta * s[i];
}

// s = -g
__global__ void k_negate(float* __restrict__ s, const float* __restrict__ g, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = -g[i];
}

extern "C" {

int ccm_dots_partial_len(void) { return 3 * VO_BLOCKS; }

// out[0..nd) = <a, b_i>; part: scratch of ccm_dots_partial_len() doubles
int ccm_launch_dots(const float* a, const float* b0, const float* b1, const float* b2, long n, int nd,
                    double* part, double* out, hipStream_t s) {
  if (nd < 1 || nd > 3 || n <= 0) return CCM_ERR_ARG;
  long want = (n + 255) / 256;
  const int nblk = (int)(want < VO_BLOCKS ? want : VO_BLOCKS);
  hipLaunchKernelGGL(k_dots_partial, dim3(nblk), dim3(256), 0, s, a, b0, b1 ? b1 : b0, b2 ? b2 : b0, n, nd,
                     part);
  CCM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dots_final, dim3(1), dim3(256), 0, s, part, nblk, nd, out);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_axpy(float* x, const float* sdir, float alpha, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, sdir, alpha, n);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_direction(float* sdir, const float* g, float beta, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_direction, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sdir, g, beta, n);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_negate(float* sdir, const float* g, long n, hipStream_t s) {
  hipLaunchKernelGGL(k_negate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, sdir, g, n);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"
