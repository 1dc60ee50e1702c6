// Contact scores from the couplings on the device (util.c:16-139): the
// Frobenius norm of each mean-centred 21x21 coupling block, the average
// product correction and the optional [0,1] rescaling.  Every value is formed
// with the same float operations, in the same order, as the host functions
// ccm_sum_submatrices / ccm_apc / ccm_normalize, with multiply-add fusion off
// (this file is built with -ffp-contract=off), so the device map equals the
// host map bit for bit; only order-free steps
// (minimum and maximum) are reduced in parallel.
#include "common.h"

#include <mutex>
#include <stdio.h>

#include "ccmpred_amd_kernels.h"

namespace {

constexpr int SS_NT = 256;  // sum_submatrices: threads per block (over j)
constexpr int AP_NT = 256;  // APC: threads for the column means and the sweeps
constexpr int AP_TOT = 64;  // one extra wave whose lane 0 forms the ordered total

// One thread per (k, j).  W(b,k,a,j) = x2[((b*L + k)*32 + a)*L + j], so the
// threads of a wave (consecutive j) read one contiguous segment per (a, b).
__global__ void __launch_bounds__(SS_NT) k_sum_submatrices(const float* __restrict__ x2, int L,
                                                           float* __restrict__ out) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * SS_NT + threadIdx.x;
  const int k = blockIdx.y;
  if (j >= L) return;
  const long Ll = L;
  float mean = 0.f;
  for (int a = 0; a < CCM_NA; a++)
    for (int b = 0; b < CCM_NA; b++) mean += x2[((b * Ll + k) * CCM_NAP + a) * Ll + j];
  mean /= (float)(CCM_NA * CCM_NA);
  float acc = 0.f;
  for (int a = 0; a < CCM_NA - 1; a++)
    for (int b = 0; b < CCM_NA - 1; b++) {
      const float d = x2[((b * Ll + k) * CCM_NAP + a) * Ll + j] - mean;
      acc += d * d;
    }
  out[(long)k * Ll + j] = k == j ? 0.f : (float)__builtin_sqrt((double)acc);
}

constexpr int AP_ALL = AP_NT + AP_TOT;

// Block-wide reduction of (lo, hi) with the host's strict comparisons; every
// thread of the block calls it (the total's wave with neutral values).
__device__ void block_minmax(float& lo, float& hi, float* red) {
  for (int off = 32; off > 0; off >>= 1) {
    const float l2 = __shfl_down(lo, off, 64), h2 = __shfl_down(hi, off, 64);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[2 * wid] = lo;
    red[2 * wid + 1] = hi;
  }
  __syncthreads();
  lo = red[0];
  hi = red[1];
  for (int w = 1; w < AP_ALL / 64; w++) {
    lo = red[2 * w] < lo ? red[2 * w] : lo;
    hi = red[2 * w + 1] > hi ? red[2 * w + 1] : hi;
  }
}

// One block per map.  Column means (one thread per column, rows in order) and
// the total (one lane of an extra wave, row-major order) as the host forms
// them; then the correction, the minimum over i < j, the shift and, if asked,
// the rescaling.  colmean: L floats of scratch.
__global__ void __launch_bounds__(AP_ALL) k_apc(float* __restrict__ mat, int L, int apc,
                                                int normalize, float* __restrict__ colmean) {
#pragma clang fp contract(off)
  __shared__ float red[2 * (AP_ALL / 64)];
  __shared__ float s_total;
  const long Ll = L, n = Ll * Ll;
  const int t = threadIdx.x;
  const bool sweep = t < AP_NT;  // the total's wave takes no part in the sweeps
  if (apc) {
    if (sweep) {
      for (int j = t; j < L; j += AP_NT) {
        float m = 0.f;
        for (long i = 0; i < Ll; i++) m += mat[i * Ll + j] / (float)L;
        colmean[j] = m;
      }
    } else if (t == AP_NT) {
      float total = 0.f;
      for (long e = 0; e < n; e++) total += mat[e];
      s_total = total / (float)n;
    }
    __syncthreads();
    const float total = s_total;
    float lo = __builtin_inff(), hi = -__builtin_inff();
    if (sweep) {
      for (long i = 0; i < Ll; i++)
        for (long j = t; j < Ll; j += AP_NT) {
          const float v = mat[i * Ll + j] - (colmean[i] * colmean[j]) / total;
          mat[i * Ll + j] = v;
          if (j > i && v < lo) lo = v;
        }
    }
    block_minmax(lo, hi, red);
    if (sweep) {
      for (long i = 0; i < Ll; i++)
        for (long j = t; j < Ll; j += AP_NT) mat[i * Ll + j] = i == j ? 0.f : mat[i * Ll + j] - lo;
    }
    __syncthreads();
  }
  if (normalize && L > 1) {
    float lo = mat[1], hi = mat[1];
    if (sweep) {
      for (long i = 0; i < Ll; i++)
        for (long j = t; j < Ll; j += AP_NT) {
          if (i == j) continue;
          const float v = mat[i * Ll + j];
          if (v < lo) lo = v;
          if (v > hi) hi = v;
        }
    }
    block_minmax(lo, hi, red);
    const float range = hi - lo;
    if (sweep) {
      for (long i = 0; i < Ll; i++)
        for (long j = t; j < Ll; j += AP_NT)
          mat[i * Ll + j] = i == j ? 0.f : (mat[i * Ll + j] - lo) / range;
    }
  }
}

}  // namespace

extern "C" int ccm_contact_map_work_len(int L) { return L < 1 ? 0 : L; }

extern "C" int ccm_launch_contact_map(const float* d_x, int L, int apc, int normalize, float* d_out,
                                      float* d_work, ccm_stream_t s) {
  if (!d_x || !d_out || L < 1 || (apc && !d_work) || (long)L * L > (1L << 31)) return CCM_ERR_ARG;
  hipStream_t st = (hipStream_t)s;
  const long nsp = 20L * L + 32 - (20L * L) % 32;
  dim3 grid((L + SS_NT - 1) / SS_NT, L);
  hipLaunchKernelGGL(k_sum_submatrices, grid, dim3(SS_NT), 0, st, d_x + nsp, L, d_out);
  CCM_LAUNCH_CHECK();
  if (apc || normalize) {
    hipLaunchKernelGGL(k_apc, dim3(1), dim3(AP_ALL), 0, st, d_out, L, apc, normalize,
                       d_work);
    CCM_LAUNCH_CHECK();
  }
  return CCM_OK;
}

// ---- allocation-free form --------------------------------------------------
// The APC's column-mean scratch (ccm_contact_map_work_len(L) = L floats) in a
// fixed device array instead of a caller's or the library's allocation.  Each
// GPU holds its own copy of the array until the process ends; its address is
// looked up once per device (with that device current) and kept.  Every call
// on one GPU shares that GPU's array: the kernels write each element before
// they read it, so nothing needs clearing, but calls that overlap on one GPU
// (two streams, two threads) race on it.
namespace {
constexpr int NOALLOC_MAX_L = CCM_CONTACT_MAP_NOALLOC_MAX_L;
constexpr int NOALLOC_MAX_DEV = 64;
__device__ __align__(256) float scratch[NOALLOC_MAX_L];

std::mutex scratch_mu;
float* scratch_ptr[NOALLOC_MAX_DEV];  // device address of `scratch`, per device

// device address of `scratch` on the current device (nullptr on error)
float* scratch_on_current_device(int dev) {
  if (dev < 0 || dev >= NOALLOC_MAX_DEV) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    scratch_ptr[dev] = static_cast<float*>(p);
  }
  return scratch_ptr[dev];
}
}  // namespace

extern "C" int ccm_contact_map_noalloc_max_l(void) { return NOALLOC_MAX_L; }

extern "C" int ccm_launch_contact_map_noalloc(const float* d_x, int L, int apc, int normalize,
                                              float* d_out, ccm_stream_t s) {
  if (L > NOALLOC_MAX_L) {
    fprintf(stderr,
            "ccm_launch_contact_map_noalloc: L = %d exceeds the fixed scratch capacity of %d "
            "columns (use ccm_launch_contact_map with a work buffer)\n",
            L, NOALLOC_MAX_L);
    return CCM_ERR_SIZE;
  }
  if (!d_x || !d_out || L < 1) return CCM_ERR_ARG;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return CCM_ERR_LAUNCH;
  if (s) {
    // the scratch address is per device: the stream must be the current device's
    hipDevice_t sdev = -1;
    if (hipStreamGetDevice((hipStream_t)s, &sdev) != hipSuccess || (int)sdev != dev)
      return CCM_ERR_ARG;
  }
  float* work = scratch_on_current_device(dev);
  if (!work) return CCM_ERR_LAUNCH;
  return ccm_launch_contact_map(d_x, L, apc, normalize, d_out, work, s);
}
