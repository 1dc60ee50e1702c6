// Shared device helpers for the gfx950 kernels of the PLL path.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define CCM_NA 21   // 20 amino acids + gap (gap index 20)
#define CCM_NAP 32  // padded alphabet of the reference's default PADDING build

// Status codes returned by every launcher (0 = success).
#define CCM_OK 0
#define CCM_ERR_ARG -2
#define CCM_ERR_LAUNCH -3
#define CCM_ERR_SIZE -5  // beyond the fixed capacity of an allocation-free entry point

#define CCM_LAUNCH_CHECK()                                  \
  do {                                                      \
    hipError_t e_ = hipGetLastError();                      \
    if (e_ != hipSuccess) return CCM_ERR_LAUNCH;            \
  } while (0)

// 8 XCDs on MI355X: consecutive workgroup ids are dealt round-robin to the
// XCDs.  Remap a hardware id so that the workgroups one XCD runs back to back
// get consecutive logical ids (and therefore share operand tiles in that
// XCD's L2).  `total` must be a multiple of 8; the grid is padded to it.
__device__ __forceinline__ uint32_t ccm_xcd_remap(uint32_t hw, uint32_t total) {
  const uint32_t per = total >> 3;
  return (hw & 7u) * per + (hw >> 3);
}

__device__ __forceinline__ double ccm_wave_sum_f64(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ float ccm_wave_sum_f32(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Wave-wide sums through DPP lane moves (VALU only, no LDS round trips as in
// the shuffles above); the total is valid in lane 63.  Fixed order: the two
// quad swaps, half-row and row mirrors (every lane then holds its row's
// sum), row 0 into row 1 and row 2 into row 3 (row_bcast:15), rows 0-1 into
// rows 2-3 (row_bcast:31).
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ float ccm_dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}

__device__ __forceinline__ float ccm_wave_sum_dpp_f32(float v) {
  v += ccm_dpp_f32<0xB1>(v);       // quad_perm [1,0,3,2]
  v += ccm_dpp_f32<0x4E>(v);       // quad_perm [2,3,0,1]
  v += ccm_dpp_f32<0x141>(v);      // row_half_mirror
  v += ccm_dpp_f32<0x140>(v);      // row_mirror
  v += ccm_dpp_f32<0x142, 0xA>(v); // row_bcast:15 into rows 1 and 3
  v += ccm_dpp_f32<0x143, 0xC>(v); // row_bcast:31 into rows 2 and 3
  return v;
}

// sum over each row of 16 lanes, held by every lane of the row
__device__ __forceinline__ float ccm_row_sum_dpp_f32(float v) {
  v += ccm_dpp_f32<0xB1>(v);
  v += ccm_dpp_f32<0x4E>(v);
  v += ccm_dpp_f32<0x141>(v);
  v += ccm_dpp_f32<0x140>(v);
  return v;
}

template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ double ccm_dpp_f64(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ __forceinline__ double ccm_wave_sum_dpp_f64(double v) {
  v += ccm_dpp_f64<0xB1>(v);
  v += ccm_dpp_f64<0x4E>(v);
  v += ccm_dpp_f64<0x141>(v);
  v += ccm_dpp_f64<0x140>(v);
  v += ccm_dpp_f64<0x142, 0xA>(v);
  v += ccm_dpp_f64<0x143, 0xC>(v);
  return v;
}

__device__ __forceinline__ double ccm_row_sum_dpp_f64(double v) {
  v += ccm_dpp_f64<0xB1>(v);
  v += ccm_dpp_f64<0x4E>(v);
  v += ccm_dpp_f64<0x141>(v);
  v += ccm_dpp_f64<0x140>(v);
  return v;
}

// Block-wide FP64 sum with a fixed reduction order; result valid in thread 0.
template <int NT>
__device__ __forceinline__ double ccm_block_sum_f64(double v, double* lds) {
  v = ccm_wave_sum_f64(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < NT / 64; i++) t += lds[i];
  }
  __syncthreads();
  return t;
}

// float -> bf16 (round to nearest even), as raw bits
// float -> bf16 (round to nearest even), as raw bits: the gfx950 conversion
// instruction v_cvt_pk_bf16_f32, one value or a packed pair (a in the low
// half)
typedef __bf16 ccm_bf16x2_t __attribute__((ext_vector_type(2)));
typedef float ccm_floatx2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t ccm_f2bf2(float a, float b) {
  const ccm_floatx2_t v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, ccm_bf16x2_t));
}
__device__ __forceinline__ uint16_t ccm_f2bf(float f) {
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}

__device__ __forceinline__ float ccm_bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

static inline long ccm_nsp(int L) {
  long ns = (long)L * (CCM_NA - 1);
  return ns + CCM_NAP - (ns % CCM_NAP);
}

// ---- grouped launches over a batch of alignments ----
#include "ccmpred_amd_batch.h"

// index i of the group entry whose block range [offs[i], offs[i+1]) holds
// blk (offs ascending, offs[0] = 0, blk < offs[np]); uniform per block
__device__ __forceinline__ int ccm_group_find(const uint32_t* offs, int np, uint32_t blk) {
  int lo = 0, hi = np;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offs[mid] <= blk)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}
