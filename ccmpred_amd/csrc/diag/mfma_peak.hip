// Measured MFMA throughput on this GPU: register-only loops of the two
// instructions the PLL contractions issue, v_mfma_f32_32x32x16_bf16 (dense,
// the gradient) and v_smfmac_f32_32x32x32_bf16 (2:4 structured-sparse, the
// forward), with operands like the contractions' (one dense operand of
// small bf16 values, one 0/1 operand) and with every bit of both operands
// varying.  No memory is read in the loop; each wave writes one float at the
// end so that no MFMA is dead.  Prints one JSON line: TFLOP/s per
// instruction and operand kind (dense-equivalent flops: 2*32*32*16 per
// dense and 2*32*32*32 per sparse instruction) at 8 waves per SIMD, and
// the same at 1, 2 and 4 waves per SIMD (keys ending _1w, _2w, _4w).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdint>

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x16_t __attribute__((ext_vector_type(16)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int ACC = 8;        // independent accumulators per wave
constexpr int ITERS = 2048;   // loop trips; ACC instructions per trip

__device__ __forceinline__ uint32_t mix(uint32_t x) {
  x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
  return x;
}

// kind 0: a = small bf16 values, b = bf16 0/1; kind 1: every bit random
// (exponents kept finite)
__device__ __forceinline__ uint4 operand(uint32_t seed, int kind, bool onehot) {
  uint4 v;
  uint32_t* p = reinterpret_cast<uint32_t*>(&v);
  for (int i = 0; i < 4; i++) {
    const uint32_t r = mix(seed * 4 + i);
    if (kind == 0)
      p[i] = onehot ? ((r & 7u) == 0 ? 0x3F80u : 0u) | ((r & 0x70u) == 0 ? 0x3F800000u : 0u)
                    : (r & 0x007F007Fu) | 0x3C003C00u;
    else
      p[i] = r & 0xBFFFBFFFu;
  }
  return v;
}

__global__ __launch_bounds__(256) void k_peak_dense(int kind, float* out) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  const bf16x8_t a = __builtin_bit_cast(bf16x8_t, operand(t, kind, false));
  const bf16x8_t b = __builtin_bit_cast(bf16x8_t, operand(t + 0x9e3779b9u, kind, true));
  floatx16 acc[ACC];
  for (int j = 0; j < ACC; j++) acc[j] = (floatx16)0.0f;
  for (int it = 0; it < ITERS; it++)
#pragma unroll
    for (int j = 0; j < ACC; j++) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[j], 0, 0, 0);
  float s = 0.f;
  for (int j = 0; j < ACC; j++) s += acc[j][0];
  if ((threadIdx.x & 63) == 0) out[t >> 6] = s;
}

__global__ __launch_bounds__(256) void k_peak_sparse(int kind, float* out) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  const bf16x8_t a = __builtin_bit_cast(bf16x8_t, operand(t, kind, kind == 0));
  const uint4 b0 = operand(t + 0x9e3779b9u, kind, false), b1 = operand(t + 0x7f4a7c15u, kind, false);
  struct { uint4 lo, hi; } bb{b0, b1};
  const bf16x16_t b = __builtin_bit_cast(bf16x16_t, bb);
  const int idx = (int)(mix(t) & 0xEEEEu);  // per value pair: positions 2 and 3 or 0 and 1... (any)
  floatx16 acc[ACC];
  for (int j = 0; j < ACC; j++) acc[j] = (floatx16)0.0f;
  for (int it = 0; it < ITERS; it++)
#pragma unroll
    for (int j = 0; j < ACC; j++) acc[j] = __builtin_amdgcn_smfmac_f32_32x32x32_bf16(a, b, acc[j], idx, 0, 0);
  float s = 0.f;
  for (int j = 0; j < ACC; j++) s += acc[j][0];
  if ((threadIdx.x & 63) == 0) out[t >> 6] = s;
}

#define CK(x)                                                                        \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));                 \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

int main() {
  int dev = 0, cus = 0;
  CK(hipGetDevice(&dev));
  CK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  float* out = nullptr;
  CK(hipMalloc(&out, sizeof(float) * (size_t)cus * 8 * 4));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const char* names[2] = {"contraction_like", "random"};
  // the clocks rise over the first second or so of load: 1.5 s of the dense
  // loop before anything is timed (one warm-up launch per measurement is
  // not enough; the first configuration measured would read low)
  {
    float spent = 0.f;
    while (spent < 1500.f) {
      CK(hipEventRecord(e0, 0));
      hipLaunchKernelGGL(k_peak_dense, dim3(cus * 8), dim3(256), 0, 0, 0, out);
      CK(hipGetLastError());
      CK(hipEventRecord(e1, 0));
      CK(hipEventSynchronize(e1));
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      spent += ms;
    }
  }
  // waves per SIMD: 8 (the headline figures), then 1, 2 and 4 (the
  // contraction kernels run two per SIMD)
  printf("{\"cus\": %d", cus);
  const int wps_list[4] = {8, 1, 2, 4};
  for (int w = 0; w < 4; w++) {
  const int wps = wps_list[w];
  const int blocks = cus * wps;  // wps workgroups of 4 waves per CU
  const double waves = (double)blocks * 4;
  for (int sparse = 0; sparse < 2; sparse++)
    for (int kind = 0; kind < 2; kind++) {
      float best = 1e30f;
      for (int rep = 0; rep < 5; rep++) {
        CK(hipEventRecord(e0, 0));
        if (sparse)
          hipLaunchKernelGGL(k_peak_sparse, dim3(blocks), dim3(256), 0, 0, kind, out);
        else
          hipLaunchKernelGGL(k_peak_dense, dim3(blocks), dim3(256), 0, 0, kind, out);
        CK(hipGetLastError());
        CK(hipEventRecord(e1, 0));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0 && ms < best) best = ms;
      }
      const double flops = waves * ITERS * ACC * (sparse ? 65536.0 : 32768.0);
      if (wps == 8)
        printf(", \"%s_%s_tflops\": %.1f", sparse ? "smfmac_32x32x32_bf16" : "mfma_32x32x16_bf16", names[kind],
               flops / (best * 1e-3) / 1e12);
      else
        printf(", \"%s_%s_tflops_%dw\": %.1f", sparse ? "smfmac_32x32x32_bf16" : "mfma_32x32x16_bf16",
               names[kind], wps, flops / (best * 1e-3) / 1e12);
    }
  }
  printf("}\n");
  CK(hipFree(out));
  return 0;
}
