// Measured HBM streaming rates on this GPU: read-only, write-only and copy
// kernels over 1 GiB buffers (16-byte accesses, grid-stride loops), with
// plain and nontemporal accesses, at several grid sizes; the best rate of
// each kind over the grids.  The memory-bound steps of the PLL path (the
// coupling conversion, the gradient assembly) are measured against these.
// Prints one JSON line (TB/s; a copy counts its bytes read and written).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdint>

typedef float f4 __attribute__((ext_vector_type(4)));

template <int NT>
__global__ __launch_bounds__(256) void k_read(const f4* __restrict__ a, size_t n, float* out) {
  f4 s = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    s += NT ? __builtin_nontemporal_load(a + i) : a[i];
  if (s.x == 12345.f) out[0] = s.y;  // never true for the zero-filled input; keeps the loads
}

template <int NT>
__global__ __launch_bounds__(256) void k_write(f4* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const f4 v = (float)i;
    if (NT)
      __builtin_nontemporal_store(v, b + i);
    else
      b[i] = v;
  }
}

template <int NT>
__global__ __launch_bounds__(256) void k_copy(const f4* __restrict__ a, f4* __restrict__ b, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const f4 v = NT ? __builtin_nontemporal_load(a + i) : a[i];
    if (NT)
      __builtin_nontemporal_store(v, b + i);
    else
      b[i] = v;
  }
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
  const size_t bytes = (size_t)1 << 30, n = bytes / sizeof(f4);
  f4 *a = nullptr, *b = nullptr;
  float* out = nullptr;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMalloc(&out, sizeof(float)));
  CK(hipMemset(a, 0, bytes));
  CK(hipMemset(b, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  // about a second of copies before anything is timed (clocks ramp up)
  {
    float spent = 0.f;
    while (spent < 1000.f) {
      CK(hipEventRecord(e0, 0));
      hipLaunchKernelGGL(k_copy<0>, dim3(2048), dim3(256), 0, 0, a, b, n);
      CK(hipGetLastError());
      CK(hipEventRecord(e1, 0));
      CK(hipEventSynchronize(e1));
      float ms = 0.f;
      CK(hipEventElapsedTime(&ms, e0, e1));
      spent += ms;
    }
  }
  const char* names[6] = {"read", "read_nt", "write", "write_nt", "copy", "copy_nt"};
  double best_rate[6] = {0, 0, 0, 0, 0, 0};
  const int grids[5] = {1024, 2048, 4096, 8192, 16384};
  for (int gi = 0; gi < 5; gi++)
    for (int kind = 0; kind < 6; kind++) {
      const dim3 g(grids[gi]), t(256);
      float best = 1e30f;
      for (int rep = 0; rep < 6; rep++) {
        CK(hipEventRecord(e0, 0));
        switch (kind) {
          case 0: hipLaunchKernelGGL(k_read<0>, g, t, 0, 0, a, n, out); break;
          case 1: hipLaunchKernelGGL(k_read<1>, g, t, 0, 0, a, n, out); break;
          case 2: hipLaunchKernelGGL(k_write<0>, g, t, 0, 0, b, n); break;
          case 3: hipLaunchKernelGGL(k_write<1>, g, t, 0, 0, b, n); break;
          case 4: hipLaunchKernelGGL(k_copy<0>, g, t, 0, 0, a, b, n); break;
          default: hipLaunchKernelGGL(k_copy<1>, g, t, 0, 0, a, b, n); break;
        }
        CK(hipGetLastError());
        CK(hipEventRecord(e1, 0));
        CK(hipEventSynchronize(e1));
        float ms = 0.f;
        CK(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0 && ms < best) best = ms;
      }
      const double rate = (kind >= 4 ? 2.0 : 1.0) * (double)bytes / (best * 1e-3) / 1e12;
      if (rate > best_rate[kind]) best_rate[kind] = rate;
    }
  printf("{\"bytes\": %zu", bytes);
  for (int k = 0; k < 6; k++) printf(", \"%s_tbs\": %.2f", names[k], best_rate[k]);
  printf("}\n");
  CK(hipFree(a));
  CK(hipFree(b));
  CK(hipFree(out));
  return 0;
}
