// Sequence reweighting as an exact integer kernel.
//
// Semantics follow reweighting.c:9-58: with idthres = (int)ceil(theta*L)
// (a float product), count_n = #{m != n : identities(n, m) > idthres};
// the reference's pair loop also visits the diagonal pair (n, n) and adds 2
// to the float accumulator when L > idthres; the weight is
// (float)(1. / (acc - 1)) with the division in double.
//
// Each MSA row is packed into five bit planes (bits 0..4 of the residue
// index), 32 columns per word.  Two residues differ iff any plane bit
// differs, so identities(n, m) = L - popcount(OR_p (plane_p(n) ^ plane_p(m)))
// summed over words.  Bits past L are zero in every row and never differ.
#include "common.h"

#define RW_TILE 64   // m rows per LDS tile
#define RW_WCH 8     // words per chunk

__global__ __launch_bounds__(256) void k_pack_planes(const uint8_t* __restrict__ msa, int L, int N,
                                                     int Wp, uint32_t* __restrict__ planes) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)N * Wp) return;
  const int n = (int)(t / Wp), w = (int)(t % Wp);
  uint32_t p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0;
  const uint8_t* row = msa + (long)n * L;
  for (int i = 0; i < 32; i++) {
    const int k = w * 32 + i;
    if (k < L) {
      const uint32_t v = row[k];
      p0 |= ((v >> 0) & 1u) << i;
      p1 |= ((v >> 1) & 1u) << i;
      p2 |= ((v >> 2) & 1u) << i;
      p3 |= ((v >> 3) & 1u) << i;
      p4 |= ((v >> 4) & 1u) << i;
    }
  }
  uint32_t* dst = planes + (long)n * 5 * Wp + w;
  dst[0] = p0;
  dst[Wp] = p1;
  dst[2 * Wp] = p2;
  dst[3 * Wp] = p3;
  dst[4 * Wp] = p4;
}

// grid.x: blocks of 256 rows n; grid.y: slices of the m range.
__global__ __launch_bounds__(256) void k_identity_counts(const uint32_t* __restrict__ planes, int L,
                                                         int N, int Wp, int idthres, int mslice,
                                                         int32_t* __restrict__ counts) {
  __shared__ uint32_t tile[RW_TILE * 5 * RW_WCH];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int m_begin = blockIdx.y * mslice;
  const int m_end = min(N, m_begin + mslice);
  const bool live = i < N;
  int cnt = 0;

  for (int mt = m_begin; mt < m_end; mt += RW_TILE) {
    int mism[RW_TILE];
#pragma unroll
    for (int m = 0; m < RW_TILE; m++) mism[m] = 0;

    for (int w0 = 0; w0 < Wp; w0 += RW_WCH) {
      __syncthreads();
      // stage planes of rows mt..mt+63, words w0..w0+7
      for (int e = threadIdx.x; e < RW_TILE * 5 * RW_WCH; e += 256) {
        const int m = e / (5 * RW_WCH), r = e % (5 * RW_WCH);
        const int p = r / RW_WCH, w = r % RW_WCH;
        const int mg = mt + m;
        tile[e] = (mg < N && w0 + w < Wp) ? planes[((long)mg * 5 + p) * Wp + w0 + w] : 0u;
      }
      __syncthreads();
      uint32_t own[5 * RW_WCH];
#pragma unroll
      for (int p = 0; p < 5; p++)
#pragma unroll
        for (int w = 0; w < RW_WCH; w++)
          own[p * RW_WCH + w] =
              (live && w0 + w < Wp) ? planes[((long)i * 5 + p) * Wp + w0 + w] : 0u;
#pragma unroll
      for (int m = 0; m < RW_TILE; m++) {
        const uint32_t* t = tile + m * 5 * RW_WCH;
        int acc = 0;
#pragma unroll
        for (int w = 0; w < RW_WCH; w++) {
          const uint32_t d = (own[w] ^ t[w]) | (own[RW_WCH + w] ^ t[RW_WCH + w]) |
                             (own[2 * RW_WCH + w] ^ t[2 * RW_WCH + w]) |
                             (own[3 * RW_WCH + w] ^ t[3 * RW_WCH + w]) |
                             (own[4 * RW_WCH + w] ^ t[4 * RW_WCH + w]);
          acc += __popc(d);
        }
        mism[m] += acc;
      }
    }
#pragma unroll
    for (int m = 0; m < RW_TILE; m++) {
      const int mg = mt + m;
      cnt += (mg < m_end && mg != i && (L - mism[m]) > idthres) ? 1 : 0;
    }
  }
  if (live && cnt) atomicAdd(&counts[i], cnt);
}

__global__ void k_weights_from_counts(const int32_t* __restrict__ counts, int N, int diag,
                                      float* __restrict__ w) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const float acc = (float)(counts[n] + diag);
  w[n] = (float)(1.0 / (double)(acc - 1.0f));
}

__global__ void k_fill_f32(float* p, long n, float v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_zero_i32(int32_t* p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

extern "C" {

int ccm_reweight_words(int L) { return ((L + 31) / 32 + RW_WCH - 1) / RW_WCH * RW_WCH; }

// planes: N*5*ccm_reweight_words(L) uint32 of scratch; counts: N int32 (output).
int ccm_launch_identity_counts(const uint8_t* d_msa, int L, int N, int idthres, uint32_t* d_planes,
                               int32_t* d_counts, hipStream_t s) {
  if (L <= 0 || N <= 0) return CCM_ERR_ARG;
  const int Wp = ccm_reweight_words(L);
  const long tot = (long)N * Wp;
  hipLaunchKernelGGL(k_pack_planes, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, d_msa, L, N,
                     Wp, d_planes);
  CCM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_zero_i32, dim3((N + 255) / 256), dim3(256), 0, s, d_counts, (long)N);
  CCM_LAUNCH_CHECK();
  const int gx = (N + 255) / 256;
  int gy = (1024 + gx - 1) / gx;  // aim for ~1024 workgroups
  int mslice = (N + gy - 1) / gy;
  mslice = (mslice + RW_TILE - 1) / RW_TILE * RW_TILE;
  gy = (N + mslice - 1) / mslice;
  hipLaunchKernelGGL(k_identity_counts, dim3(gx, gy), dim3(256), 0, s, d_planes, L, N, Wp, idthres,
                     mslice, d_counts);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_weights_from_counts(const int32_t* d_counts, int N, int L, int idthres, float* d_w,
                                   hipStream_t s) {
  if (N <= 0) return CCM_ERR_ARG;
  const int diag = L > idthres ? 2 : 0;
  hipLaunchKernelGGL(k_weights_from_counts, dim3((N + 255) / 256), dim3(256), 0, s, d_counts, N, diag,
                     d_w);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_fill_f32(float* d_p, long n, float v, hipStream_t s) {
  if (n <= 0) return CCM_OK;
  hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_p, n, v);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"
