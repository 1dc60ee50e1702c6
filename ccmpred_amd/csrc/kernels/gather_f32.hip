// Plain FP32 one-hot x coupling contractions of the PLL path.
//
// Forward (evaluate_cpu.c:55-75, d_compute_pc cu:30-107):
//   Z[n][a*L+j] = [a<20] V(j,a) + sum_k W(x_nk, k, a, j)
// Every x_nk selects one coupling row (x_nk, k) of 21L floats; the kernel
// stages, per chunk of 4 columns k, all 21 candidate rows of a 128-column
// slab in LDS and lets each wave gather them for 16 sequences.
//
// Pair gradient (evaluate_cpu.c:112-126, d_compute_edge_gradients cu:279-312):
//   G[(b*L+k)][c] = sum_n [x_nk = b] D[n][c],  D = w_n (P_n - onehot_n)
// Each wave owns 4 columns k and scatters D rows into wave-private LDS
// accumulators of 21 x 64 columns per k, so the result is deterministic.
#include "common.h"

#define GF_SLAB 128  // forward: columns per workgroup (2 per lane)
#define GF_TK 4      // forward: columns k per LDS chunk
#define GF_SEQ 16    // forward: sequences per wave
#define GF_SB 4      // forward: sequences whose LDS reads are in flight together
#define GF_NT 1024   // forward: threads per workgroup (256 sequences)
#define GF_NS (GF_NT / 64 * GF_SEQ)

// One workgroup: a 128-column slab for 256 sequences, 16 waves of 16
// sequences.  Per chunk of GF_TK columns k the 21 candidate rows of the
// slab are staged in LDS; the next chunk's rows and one-hot offsets are
// loaded into registers while the current one is summed, and each wave
// keeps the LDS reads of GF_SB sequences in flight before it adds them (the
// adds of one sequence stay in k order, so the sums are those of a
// sequence-at-a-time loop bit for bit).
__global__ __launch_bounds__(GF_NT) void k_gather_fwd_f32(const uint8_t* __restrict__ msap, int Lp,
                                                          const float* __restrict__ x1,
                                                          const float* __restrict__ x2, int L, int N,
                                                          float* __restrict__ Z, long ldz,
                                                          uint32_t tiles_n, uint32_t tiles,
                                                          uint32_t total) {
  constexpr int NE = GF_TK * CCM_NA * 64;      // staged float2 per chunk
  constexpr int PER = (NE + GF_NT - 1) / GF_NT;
  __shared__ float2 wt[NE];                    // [kk][b][lane]
  __shared__ uint4 offs[GF_NS];                // per sequence: byte offsets for kk = 0..3
  const uint32_t q = ccm_xcd_remap(blockIdx.x, total);
  if (q >= tiles) return;
  const int ct = q / tiles_n, nt = q % tiles_n;
  const int c0 = ct * GF_SLAB, n0 = nt * GF_NS;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long C21 = (long)CCM_NA * L, ld2 = (long)CCM_NAP * L;
  const int c = c0 + 2 * lane;

  // seed with the singles (a < 20), gap block starts at 0
  float2 init = make_float2(0.f, 0.f);
  if (c < 20L * L) init.x = x1[c];
  if (c + 1 < 20L * L) init.y = x1[c + 1];
  float2 acc[GF_SEQ];
#pragma unroll
  for (int i = 0; i < GF_SEQ; i++) acc[i] = init;

  float2 pf[PER];
  uint4 po = make_uint4(0u, 0u, 0u, 0u);
  // the rows and offsets of the chunk at k0 into registers
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = tid + u * GF_NT;
      float2 v = make_float2(0.f, 0.f);
      if (e < NE) {
        const int kk = e / (CCM_NA * 64), r = e % (CCM_NA * 64);
        const int b = r / 64, cc = c0 + 2 * (r % 64);
        const int k = k0 + kk;
        if (k < L) {
          const float* src = x2 + (long)(b * L + k) * ld2 + cc;
          if (cc + 1 < C21) {
            v = *reinterpret_cast<const float2*>(src);
          } else if (cc < C21) {
            v.x = src[0];
          }
        }
      }
      pf[u] = v;
    }
    if (tid < GF_NS) {
      const int n = n0 + tid;
      uint32_t o[GF_TK];
#pragma unroll
      for (int kk = 0; kk < GF_TK; kk++) {
        const int k = k0 + kk;
        const uint32_t x = (n < N && k < L) ? (uint32_t)msap[(long)n * Lp + k] : 0u;
        o[kk] = (uint32_t)((kk * CCM_NA + x) * 64) * 8u;
      }
      po = make_uint4(o[0], o[1], o[2], o[3]);
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < L; k0 += GF_TK) {
    __syncthreads();  // the previous chunk is no longer read
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = tid + u * GF_NT;
      if (e < NE) wt[e] = pf[u];
    }
    if (tid < GF_NS) offs[tid] = po;
    __syncthreads();
    if (k0 + GF_TK < L) fetch(k0 + GF_TK);
    const char* base = reinterpret_cast<const char*>(wt) + lane * 8;
#pragma unroll
    for (int i0 = 0; i0 < GF_SEQ; i0 += GF_SB) {
      float2 v[GF_SB][GF_TK];
#pragma unroll
      for (int s = 0; s < GF_SB; s++) {
        const uint4 o = offs[wv * GF_SEQ + i0 + s];
        v[s][0] = *reinterpret_cast<const float2*>(base + o.x);
        v[s][1] = *reinterpret_cast<const float2*>(base + o.y);
        v[s][2] = *reinterpret_cast<const float2*>(base + o.z);
        v[s][3] = *reinterpret_cast<const float2*>(base + o.w);
      }
#pragma unroll
      for (int s = 0; s < GF_SB; s++) {
        float2 a = acc[i0 + s];
#pragma unroll
        for (int kk = 0; kk < GF_TK; kk++) {
          a.x += v[s][kk].x;
          a.y += v[s][kk].y;
        }
        acc[i0 + s] = a;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < GF_SEQ; i++) {
    const int n = n0 + wv * GF_SEQ + i;
    if (n < N) {
      float* dst = Z + (long)n * ldz + c;
      if (c + 1 < C21) {
        *reinterpret_cast<float2*>(dst) = acc[i];
      } else if (c < C21) {
        dst[0] = acc[i].x;
      }
    }
  }
}

#define GG_KW 4  // gradient: columns k per wave (16 per workgroup)

__global__ __launch_bounds__(256) void k_gather_grad_f32(const float* __restrict__ D, long ldd,
                                                         const uint8_t* __restrict__ msap, int Lp,
                                                         int L, int N, float* __restrict__ G,
                                                         long ldg, int accumulate, uint32_t tiles_k,
                                                         uint32_t tiles, uint32_t total) {
  // wave-private accumulators, one array per k slot so the compiler may
  // overlap their read-modify-write chains
  __shared__ float a0[4][CCM_NA * 64];
  __shared__ float a1[4][CCM_NA * 64];
  __shared__ float a2[4][CCM_NA * 64];
  __shared__ float a3[4][CCM_NA * 64];
  const uint32_t q = ccm_xcd_remap(blockIdx.x, total);
  if (q >= tiles) return;
  const int cs = q / tiles_k, kt = q % tiles_k;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long C21 = (long)CCM_NA * L;
  const int c = cs * 64 + lane;
  const int kb = kt * 16 + wv * GG_KW;
  const bool cok = c < C21;

  float* s0 = a0[wv];
  float* s1 = a1[wv];
  float* s2 = a2[wv];
  float* s3 = a3[wv];
  for (int b = 0; b < CCM_NA; b++) {
    s0[b * 64 + lane] = 0.f;
    s1[b * 64 + lane] = 0.f;
    s2[b * 64 + lane] = 0.f;
    s3[b * 64 + lane] = 0.f;
  }
  // the wave's 4 columns k are consecutive; msap rows are 32-byte aligned
  const bool k0ok = kb < L, k1ok = kb + 1 < L, k2ok = kb + 2 < L, k3ok = kb + 3 < L;
  if (k0ok) {
    for (int n0 = 0; n0 < N; n0 += 8) {
      float d[8];
      uint32_t xs[8];
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const int n = n0 + i;
        d[i] = (n < N && cok) ? D[(long)n * ldd + c] : 0.f;
        xs[i] = (n < N) ? *reinterpret_cast<const uint32_t*>(msap + (long)n * Lp + kb) : 0u;
      }
#pragma unroll
      for (int i = 0; i < 8; i++) {
        const uint32_t x = xs[i];
        const int o0 = (int)(x & 0xffu) * 64 + lane;
        const int o1 = (int)((x >> 8) & 0xffu) * 64 + lane;
        const int o2 = (int)((x >> 16) & 0xffu) * 64 + lane;
        const int o3 = (int)(x >> 24) * 64 + lane;
        s0[o0] += d[i];
        if (k1ok) s1[o1] += d[i];
        if (k2ok) s2[o2] += d[i];
        if (k3ok) s3[o3] += d[i];
      }
    }
  }
  if (cok) {
    for (int b = 0; b < CCM_NA; b++) {
      float* g0 = G + (long)(b * L + kb) * ldg + c;
      if (k0ok) g0[0] = accumulate ? g0[0] + s0[b * 64 + lane] : s0[b * 64 + lane];
      if (k1ok) g0[ldg] = accumulate ? g0[ldg] + s1[b * 64 + lane] : s1[b * 64 + lane];
      if (k2ok) g0[2 * ldg] = accumulate ? g0[2 * ldg] + s2[b * 64 + lane] : s2[b * 64 + lane];
      if (k3ok) g0[3 * ldg] = accumulate ? g0[3 * ldg] + s3[b * 64 + lane] : s3[b * 64 + lane];
    }
  }
}

// D = -w_n * onehot_n: run through the gradient kernel this yields
// G = -H2 with H2[(b,k)][(a,j)] = sum_n w_n [x_nk=b][x_nj=a], the data term
// of the reference's init-time histograms (cu:350-387).
__global__ void k_neg_weighted_onehot(const uint8_t* __restrict__ msap, int Lp, const float* __restrict__ w,
                                      int L, int N, float* __restrict__ D, long ldd) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long C21 = (long)CCM_NA * L;
  if (t >= (long)N * C21) return;
  const int n = (int)(t / C21);
  const int cidx = (int)(t % C21);
  const int a = cidx / L, j = cidx % L;
  D[(long)n * ldd + cidx] = (msap[(long)n * Lp + j] == a) ? -w[n] : 0.f;
}

static inline uint32_t round8(uint32_t v) { return (v + 7u) & ~7u; }

extern "C" {

// Z (N x ldz floats, ldz even and >= 21L) <- V + sum_k W rows
int ccm_launch_gather_fwd_f32(const uint8_t* d_msap, int Lp, const float* d_x, int L, int N,
                              float* d_z, long ldz, hipStream_t s) {
  if (L <= 0 || N <= 0 || (ldz & 1) || ldz < 21L * L || (Lp % 32) || Lp < L) return CCM_ERR_ARG;
  const float* x1 = d_x;
  const float* x2 = d_x + ccm_nsp(L);
  const uint32_t tiles_c = (uint32_t)((21L * L + GF_SLAB - 1) / GF_SLAB);
  const uint32_t tiles_n = (uint32_t)((N + GF_NS - 1) / GF_NS);
  const uint32_t tiles = tiles_c * tiles_n, total = round8(tiles);
  hipLaunchKernelGGL(k_gather_fwd_f32, dim3(total), dim3(GF_NT), 0, s, d_msap, Lp, x1, x2, L, N, d_z,
                     ldz, tiles_n, tiles, total);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

// G (21L x ldg floats) (+)= sum_n onehot_n^T D_n ; D is N x ldd
int ccm_launch_gather_grad_f32(const float* d_d, long ldd, const uint8_t* d_msap, int Lp, int L,
                               int N, float* d_g, long ldg, int accumulate, hipStream_t s) {
  if (L <= 0 || N <= 0 || ldg < 21L * L || ldd < 21L * L || (Lp % 32) || Lp < L) return CCM_ERR_ARG;
  const uint32_t tiles_c = (uint32_t)((21L * L + 63) / 64);
  const uint32_t tiles_k = (uint32_t)((L + 15) / 16);
  const uint32_t tiles = tiles_c * tiles_k, total = round8(tiles);
  hipLaunchKernelGGL(k_gather_grad_f32, dim3(total), dim3(256), 0, s, d_d, ldd, d_msap, Lp, L, N, d_g,
                     ldg, accumulate, tiles_k, tiles, total);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_neg_weighted_onehot(const uint8_t* d_msap, int Lp, const float* d_w, int L, int N,
                                   float* d_d, long ldd, hipStream_t s) {
  const long tot = (long)N * 21L * L;
  hipLaunchKernelGGL(k_neg_weighted_onehot, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, d_msap,
                     Lp, d_w, L, N, d_d, ldd);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Pair gradient as a segmented gather (deterministic, no read-modify-write).
//
// Sequences are taken in chunks of 128.  For every chunk and column k a
// precomputed record lists the chunk's rows grouped by residue b = x_nk
// (ascending), each group padded to a multiple of 4 with row 128 (a zero
// row), plus the group starts:
//   bytes [0, 192): row indices;  bytes [192, 214): start of group b, b = 0..21
// A workgroup of 16 waves stages 128 rows of a 128-column slab of D in LDS
// (the next chunk's rows and records are loaded into registers meanwhile)
// and each wave sums, for its column k and every residue b, the rows of
// group b into a register accumulator: G[(b*L + k)][c] = sum over the group
// of D[n][c].
#define GS_NC 128
#define GS_REC 256
#define GS_NT 1024  // threads: one wave per column k, 16 columns k per workgroup

__global__ __launch_bounds__(GS_NT) void k_gather_grad_seg_f32(
    const float* __restrict__ D, long ldd, const uint8_t* __restrict__ meta, int L, int N,
    float* __restrict__ G, long ldg, int accumulate, uint32_t tiles_k, uint32_t tiles, uint32_t total) {
  constexpr int KT = GS_NT / 64;                      // columns k per workgroup
  constexpr int PD = GS_NC * 64 / GS_NT;              // staged float2 per thread
  constexpr int NM = KT * GS_REC / 4;                 // record words per chunk
  constexpr int PM = (NM + GS_NT - 1) / GS_NT;
  __shared__ float2 dt[(GS_NC + 1) * 64];   // chunk rows (+ the zero row) x 128 columns
  __shared__ uint32_t mt[NM];               // records of the workgroup's columns k
  const uint32_t q = ccm_xcd_remap(blockIdx.x, total);
  if (q >= tiles) return;
  const int cs = q / tiles_k, kt = q % tiles_k;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long C21 = (long)CCM_NA * L;
  const int c0 = cs * 128;
  const int col = c0 + 2 * lane;
  const int kbase = kt * KT;

  float2 acc[CCM_NA];
#pragma unroll
  for (int b = 0; b < CCM_NA; b++) acc[b] = make_float2(0.f, 0.f);

  if (wv == 0) dt[GS_NC * 64 + lane] = make_float2(0.f, 0.f);
  const int nch = (N + GS_NC - 1) / GS_NC;
  float2 pd[PD];
  uint32_t pm[PM];
  // the rows and records of chunk ch into registers
  auto fetch = [&](int ch) {
#pragma unroll
    for (int u = 0; u < PD; u++) {
      const int e = tid + u * GS_NT;
      const int r = e >> 6, cc = c0 + 2 * (e & 63);
      const long n = (long)ch * GS_NC + r;
      float2 v = make_float2(0.f, 0.f);
      if (n < N) {
        const float* src = D + n * ldd + cc;
        if (cc + 1 < C21)
          v = *reinterpret_cast<const float2*>(src);
        else if (cc < C21)
          v.x = src[0];
      }
      pd[u] = v;
    }
#pragma unroll
    for (int u = 0; u < PM; u++) {
      const int e = tid + u * GS_NT;
      uint32_t w = 0u;
      if (e < NM) {
        const int kk = e / (GS_REC / 4), ww = e % (GS_REC / 4);
        const int k = kbase + kk;
        w = k < L ? reinterpret_cast<const uint32_t*>(meta + ((long)ch * L + k) * GS_REC)[ww] : 0u;
      }
      pm[u] = w;
    }
  };
  fetch(0);
  const int k = kbase + wv;
  for (int ch = 0; ch < nch; ch++) {
    __syncthreads();  // the previous chunk is no longer read
#pragma unroll
    for (int u = 0; u < PD; u++) dt[tid + u * GS_NT] = pd[u];
#pragma unroll
    for (int u = 0; u < PM; u++) {
      const int e = tid + u * GS_NT;
      if (e < NM) mt[e] = pm[u];
    }
    __syncthreads();
    if (ch + 1 < nch) fetch(ch + 1);
    if (k >= L) continue;
    const uint32_t* rec = mt + wv * (GS_REC / 4);
    const uint32_t pos4 = rec[lane];  // packed row indices 4*lane .. 4*lane+3
    uint32_t sg[6];
#pragma unroll
    for (int t = 0; t < 6; t++) sg[t] = __builtin_amdgcn_readfirstlane(rec[48 + t]);
#pragma unroll
    for (int b = 0; b < CCM_NA; b++) {
      const int s = (sg[b >> 2] >> (8 * (b & 3))) & 0xff;
      const int e = (sg[(b + 1) >> 2] >> (8 * ((b + 1) & 3))) & 0xff;
      float2 a = acc[b];
      for (int i = s; i < e; i += 4) {
        const uint32_t r4 = __builtin_amdgcn_readlane(pos4, i >> 2);
        const float2 v0 = dt[(r4 & 0xffu) * 64 + lane];
        const float2 v1 = dt[((r4 >> 8) & 0xffu) * 64 + lane];
        const float2 v2 = dt[((r4 >> 16) & 0xffu) * 64 + lane];
        const float2 v3 = dt[(r4 >> 24) * 64 + lane];
        a.x += v0.x; a.y += v0.y;
        a.x += v1.x; a.y += v1.y;
        a.x += v2.x; a.y += v2.y;
        a.x += v3.x; a.y += v3.y;
      }
      acc[b] = a;
    }
  }
  if (k < L) {
#pragma unroll
    for (int b = 0; b < CCM_NA; b++) {
      float* dst = G + (long)(b * L + k) * ldg + col;
      if (col + 1 < C21) {
        float2 v = acc[b];
        if (accumulate) {
          const float2 o = *reinterpret_cast<const float2*>(dst);
          v.x += o.x;
          v.y += o.y;
        }
        *reinterpret_cast<float2*>(dst) = v;
      } else if (col < C21) {
        dst[0] = accumulate ? dst[0] + acc[b].x : acc[b].x;
      }
    }
  }
}

extern "C" {

int ccm_seg_meta_bytes(int L, int N) { return ((N + GS_NC - 1) / GS_NC) * L * GS_REC; }

// G (21L x ldg floats, ldg even) (+)= sum_n onehot_n^T D_n over the N
// sequences the chunk records in d_meta describe (N x ldd floats of D)
int ccm_launch_gather_grad_seg_f32(const float* d_d, long ldd, const uint8_t* d_meta, int L, int N,
                                   float* d_g, long ldg, int accumulate, hipStream_t s) {
  if (L <= 0 || N <= 0 || ldg < 21L * L || ldd < 21L * L || (ldg & 1) || (ldd & 1)) return CCM_ERR_ARG;
  const uint32_t tiles_c = (uint32_t)((21L * L + 127) / 128);
  const uint32_t tiles_k = (uint32_t)((L + GS_NT / 64 - 1) / (GS_NT / 64));
  const uint32_t tiles = tiles_c * tiles_k, total = round8(tiles);
  hipLaunchKernelGGL(k_gather_grad_seg_f32, dim3(total), dim3(GS_NT), 0, s, d_d, ldd, d_meta, L, N, d_g,
                     ldg, accumulate, tiles_k, tiles, total);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"
