// Helpers of the direct-to-LDS MFMA kernels (mfma_bf16.hip, sparse_fwd.hip):
// explicit wait counts, the LDS hand-over barrier, LDS swizzles, buffer
// resources, LDS stages, and the column softmax that the fused forward
// contractions run as their epilogue.
#pragma once

#include "common.h"

// s_waitcnt with only the vector-memory count limited to N (gfx9 encoding:
// vmcnt bits [3:0] and [15:14], expcnt [6:4] and lgkmcnt [11:8] at maximum)
template <int N>
__device__ __forceinline__ void ccm_wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// workgroup barrier for LDS hand-over: wait for this wave's LDS operations
// (lgkmcnt(0)), then s_barrier, between two empty asm statements with a
// memory clobber that keep the compiler from moving memory accesses across.
// Written with the builtins, the LDS wait is visible to the compiler's own
// wait insertion; no vector-memory wait is implied (a release fence would
// add one for the direct-to-LDS loads of later stages still in flight).
__device__ __forceinline__ void ccm_lds_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt at their maxima
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// physical granule of logical granule g in row r of a dense stage with G
// granules per row (G = 8: rows of 128 bytes; G = 16: rows of 256 bytes)
template <int G>
__device__ __forceinline__ int swz_a(int r, int g) {
  return G == 8 ? (g ^ ((r >> 1) & 7)) : (g ^ (r & 15));
}
// physical granule of logical granule g of column c of a byte stage with GB
// granules per column (GB = 4: 64 bytes per column)
template <int GB>
__device__ __forceinline__ int swz_b(int c, int g) {
  return g ^ ((c >> 2) & (GB - 1));
}

// Buffer resources and the buffer form of the direct-to-LDS load.  The
// resource type and builtins exist in the device compilation only; the
// host pass, which never runs device code, sees inert stand-ins.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __amdgpu_buffer_rsrc_t ccm_rsrc_t;
__device__ __forceinline__ ccm_rsrc_t ccm_make_rsrc(const void* base) {
  // raw buffer, no stride, no bounds limit beyond the 31-bit offsets used
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, 0x7FFFFFFF, 0x00020000);
}
// 16 bytes per lane from base + voff + soff into LDS at lds + 16 * lane
__device__ __forceinline__ void ccm_buffer_load_lds16(ccm_rsrc_t r, void* lds, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 16, voff, soff, 0, 0);
}
// 4 bytes per lane from base + voff + soff into LDS at lds + 4 * lane
__device__ __forceinline__ void ccm_buffer_load_lds4(ccm_rsrc_t r, void* lds, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (__attribute__((address_space(3))) void*)lds, 4, voff, soff, 0, 0);
}
// one 16-bit store at base + voff + soff (32-bit offsets, no 64-bit address
// arithmetic per store)
__device__ __forceinline__ void ccm_buffer_store16(ccm_rsrc_t r, uint16_t v, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b16(v, r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void ccm_buffer_store32(ccm_rsrc_t r, uint32_t v, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)voff, (int)soff, 0);
}
#else
typedef const void* ccm_rsrc_t;
__device__ __forceinline__ ccm_rsrc_t ccm_make_rsrc(const void* base) { return base; }
__device__ __forceinline__ void ccm_buffer_load_lds16(ccm_rsrc_t, void*, uint32_t, uint32_t) {}
__device__ __forceinline__ void ccm_buffer_load_lds4(ccm_rsrc_t, void*, uint32_t, uint32_t) {}
__device__ __forceinline__ void ccm_buffer_store16(ccm_rsrc_t, uint16_t, uint32_t, uint32_t) {}
__device__ __forceinline__ void ccm_buffer_store32(ccm_rsrc_t, uint32_t, uint32_t, uint32_t) {}
#endif

// LDS stages of the direct-to-LDS kernel: one __shared__ object per stage
// (dense tile, then the byte tile), so that the compiler can tell a stage
// being read from the stages being filled and does not drain all
// outstanding loads before every LDS read
template <int BYTES, int J>
__device__ __forceinline__ uint8_t* dtl_stage() {
  __shared__ __attribute__((aligned(1024))) uint8_t s[BYTES];
  return s;
}

// What the fused forward (MODE 2) does with a tile of logits instead of
// storing it: the column softmax of softmax_cols_block for the tile's
// sequences, its probability operand Dt and its objective and node-gradient
// partials, per block of 128 sequences.
struct ccm_fwd_epi {
  const float* x1;       // singles V(j,a) = x1[a*L + j]
  const uint8_t* msapt;  // transposed padded MSA at the chunk's first sequence
  long ldmt;
  const float* w;        // weights from the chunk's first sequence
  int L, Nc;             // alignment columns; real sequences in the chunk
  uint16_t* dt;          // rows a*Lp + j, lddt apart; lo part at lo_off
  long lddt, lo_off;
  int passes, mode;      // Dt passes (1 | 2); mode 0: objective only
  int blk0;              // global index of the chunk's first 128-sequence block
  double* fpart;         // [block][L]
  double* g1part;        // [block][20][L]
};


// The column softmax of one round of 128 sequences of a fused forward tile
// (softmax_cols_block of softmax.hip, cu:30-107 and cu:114-141 of the
// reference): the tile's logits for those sequences sit in four LDS buffers
// b0..b3 of 32 sequences each (row R = a*JG + jj for residue a and tile
// column jj, row stride ESTR floats).  Item it of the round is alignment
// column jj = it / 2 for the 64 sequences of half hc = it % 2, one sequence
// per lane; waves take items it = w, w + NW, ...  Writes Dt (mode 1) and,
// per row of 16 lanes, the objective and node-gradient sums into fsum and
// gsum (slot ((r*2 + hc)*4 + row)*JG + jj).
template <int JG, int NW, int ESTR>
__device__ __forceinline__ void fwd_softmax_round(const float* b0, const float* b1, const float* b2,
                                                  const float* b3, const float* xs, float* gsum,
                                                  double* fsum, const ccm_fwd_epi& epi, int mt, int n0,
                                                  int Ncol, int r, int wave_u, int lane,
                                                  ccm_rsrc_t drsrc, uint32_t dstep, uint32_t dlo) {
        // Ncol is a multiple of 128: a round's sequences are all inside the
        // chunk or all past it
        const bool in = n0 + r * 128 < Ncol;
        for (int it = wave_u; in && it < 2 * JG; it += NW) {
          const int jj = it >> 1, hc = it & 1, j = mt * JG + jj;
          if (j >= epi.L) break;
          const int n = n0 + r * 128 + hc * 64 + lane;  // sequence within the chunk
          const bool ok = n < epi.Nc;
          const float* bp = (hc ? ((lane >> 5) ? b3 : b2) : ((lane >> 5) ? b1 : b0)) + jj * ESTR + (lane & 31);
          // padding sequences (past Nc) carry the finite logits of an all-255
          // MSA row and weight 0
          float z[CCM_NA];
  #pragma unroll
          for (int a = 0; a < CCM_NA; a++) {
            float v = bp[a * JG * ESTR];
            if (a < 20) v += xs[a * JG + jj];
            z[a] = v;
          }
          const int xb = ok ? epi.msapt[(long)j * epi.ldmt + n] : 0;
          const float wv = ok ? epi.w[n] : 0.f;
          // the observed residue's logit, read once more instead of selected
          const float zx = bp[xb * JG * ESTR] + (xb < 20 ? xs[xb * JG + jj] : 0.f);
          float mx = -INFINITY;
  #pragma unroll
          for (int a = 0; a < CCM_NA; a++) mx = fmaxf(mx, z[a]);
          float ssum = 0.f;
  #pragma unroll
          for (int a = 0; a < CCM_NA; a++) {
            z[a] = __expf(z[a] - mx);
            ssum += z[a];
          }
          double fl = ok ? (double)(wv * (mx + __logf(ssum) - zx)) : 0.0;
          const int slot = ((r * 2 + hc) * 4 + (lane >> 4)) * JG + jj;
          if (epi.mode) {
            const float ws = wv / ssum;
            float g1[20];
            const uint32_t voff = (uint32_t)(((long)j * epi.lddt + n) * 2);  // row j (residue 0)
  #pragma unroll
            for (int a = 0; a < CCM_NA; a++) {
              const float qv = z[a] * ws;
              const uint16_t h = ccm_f2bf(qv);
              ccm_buffer_store16(drsrc, h, voff, a * dstep);
              // residual of the exact product (explicit fused multiply-add, as
              // in softmax_cols)
              if (epi.passes > 1)
                ccm_buffer_store16(drsrc, ccm_f2bf(__builtin_fmaf(z[a], ws, -ccm_bf2f(h))), voff, a * dstep + dlo);
              if (a < 20) g1[a] = qv;  // the data term -w [x = a] is added once, at assembly
            }
  #pragma unroll
            for (int a = 0; a < 20; a++) g1[a] = ccm_row_sum_dpp_f32(g1[a]);
            if ((lane & 15) == 15) {
  #pragma unroll
              for (int a = 0; a < 20; a++) gsum[slot * 20 + a] = g1[a];
            }
          }
          fl = ccm_row_sum_dpp_f64(fl);
          if ((lane & 15) == 15) fsum[slot] = fl;
        }
}

// The partials per block of 128 sequences (one round) from its eight row
// sums, in a fixed order that no tile shape changes: g1part[blk][a][j]
// (mode 1) and fpart[blk][j], blk = blk0 + (n0 + r*128) / 128.
template <int JG, int NT, int ROUNDS>
__device__ __forceinline__ void fwd_softmax_partials(const float* gsum, const double* fsum,
                                                     const ccm_fwd_epi& epi, int mt, int n0, int Ncol,
                                                     int tid) {
      // partials per block of 128 sequences (one round): its eight row sums
      for (int e = tid; e < ROUNDS * JG * CCM_NA; e += NT) {
        const int r = e / (JG * CCM_NA), rem = e - r * (JG * CCM_NA);
        const int jj = rem / CCM_NA, a = rem - jj * CCM_NA;
        const int j = mt * JG + jj, nb = n0 + r * 128;
        if (j >= epi.L || nb >= Ncol) continue;
        const long gb = (long)epi.blk0 + nb / 128;
        const int s0i = r * 8 * JG + jj;
        if (a < 20) {
          if (epi.mode) {
            double t = 0.0;
  #pragma unroll
            for (int u = 0; u < 8; u++) t += (double)gsum[(s0i + u * JG) * 20 + a];
            epi.g1part[(gb * 20 + a) * epi.L + j] = t;
          }
        } else {
          double t = 0.0;
  #pragma unroll
          for (int u = 0; u < 8; u++) t += fsum[s0i + u * JG];
          epi.fpart[gb * epi.L + j] = t;
        }
      }
}
