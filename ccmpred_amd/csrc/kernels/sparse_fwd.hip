// The forward contraction on gfx950's 2:4 structured-sparse MFMA, with the
// column softmax as its epilogue (evaluate_cpu.c:55-109; cu:30-141).
//
// Z[n][c] = sum_r onehot[n][r] * W[r][c] with the one-hot operand as the
// sparse A operand of v_smfmac_f32_32x32x32_bf16 (rows = sequences n) and
// the couplings as the dense B operand (columns c = a*Lp + j, the rows of
// the coupling matrix in residue-major order).  The K dimension is laid out
// in groups of four slots: slot 0 is an explicit zero, slots 1..3 hold
// residues 3t, 3t+1, 3t+2 of one coupling column k, so the 21 residues of a
// column take 7 groups (28 slots).  Per group the sparse operand stores two
// values with 2-bit positions; they are always (0.0 at position 0,
// 1.0 at position p): p = x_nk - 3t + 1 when sequence n has residue
// x_nk in group t of column k, else p = 0, and the 1.0 then meets the zero
// slot.  The operand's values are therefore constants, and only the
// positions vary: 16 bits per lane and instruction, precomputed once from
// the MSA (k_sparse_index).  One instruction covers 8 groups, i.e. 32 K
// slots at the issue cost of a dense 32x32x16 instruction: 14 issue slots
// per column against 21 for the dense one-hot, and no one-hot arithmetic in
// the loop.  With the coupling split into a hi and a lo BF16 half (two
// passes), each instruction's 32 slots are stored once with the hi halves
// and once with the lo halves, and both use the same positions.
//
// Lane layout of the instruction (verified on the device): A lane l holds
// row l % 32 and groups 4 (l / 32) .. 4 (l / 32) + 3 (K slots 16 (l/32) ..
// +16), value v in group v / 2, its position in bits 2v..2v+1 of the
// selected 16-bit half of the index operand; B lane l holds column l % 32
// and K slots 8 (l/32) .. +8 (values 0..7) and 16 + 8 (l/32) .. +8
// (values 8..15); D lane l element i is row (i/4)*8 + 4 (l/32) + i%4,
// column l % 32.
#include "dtl.h"

#include <type_traits>

#include "ccmpred_amd_kernels.h"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x16_t __attribute__((ext_vector_type(16)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

namespace {

struct frag32 {  // the two granules of a B fragment, values 0..7 then 8..15
  uint4 lo, hi;
};

constexpr int SP_GROUPS = 7;           // groups of 4 K slots per coupling column
constexpr int SP_KC = 4 * SP_GROUPS;   // K slots per coupling column
constexpr int SP_TK = 64;              // K slots per LDS stage (two instructions)

// positions of instruction i for the lane half q of sequence row x (the
// row's residues, 255 past the alignment): 4 groups, 4 bits each
__device__ __forceinline__ uint32_t sparse_word(const uint8_t* __restrict__ x, int i, int q) {
  uint32_t w = 0;
#pragma unroll
  for (int g = 0; g < 4; g++) {
    const int G = 8 * i + 4 * q + g, k = G / SP_GROUPS, t = G - k * SP_GROUPS;
    const int b = x ? x[k] : 255;
    const int p = (b < CCM_NA && b / 3 == t) ? b - 3 * t + 1 : 0;
    w |= (uint32_t)(p << 2) << (4 * g);
  }
  return w;
}

// idx[(st*2 + q)*ldidx + n]: the positions of step st (two instructions:
// low and high 16 bits) for lane half q of sequence n.  One pass: step st
// holds instructions 2 st and 2 st + 1; two passes: instruction st twice
// (its hi and lo halves are the step's two instructions).
__global__ __launch_bounds__(256) void k_sparse_index(const uint8_t* __restrict__ msap, int Lp, int Np,
                                                      int passes, int nsteps, uint32_t* __restrict__ idx,
                                                      long ldidx) {
  const long total = (long)nsteps * 2 * ldidx;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
    const long n = e % ldidx, sq = e / ldidx;
    const int q = (int)(sq & 1), st = (int)(sq >> 1);
    const uint8_t* x = n < Np ? msap + n * Lp : nullptr;
    const int i0 = passes == 1 ? 2 * st : st, i1 = passes == 1 ? 2 * st + 1 : st;
    idx[e] = sparse_word(x, i0, q) | (sparse_word(x, i1, q) << 16);
  }
}

// couplings -> Ws (bf16): row c = a*Lp + j, K slot of column k, group t,
// position p >= 1 holds W(3t + p - 1, k, a, j) (zero for k >= L); each
// instruction's 32 slots are followed by their lo halves when PASSES = 2.
// Ws row c is column C = a*L + j of the x layout, so the conversion is a
// transpose, done by tiles through LDS: a workgroup reads the CW x-layout
// columns C0.. of the 168 rows (b, k) of CWS_CK = 8 coupling columns k
// with 16-byte loads (CW/4 lanes per row segment), stages them as BF16
// (hi, and lo for two passes) in LDS with row = C and the 168 values of a
// row in group order, and writes each Ws row's 56 groups (7 instructions)
// as contiguous 16-byte stores.  Rows of padding columns j >= L are not
// written (zeroed once).
constexpr int CWS_CK = 8;                   // coupling columns per tile
constexpr int CWS_R = CCM_NA * CWS_CK;      // staged values per Ws row (56 groups of 3)
constexpr int CWS_LD = CWS_R + 2;           // uint16 per LDS row: 340 bytes, 4-byte aligned
// nontemporal 16-byte load (streaming data: see convert_ws_pad_tile)
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 cws_load4(const float* p) {
  const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
  return make_float4(t.x, t.y, t.z, t.w);
}

// block sum of a 256-thread workgroup's FP64 terms into *out (thread 0).
// In a workgroup of 512 threads the halves are two such blocks (waves 0-3
// and 4-7), each with its own out (null: no store); every thread calls it.
__device__ __forceinline__ void cws_block_sum(double r, double* __restrict__ out) {
  __shared__ double red[8];
  r = ccm_wave_sum_f64(r);
  const int w = (int)(threadIdx.x >> 6);
  if ((threadIdx.x & 63) == 0) red[w] = r;
  __syncthreads();
  if ((threadIdx.x & 255) == 0 && out) {
    const int h = 4 * (w >> 2);
    *out = (red[h] + red[h + 1]) + (red[h + 2] + red[h + 3]);
  }
}

// The padding columns a >= 21 of the pair block: their gradient is
// 2 lambda_pair x and their regularisation 0.5 lambda_pair x^2
// (evaluate_cpu.c:158-161).  Tile (bxp, by): the 168 rows (b, k),
// k in [8 by, 8 by + 8), over the 128 x-layout columns from
// Cs = 4 floor(21 L / 4) + 128 bxp (16-byte granules; in the first one
// only the columns >= 21 L count), every load of the tile issued first.
// Each 256 threads take one tile (tile false: none; the sum is still joined).
constexpr int CWP = 128;
__device__ __forceinline__ void convert_ws_pad_tile(const float* __restrict__ x2, float* __restrict__ g2, int L,
                                                    float lam2, int bxp, int by, bool tile,
                                                    double* __restrict__ regpart) {
  constexpr int F4 = CWP / 4, NE = CWS_R * F4, PER = NE / 256;
  static_assert(NE % 256 == 0, "granules per thread");
  const int C21 = CCM_NA * L;
  const long ld2 = (long)CCM_NAP * L;
  const int k0 = by * CWS_CK, Cs = (C21 & ~3) + bxp * CWP;
  const int t = (int)(threadIdx.x & 255);
  const bool al = ((uintptr_t)x2 & 15u) == 0 && ((uintptr_t)g2 & 15u) == 0;
  const float two_l = 2.0f * lam2, half_l = 0.5f * lam2;
  float4 v[PER];
  long off[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int e = t + 256 * u, row = e / F4, C = Cs + 4 * (e - row * F4);
    const int b = row / CWS_CK, k = k0 + (row - b * CWS_CK);
    off[u] = (tile && k < L && C < ld2) ? (long)(b * L + k) * ld2 + C : -1;
    v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (off[u] >= 0) {
      // streaming accesses (x and g here are touched once per evaluation):
      // they leave the caches to the contractions' operands
      const float* src = x2 + off[u];
      if (al) {
        v[u] = cws_load4(src);
      } else {
        v[u] = make_float4(__builtin_nontemporal_load(src), __builtin_nontemporal_load(src + 1),
                           __builtin_nontemporal_load(src + 2), __builtin_nontemporal_load(src + 3));
      }
    }
  }
  double r = 0.0;
#pragma unroll
  for (int u = 0; u < PER; u++) {
    if (off[u] < 0) continue;
    const int C = Cs + 4 * ((t + 256 * u) % F4);
    const float q[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
    float o[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      o[c] = 0.f + two_l * q[c];
      if (C + c >= C21) r += (double)(half_l * q[c] * q[c]);
    }
    float* dst = g2 + off[u];
    if (C >= C21 && al) {
      const f4v ov = {o[0], o[1], o[2], o[3]};
      __builtin_nontemporal_store(ov, reinterpret_cast<f4v*>(dst));
    } else {
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (C + c >= C21) __builtin_nontemporal_store(o[c], dst + c);
    }
  }
  cws_block_sum(r, tile ? regpart : nullptr);
}

// the singles' regularisation lambda_single sum V^2 over the 20 L singles
// (evaluate_cpu.c:153-156), one block
__device__ __forceinline__ void convert_ws_singles_reg(const float* __restrict__ x1, int L, float lam1,
                                                       double* __restrict__ out) {
  double r = 0.0;
  for (int v = (int)threadIdx.x; v < 20 * L; v += 256) r += (double)(lam1 * x1[v] * x1[v]);
  cws_block_sum(r, out);
}

template <int PASSES>
struct cws_shape {
  static constexpr int CW = PASSES == 1 ? 128 : 64;  // x-layout columns (Ws rows) per tile
};

template <int PASSES>
__device__ __forceinline__ void convert_ws_block(const float* __restrict__ x2, int L, int Lp,
                                                 uint16_t* __restrict__ ws, long lds, int bx, int by,
                                                 float half_l, double* __restrict__ regpart) {
  constexpr int CW = cws_shape<PASSES>::CW;
  constexpr int F4 = CW / 4;                  // 16-byte loads per row segment
  constexpr int NE = CWS_R * F4;              // loads of the tile
  constexpr int PER = (NE + 255) / 256;       // per thread
  constexpr int Q = 4 * SP_GROUPS;            // 16-byte stores per Ws row and pass: 56 groups
  __shared__ __attribute__((aligned(16))) uint16_t sh[PASSES][CW * CWS_LD];
  const int C21 = CCM_NA * L;
  const long ld2 = (long)CCM_NAP * L;
  const int C0 = bx * CW, k0 = by * CWS_CK;
  const bool al = ((uintptr_t)x2 & 15u) == 0;
  // all loads first: element e = staged row (b, kk) x 16-byte column f
  float4 v[PER];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int e = (int)threadIdx.x + 256 * u;
    v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (e < NE) {
      const int row = e / F4, f = e - row * F4;
      const int b = row / CWS_CK, kk = row - b * CWS_CK;
      const int k = k0 + kk, C = C0 + 4 * f;
      if (k < L && C < C21) {
        // (the row is 32 L floats long: C + 3 stays inside it); nontemporal:
        // x is read once here and not again before the assembly (0.044 ->
        // 0.042 ms at L=250 against plain loads, same values)
        const float* src = x2 + (long)(b * L + k) * ld2 + C;
        v[u] = al ? cws_load4(src) : make_float4(src[0], src[1], src[2], src[3]);
      }
    }
  }
  if (regpart) {
    // the objective's regularisation over the values this block read
    // (each coupling of the a < 21 block is read by one block; a 16-byte
    // load at the last columns also holds up to three padding columns
    // C >= 21 L, which the padding blocks count)
    double r = 0.0;
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int e = (int)threadIdx.x + 256 * u;
      const int C = C0 + 4 * (e % F4);
      const float q0 = v[u].x, q1 = C + 1 < C21 ? v[u].y : 0.f;
      const float q2 = C + 2 < C21 ? v[u].z : 0.f, q3 = C + 3 < C21 ? v[u].w : 0.f;
      r += ((double)(half_l * q0 * q0) + (double)(half_l * q1 * q1)) +
           ((double)(half_l * q2 * q2) + (double)(half_l * q3 * q3));
    }
    cws_block_sum(r, regpart);
  }
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int e = (int)threadIdx.x + 256 * u;
    if (e < NE) {
      const int row = e / F4, f = e - row * F4;
      const int b = row / CWS_CK, kk = row - b * CWS_CK;
      const int o = 4 * f * CWS_LD + kk * CCM_NA + b;
      const uint32_t h01 = ccm_f2bf2(v[u].x, v[u].y), h23 = ccm_f2bf2(v[u].z, v[u].w);
      sh[0][o] = (uint16_t)h01;
      sh[0][o + CWS_LD] = (uint16_t)(h01 >> 16);
      sh[0][o + 2 * CWS_LD] = (uint16_t)h23;
      sh[0][o + 3 * CWS_LD] = (uint16_t)(h23 >> 16);
      if constexpr (PASSES > 1) {
        const uint32_t l01 = ccm_f2bf2(v[u].x - ccm_bf2f((uint16_t)h01), v[u].y - ccm_bf2f((uint16_t)(h01 >> 16)));
        const uint32_t l23 = ccm_f2bf2(v[u].z - ccm_bf2f((uint16_t)h23), v[u].w - ccm_bf2f((uint16_t)(h23 >> 16)));
        sh[1][o] = (uint16_t)l01;
        sh[1][o + CWS_LD] = (uint16_t)(l01 >> 16);
        sh[1][o + 2 * CWS_LD] = (uint16_t)l23;
        sh[1][o + 3 * CWS_LD] = (uint16_t)(l23 >> 16);
      }
    }
  }
  __syncthreads();
  // store q of a Ws row: groups 2q, 2q+1, i.e. staged values 6q .. 6q+5,
  // as the slots (0, v0, v1, v2, 0, v3, v4, v5); consecutive lanes take
  // consecutive q, so that a row's stores are contiguous
  for (int e = (int)threadIdx.x; e < CW * Q; e += 256) {
    const int cl = e / Q, q = e - cl * Q, C = C0 + cl;
    if (C >= C21) continue;
    const int a = C / L;
    uint16_t* row = ws + ((long)a * Lp + (C - a * L)) * lds;
#pragma unroll
    for (int p = 0; p < PASSES; p++) {
      const uint32_t* w = reinterpret_cast<const uint32_t*>(&sh[p][cl * CWS_LD + 6 * q]);
      const uint32_t w0 = w[0], w1 = w[1], w2 = w[2];
      const uint4 o = make_uint4(w0 << 16, (w0 >> 16) | (w1 << 16), w1 & 0xFFFF0000u, w2);
      // instruction 7 (k0 / 8) + q / 4 of the row, its 16-byte quarter q % 4
      const long slot = PASSES == 1 ? 28L * k0 + 8 * q
                                    : 64L * (7 * (k0 / CWS_CK) + q / 4) + 8 * (q % 4) + 32 * p;
      *reinterpret_cast<uint4*>(row + slot) = o;
    }
  }
}

// blocks of one conversion: x-layout column tiles by coupling-column tiles
__host__ __device__ __forceinline__ int convert_ws_tiles_c(int L, int passes) {
  const int cw = passes == 1 ? cws_shape<1>::CW : cws_shape<2>::CW;
  return (CCM_NA * L + cw - 1) / cw;
}
__host__ __device__ __forceinline__ int convert_ws_rows(int Lp) { return Lp / CWS_CK; }
// padding-column tiles of the form that also writes the padding gradient
__host__ __device__ __forceinline__ int convert_ws_tiles_p(int L) {
  return (CCM_NAP * L - ((CCM_NA * L) & ~3) + CWP - 1) / CWP;
}

// One block of the conversion grid (bx, by).  Plain form (regpart null):
// bx < tiles_c, the coupling transpose.  With the regularisation: the
// transpose blocks bx < tiles_c also write the regularisation partial
// regpart[by * (tiles_c + tiles_p) + bx]; the padding tiles bx - tiles_c
// (partials at the same index) run here when tp_run = tiles_p (else, with
// tp_run = 0, in the forward's launch); block (tiles_c + tp_run, 0) writes
// the singles' regularisation into *regsingle (the other blocks of that
// column return).
// pad_zero: x's padding is known to be zero and g's to hold zeros (a CG
// solve after its padding scan): the padding tiles write zero partials only.
template <int PASSES>
__device__ __forceinline__ void convert_ws_any(const float* __restrict__ x, long nsp, int L, int Lp,
                                               uint16_t* __restrict__ ws, long lds, float* __restrict__ g,
                                               float lam1, float lam2, double* __restrict__ regpart,
                                               double* __restrict__ regsingle, int tp_run, bool pad_zero,
                                               int bx, int by) {
  const int tc = convert_ws_tiles_c(L, PASSES), tp = convert_ws_tiles_p(L);
  if (bx < tc) {
    convert_ws_block<PASSES>(x + nsp, L, Lp, ws, lds, bx, by, 0.5f * lam2,
                             regpart ? regpart + (long)by * (tc + tp) + bx : nullptr);
  } else if (bx < tc + tp_run) {
    double* const part = regpart + (long)by * (tc + tp) + bx;
    if (pad_zero) {
      if (threadIdx.x == 0) *part = 0.0;
    } else {
      convert_ws_pad_tile(x + nsp, g + nsp, L, lam2, bx - tc, by, true, part);
    }
  } else if (by == 0) {
    convert_ws_singles_reg(x, L, lam1, regsingle);
  }
}

template <int PASSES>
__global__ __launch_bounds__(256) void k_convert_ws(const float* __restrict__ x, long nsp, int L, int Lp,
                                                    uint16_t* __restrict__ ws, long lds, float* __restrict__ g,
                                                    float lam1, float lam2, double* __restrict__ regpart,
                                                    double* __restrict__ regsingle, int tp_run) {
  convert_ws_any<PASSES>(x, nsp, L, Lp, ws, lds, g, lam1, lam2, regpart, regsingle, tp_run, false, blockIdx.x,
                         blockIdx.y);
}

// every alignment of a group (x buffers of a batch are whole alignments),
// always with the regularisation and the padding gradient
template <int PASSES>
__global__ __launch_bounds__(256) void k_convert_ws_grouped(const ccm_bprob* __restrict__ probs,
                                                            const uint32_t* __restrict__ offs,
                                                            const int32_t* __restrict__ pidx, int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  const int t = (int)(blockIdx.x - offs[gi]);
  const int tp = convert_ws_tiles_p(B.L), tcx = convert_ws_tiles_c(B.L, PASSES) + tp + 1;
  convert_ws_any<PASSES>(B.x, B.nsp, B.L, B.Lp, B.wt, B.ldwt, B.g, B.lam1, B.lam2, B.regpair, B.regsingle,
                         tp, B.ctl->pad_scanned && !B.ctl->x_pad_nz, t % tcx, t / tcx);
}

// One workgroup: TM = WM*RB*32 sequences x TN = WN*CB*32 gathered rows of Ws
// (all 21 residues of JG = TN / 21 alignment columns, as the dense fused
// forward gathers them), K in steps of 64 slots through three direct-to-LDS
// stages.  Row block rb of wave row wm holds the sequences 4*row + rb % 4 of
// the tile's 128-sequence round wm*RB/4 + rb/4, so that one 16-byte LDS read
// brings the positions of four row blocks (idx row layout
// [step][half][sequence]).  (A 1 x 8 wave grid with eight row blocks per
// wave, which halves the LDS reads of the dense fragments, measured the
// same: 0.167 against 0.163 ms.)
template <int WM, int WN, int RB, int CB>
__device__ __forceinline__ void sparse_fwd_tile(const uint16_t* __restrict__ Ws, long lds, int Lp,
                                                int nsteps, const uint32_t* __restrict__ idx, long ldidx,
                                                int Ncol, int mt, int nt, const ccm_fwd_epi& epi) {
  constexpr int NW = WM * WN;
  constexpr int TM = WM * RB * 32, TN = WN * CB * 32;
  constexpr int G = SP_TK / 8;                 // 16-byte granules per staged row
  constexpr int A_INS = TN * G / (64 * NW);    // DTL instructions per wave and stage
  constexpr int PER_STEP = A_INS + 1;          // and one 4-byte DTL load of positions
  constexpr int JG = TN / CCM_NA;
  constexpr int ESTR = 33;                     // epilogue row stride (floats)
  constexpr int W_BYTES = TN * SP_TK * 2;
  constexpr int P_BYTES = 2 * TM * 4;          // positions: [half][sequence] words
  constexpr int E_BYTES = TN * ESTR * 4;
  constexpr int STAGE_BYTES = W_BYTES + P_BYTES > E_BYTES ? W_BYTES + P_BYTES : E_BYTES;
  constexpr int RQ = RB / 4;   // groups of four row blocks per wave: 128 sequences each
  constexpr int ROUNDS = TM / 128;
  static_assert(RB % 4 == 0, "row blocks in fours: one 16-byte read of positions per four");
  static_assert(P_BYTES == NW * 256, "one 4-byte DTL load of positions per wave and stage");
  static_assert((TN * G) % (64 * NW) == 0, "dense loader mapping");
  static_assert(PER_STEP < 64, "vmcnt range");

  const int n0 = nt * TM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const int l32 = lane & 31, q = lane >> 5;
  const int wave_u = __builtin_amdgcn_readfirstlane(wave);

  // gathered rows of Ws: local row a*JG + jj <- row a*Lp + j, j = mt*JG + jj
  // (rows past 21*JG and columns j >= L repeat a row whose results are dropped)
  const ccm_rsrc_t wrsrc = ccm_make_rsrc(Ws);
  uint32_t woff[A_INS];
#pragma unroll
  for (int i = 0; i < A_INS; i++) {
    const int p = (wave * A_INS + i) * 64 + lane;
    const int r = p / G, g = swz_a<G>(r, p % G);
    const int a = r / JG < CCM_NA ? r / JG : 0;
    const int j = mt * JG + (r - (r / JG) * JG);
    const long row = (long)a * Lp + (j < epi.L ? j : epi.L - 1);
    woff[i] = (uint32_t)((row * lds + g * 8) * 2);
  }
  // positions of a step: wave w loads half w / (NW/2), sequences
  // n0 + (w % (NW/2))*64 + lane, into stage bytes W_BYTES + 256 w (i.e. word
  // [half][sequence of the tile]); a lane reads the words of its sequences
  // wm*128 + 4*l32 + (0..3) of half q
  const ccm_rsrc_t irsrc = ccm_make_rsrc(idx);
  const uint32_t ioff =
      (uint32_t)(((long)(wave / (NW / 2)) * ldidx + n0 + (wave % (NW / 2)) * 64 + lane) * 4);
  const uint32_t istep = (uint32_t)(2 * ldidx * 4);
  // row block rb of wave row wm holds the sequences 4 rho + rb % 4 of the
  // tile's 128-sequence round wm*RQ + rb/4 (rho = the block's row)
  const int pread = W_BYTES + (q * TM + wm * RQ * 128 + 4 * l32) * 4;

  floatx16 acc[RB][CB];
#pragma unroll
  for (int i = 0; i < RB; i++)
#pragma unroll
    for (int j = 0; j < CB; j++) acc[i][j] = (floatx16)0.0f;
  const bf16x8_t aconst = __builtin_bit_cast(bf16x8_t, make_uint4(0x3F800000u, 0x3F800000u, 0x3F800000u, 0x3F800000u));

  auto issue = [&](uint8_t* __restrict__ sp, int st) {
    const uint32_t c = (uint32_t)st * (SP_TK * 2);
#pragma unroll
    for (int i = 0; i < A_INS; i++) ccm_buffer_load_lds16(wrsrc, sp + (wave_u * A_INS + i) * 1024, woff[i], c);
    ccm_buffer_load_lds4(irsrc, sp + W_BYTES + wave_u * 256, ioff, (uint32_t)st * istep);
  };
  auto ld_pos = [&](const uint8_t* __restrict__ sp, uint4(&pr)[RQ]) {
#pragma unroll
    for (int h = 0; h < RQ; h++) pr[h] = *reinterpret_cast<const uint4*>(sp + pread + h * 512);
  };
  // B fragments of instruction sub (0, 1) of a stage: two granules per lane
  uint4 bf[2][CB][2];
  auto ld_frags = [&](const uint8_t* __restrict__ sp, uint4(&b)[CB][2], int sub) {
    const uint16_t* ws = reinterpret_cast<const uint16_t*>(sp);
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      const int r = (wn * CB + cb) * 32 + l32;
      b[cb][0] = *reinterpret_cast<const uint4*>(&ws[r * SP_TK + swz_a<G>(r, 4 * sub + q) * 8]);
      b[cb][1] = *reinterpret_cast<const uint4*>(&ws[r * SP_TK + swz_a<G>(r, 4 * sub + 2 + q) * 8]);
    }
  };
  auto mfmas = [&](const uint4(&b)[CB][2], const uint4(&pr)[RQ], auto sub_c) {
    constexpr int SUB = decltype(sub_c)::value;
#pragma unroll
    for (int cb = 0; cb < CB; cb++) {
      const frag32 f{b[cb][0], b[cb][1]};
      const bf16x16_t bv = __builtin_bit_cast(bf16x16_t, f);
#pragma unroll
      for (int rb = 0; rb < RB; rb++) {
        const uint4& p = pr[rb / 4];
        const int pos = (rb & 3) == 0 ? (int)p.x : (rb & 3) == 1 ? (int)p.y : (rb & 3) == 2 ? (int)p.z : (int)p.w;
        acc[rb][cb] = __builtin_amdgcn_smfmac_f32_32x32x32_bf16(aconst, bv, acc[rb][cb], pos, 0, SUB);
      }
    }
  };
  // one step on the stage at cur; nxt is the next step's stage, pc the
  // positions of this step.  Instruction 1's fragments are read before
  // instruction 0's MFMAs; before instruction 1's MFMAs the step hands over:
  // wait for the next stage, barrier, read the next step's positions and
  // first fragments, refill cur with step st+3.
  uint4 pc[RQ];
  auto step = [&](uint8_t* cur, const uint8_t* nxt, int st) {
    ld_frags(cur, bf[1], 1);
    mfmas(bf[0], pc, std::integral_constant<int, 0>{});
    uint4 pcur[RQ];
#pragma unroll
    for (int h = 0; h < RQ; h++) pcur[h] = pc[h];
    ccm_wait_vmcnt<PER_STEP>();
    ccm_lds_barrier();
    ld_frags(nxt, bf[0], 0);
    ld_pos(nxt, pc);
    __builtin_amdgcn_sched_barrier(0);
    issue(cur, st + 3 < nsteps ? st + 3 : nsteps - 1);
    mfmas(bf[1], pcur, std::integral_constant<int, 1>{});
  };

  uint8_t* s0 = dtl_stage<STAGE_BYTES, 0>();
  uint8_t* s1 = dtl_stage<STAGE_BYTES, 1>();
  uint8_t* s2 = dtl_stage<STAGE_BYTES, 2>();
  issue(s0, 0);
  issue(s1, nsteps > 1 ? 1 : 0);
  issue(s2, nsteps > 2 ? 2 : 0);
  ccm_wait_vmcnt<0>();
  ccm_lds_barrier();
  ld_frags(s0, bf[0], 0);
  ld_pos(s0, pc);
  // complete triples without conditions (one scheduling region each: with a
  // conditional step the compiler loses count of the loads in flight and
  // waits for all of them), then the last one or two steps
  int st = 0;
  for (; st + 3 <= nsteps; st += 3) {
    step(s0, s1, st);
    step(s1, s2, st + 1);
    step(s2, s0, st + 2);
  }
  if (st < nsteps) {
    step(s0, s1, st);
    if (st + 1 < nsteps) step(s1, s2, st + 1);
  }
  ccm_wait_vmcnt<0>();

  // Softmax epilogue in rounds of 128 sequences (wave rows): the waves of
  // round r put their accumulators into four LDS buffers, sequence-block
  // major as fwd_softmax_round reads them: output row rho of block rb is
  // sequence 4*rho + rb of the round, i.e. buffer rho / 8 = i / 4, column
  // 4*(rho % 8) + rb.
  ccm_lds_barrier();
  float* const b0 = reinterpret_cast<float*>(s0);
  float* const b1 = reinterpret_cast<float*>(s1);
  float* const b2 = reinterpret_cast<float*>(s2);
  float* const b3 = reinterpret_cast<float*>(dtl_stage<STAGE_BYTES, 3>());
  constexpr int NSLOT = ROUNDS * 2 * 4 * JG;
  __shared__ float gsum[NSLOT * 20];
  __shared__ double fsum[NSLOT];
  __shared__ float xs[20 * JG];
  for (int e = tid; e < 20 * JG; e += NW * 64) {
    const int a = e / JG, j = mt * JG + (e - a * JG);
    xs[e] = j < epi.L ? epi.x1[(long)a * epi.L + j] : 0.f;
  }
  const ccm_rsrc_t drsrc = ccm_make_rsrc(epi.dt);
  const uint32_t dstep = (uint32_t)((long)Lp * epi.lddt * 2), dlo = (uint32_t)(epi.lo_off * 2);
  for (int r = 0; r < ROUNDS; r++) {
#pragma unroll
    for (int rb = 0; rb < RB; rb++) {
      if (wm * RQ + rb / 4 != r) continue;
#pragma unroll
      for (int cb = 0; cb < CB; cb++) {
        const int R = (wn * CB + cb) * 32 + l32;
#pragma unroll
        for (int i = 0; i < 16; i++) {
          float* const bb = (i >> 2) == 0 ? b0 : (i >> 2) == 1 ? b1 : (i >> 2) == 2 ? b2 : b3;
          bb[R * ESTR + 16 * q + 4 * (i & 3) + (rb & 3)] = acc[rb][cb][i];
        }
      }
    }
    __syncthreads();
    fwd_softmax_round<JG, NW, ESTR>(b0, b1, b2, b3, xs, gsum, fsum, epi, mt, n0, Ncol, r, wave_u, lane,
                                    drsrc, dstep, dlo);
    __syncthreads();
  }
  fwd_softmax_partials<JG, NW * 64, ROUNDS>(gsum, fsum, epi, mt, n0, Ncol, tid);
}

// The padding tiles of the coupling conversion (convert_ws_pad_tile), run
// by workgroups past the forward's tiles, two tiles per workgroup: they are
// dispatched after every tile and take CUs that the forward's last, partial
// round of tiles leaves idle.
struct ccm_pad_job {
  const float* x2;  // x past the singles
  float* g2;        // g past the singles
  double* regpart;  // the conversion's partials: tile (bxp, by) at by * stride + tc + bxp
  float lam2;
  int tp, rows, stride, tc;
};

template <int WM, int WN, int RB, int CB>
__global__ __launch_bounds__(64 * WM * WN, 1) void k_sparse_fwd(
    const uint16_t* __restrict__ Ws, long lds, int Lp, int nsteps, const uint32_t* __restrict__ idx,
    long ldidx, int Ncol, uint32_t tiles_n, uint32_t tiles, uint32_t total, ccm_fwd_epi epi, ccm_pad_job pj) {
  static_assert(64 * WM * WN == 512, "two padding tiles per workgroup");
  if (blockIdx.x >= total) {
    const uint32_t q = 2u * (blockIdx.x - total) + (threadIdx.x >> 8);
    const bool ok = q < (uint32_t)(pj.tp * pj.rows);
    const int bxp = ok ? (int)(q % (uint32_t)pj.tp) : 0, by = ok ? (int)(q / (uint32_t)pj.tp) : 0;
    convert_ws_pad_tile(pj.x2, pj.g2, epi.L, pj.lam2, bxp, by, ok,
                        pj.regpart + (long)by * pj.stride + pj.tc + bxp);
    return;
  }
  const uint32_t qid = ccm_xcd_remap(blockIdx.x, total);
  if (qid >= tiles) return;
  sparse_fwd_tile<WM, WN, RB, CB>(Ws, lds, Lp, nsteps, idx, ldidx, Ncol, qid / tiles_n, qid % tiles_n, epi);
}

constexpr int SF_TM = 256, SF_TN = 256;

// The sparse fused forward over a group of alignments (whole alignments as
// one chunk each, x1 = x), tiles per alignment ceil(L / 12) x ceil(ncols /
// 256) in the single-alignment order, workgroups through the group's tile
// map when it has one
__global__ __launch_bounds__(512, 1) void k_sparse_fwd_grouped(const ccm_bprob* __restrict__ probs,
                                                               const uint32_t* __restrict__ offs,
                                                               const int32_t* __restrict__ pidx, int np,
                                                               uint32_t total, const uint32_t* __restrict__ qmap) {
  const uint32_t qid = qmap ? qmap[blockIdx.x] : ccm_xcd_remap(blockIdx.x, total);
  if (qid >= offs[np]) return;
  const int gi = ccm_group_find(offs, np, qid);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  const uint32_t t = qid - offs[gi];
  const uint32_t tiles_n = (uint32_t)((B.ncols + SF_TM - 1) / SF_TM);
  const ccm_fwd_epi epi{B.x, B.msapt, B.ldmt, B.w, B.L, B.N, B.dt, B.lddt, (long)B.ncols, B.passes, 1, 0,
                        B.fpart, B.g1part};
  sparse_fwd_tile<2, 4, 4, 2>(B.wt, B.ldwt, B.Lp, B.passes_fwd * SP_GROUPS * B.Lp / 16, B.sidx, B.ldsidx,
                              B.ncols, (int)(t / tiles_n), (int)(t % tiles_n), epi);
}

}  // namespace

extern "C" {

long ccm_sparse_ws_len(int Lp, int passes) { return (long)passes * SP_KC * Lp; }

int ccm_sparse_steps(int Lp, int passes) { return Lp % 64 ? 0 : passes * SP_GROUPS * Lp / 16; }

long ccm_sparse_index_ld(int Np) { return ((long)Np + SF_TM - 1) / SF_TM * SF_TM; }

int ccm_sparse_fwd_fits(int Lp, int passes_fwd, int passes, int ncols) {
  const long P = (long)CCM_NA * Lp;
  return (P + 1) * ccm_sparse_ws_len(Lp, passes_fwd) * 2 <= 0x7FFFFFFFL &&
         P * passes * (long)ncols * 2 <= 0x7FFFFFFFL;
}

int ccm_launch_sparse_index(const uint8_t* d_msap, int Lp, int Np, int passes, uint32_t* d_idx,
                            ccm_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  const int nsteps = ccm_sparse_steps(Lp, passes);
  if (!d_msap || !d_idx || nsteps <= 0 || Np < 1 || passes < 1 || passes > 2) return CCM_ERR_ARG;
  const long ld = ccm_sparse_index_ld(Np);
  const long total = (long)nsteps * 2 * ld;
  const int blocks = (int)((total + 255) / 256 < 65536 ? (total + 255) / 256 : 65536);
  hipLaunchKernelGGL(k_sparse_index, dim3(blocks), dim3(256), 0, s, d_msap, Lp, Np, passes, nsteps, d_idx, ld);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_convert_ws(const float* d_x, int L, int Lp, uint16_t* d_ws, long lds, int passes,
                          ccm_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  if (Lp % 64 || Lp < L || L < 1 || passes < 1 || passes > 2 || lds < ccm_sparse_ws_len(Lp, passes) ||
      lds % 8)
    return CCM_ERR_ARG;
  const dim3 grid(convert_ws_tiles_c(L, passes), convert_ws_rows(Lp));
  const long nsp = ccm_nsp(L);
  float* const g = nullptr;
  double* const rp = nullptr;
  if (passes == 1)
    hipLaunchKernelGGL(k_convert_ws<1>, grid, dim3(256), 0, s, d_x, nsp, L, Lp, d_ws, lds, g, 0.f, 0.f, rp, rp, 0);
  else
    hipLaunchKernelGGL(k_convert_ws<2>, grid, dim3(256), 0, s, d_x, nsp, L, Lp, d_ws, lds, g, 0.f, 0.f, rp, rp, 0);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_convert_ws_reg_parts(int L, int Lp, int passes) {
  if (L < 1 || Lp % 64 || Lp < L || passes < 1 || passes > 2) return 0;
  return (convert_ws_tiles_c(L, passes) + convert_ws_tiles_p(L)) * convert_ws_rows(Lp);
}

int ccm_launch_convert_ws_reg(const float* d_x, float* d_g, int L, int Lp, uint16_t* d_ws, long lds, int passes,
                              float lam1, float lam2, double* d_regpair, double* d_regsingle, int pad,
                              ccm_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  if (Lp % 64 || Lp < L || L < 1 || passes < 1 || passes > 2 || lds < ccm_sparse_ws_len(Lp, passes) ||
      lds % 8 || !d_x || !d_ws || !d_regpair || !d_regsingle || (pad && !d_g))
    return CCM_ERR_ARG;
  const int tp_run = pad ? convert_ws_tiles_p(L) : 0;
  const dim3 grid(convert_ws_tiles_c(L, passes) + tp_run + 1, convert_ws_rows(Lp));
  const long nsp = ccm_nsp(L);
  if (passes == 1)
    hipLaunchKernelGGL(k_convert_ws<1>, grid, dim3(256), 0, s, d_x, nsp, L, Lp, d_ws, lds, d_g, lam1, lam2,
                       d_regpair, d_regsingle, tp_run);
  else
    hipLaunchKernelGGL(k_convert_ws<2>, grid, dim3(256), 0, s, d_x, nsp, L, Lp, d_ws, lds, d_g, lam1, lam2,
                       d_regpair, d_regsingle, tp_run);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_sparse_fwd_softmax_pad(const uint16_t* d_ws, long lds, int Lp, int passes_fwd,
                                      const uint32_t* d_idx, long ldidx, int ncols, const float* d_x1,
                                      const uint8_t* d_msapt, long ldmt, const float* d_w, int L, int Nc,
                                      uint16_t* d_dt, long lddt, long lo_off, int passes, int mode, int blk0,
                                      double* d_fpart, double* d_g1part, const float* d_x, float* d_g,
                                      float lam2, double* d_regpair, ccm_stream_t stream) {
  hipStream_t s = (hipStream_t)stream;
  // the padding tiles of the conversion, when asked for
  ccm_pad_job pj{};
  uint32_t npad = 0;
  if (d_x || d_g || d_regpair) {
    if (!d_x || !d_g || !d_regpair || !(Lp % 64 == 0 && Lp >= L && L >= 1)) return CCM_ERR_ARG;
    const long nsp = ccm_nsp(L);
    const int tc = convert_ws_tiles_c(L, passes_fwd == 1 ? 1 : 2), tp = convert_ws_tiles_p(L);
    pj = ccm_pad_job{d_x + nsp, d_g + nsp, d_regpair, lam2, tp, convert_ws_rows(Lp), tc + tp, tc};
    npad = (uint32_t)((tp * convert_ws_rows(Lp) + 1) / 2);
  }
  const long P = (long)CCM_NA * Lp;
  const int nsteps = ccm_sparse_steps(Lp, passes_fwd);
  if (Lp % 64 || Lp < L || L < 1 || ncols % 128 || ncols <= 0 || Nc < 1 || Nc > ncols || nsteps <= 0 ||
      lds < ccm_sparse_ws_len(Lp, passes_fwd) || lds % 8 || !d_idx || ldidx < ncols || !d_fpart ||
      (mode && (!d_dt || !d_g1part)) || blk0 < 0)
    return CCM_ERR_ARG;
  if (mode && (passes < 1 || passes > 2 || lddt < (long)passes * ncols || (passes > 1 && lo_off < ncols)))
    return CCM_ERR_ARG;
  if (!ccm_sparse_fwd_fits(Lp, passes_fwd, mode ? passes : 1, mode ? ncols : 0)) return CCM_ERR_ARG;
  // the positions are read with 32-bit offsets from d_idx
  if ((long)nsteps * 2 * ldidx * 4 > 0x7FFFFFFFL) return CCM_ERR_ARG;
  ccm_fwd_epi epi{d_x1, d_msapt, ldmt, d_w, L, Nc, d_dt, lddt, lo_off, passes, mode, blk0, d_fpart, d_g1part};
  constexpr int JG = SF_TN / CCM_NA;
  const uint32_t tiles_n = (uint32_t)((ncols + SF_TM - 1) / SF_TM);
  const uint32_t tiles = (uint32_t)((L + JG - 1) / JG) * tiles_n, total = (tiles + 7u) & ~7u;
  // the sequence tiles read positions up to the next multiple of 256
  if (ldidx < (long)tiles_n * SF_TM) return CCM_ERR_ARG;
  hipLaunchKernelGGL((k_sparse_fwd<2, 4, 4, 2>), dim3(total + npad), dim3(512), 0, s, d_ws, lds, Lp, nsteps,
                     d_idx, ldidx, ncols, tiles_n, tiles, total, epi, pj);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_sparse_fwd_softmax(const uint16_t* d_ws, long lds, int Lp, int passes_fwd,
                                  const uint32_t* d_idx, long ldidx, int ncols, const float* d_x1,
                                  const uint8_t* d_msapt, long ldmt, const float* d_w, int L, int Nc,
                                  uint16_t* d_dt, long lddt, long lo_off, int passes, int mode, int blk0,
                                  double* d_fpart, double* d_g1part, ccm_stream_t stream) {
  return ccm_launch_sparse_fwd_softmax_pad(d_ws, lds, Lp, passes_fwd, d_idx, ldidx, ncols, d_x1, d_msapt, ldmt,
                                           d_w, L, Nc, d_dt, lddt, lo_off, passes, mode, blk0, d_fpart, d_g1part,
                                           nullptr, nullptr, 0.f, nullptr, stream);
}

// grouped forms (batched alignments; descriptors in ccmpred_amd_batch.h)
int ccm_convert_ws_blocks(int L, int Lp, int passes) {
  return (convert_ws_tiles_c(L, passes) + convert_ws_tiles_p(L) + 1) * convert_ws_rows(Lp);
}

uint32_t ccm_sparse_fwd_tile_rows(int L) { return L < 1 ? 0u : (uint32_t)((L + SF_TN / CCM_NA - 1) / (SF_TN / CCM_NA)); }

uint32_t ccm_sparse_fwd_tile_count(int L, int ncols) {
  return ncols < 1 ? 0u : ccm_sparse_fwd_tile_rows(L) * (uint32_t)((ncols + SF_TM - 1) / SF_TM);
}

int ccm_launch_convert_ws_grouped(const ccm_bprob* d_probs, ccm_group g, int passes, ccm_stream_t stream) {
  if (g.np <= 0 || g.total == 0 || passes < 1 || passes > 2) return CCM_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (passes == 1)
    hipLaunchKernelGGL(k_convert_ws_grouped<1>, dim3(g.total), dim3(256), 0, s, d_probs, g.offs, g.pidx, g.np);
  else
    hipLaunchKernelGGL(k_convert_ws_grouped<2>, dim3(g.total), dim3(256), 0, s, d_probs, g.offs, g.pidx, g.np);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_sparse_fwd_softmax_grouped(const ccm_bprob* d_probs, ccm_group g, ccm_stream_t stream) {
  if (g.np <= 0 || g.total == 0) return CCM_ERR_ARG;
  // with a tile map the grid is the map; else padded to whole XCD rounds
  const uint32_t total = g.qmap ? g.total : (g.total + 7u) & ~7u;
  hipLaunchKernelGGL(k_sparse_fwd_grouped, dim3(total), dim3(512), 0, (hipStream_t)stream, d_probs, g.offs,
                     g.pidx, g.np, total, g.qmap);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"
