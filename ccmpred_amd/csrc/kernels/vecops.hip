// Device vector arithmetic of the conjugate-gradient loop.  All dot
// products accumulate FP32 products in FP64 with a fixed reduction order
// (per-block partials, then one block sums the partials), so repeated runs
// give bitwise-identical scalars.
#include "common.h"

#define VO_BLOCKS 1024

// The padded parameter layout (ccmpred.c:144-151) in float4 units: ns4
// float4s of singles, then one row of row4 = 8 L float4s per (b, k) whose
// first act4 = ceil(21 L / 4) hold the pair entries a < 21 (the last one
// may also hold a few a = 21 entries); the rest of the row is a >= 21
// padding.  There the gradient is 2 lambda_pair x (evaluate_cpu.c:158-161),
// so where x starts at zero in that padding (ccmpred.c:497 clears it) the
// gradient, the CG direction and the previous gradient stay exactly zero
// for the whole solve: a dot product whose first operand is the gradient
// only adds exact zeros there, and a vector step leaves x unchanged, so
// skipping the padding changes no bit of either.  row4 == 0 means no
// layout (every float4 is read): so it is when ccm_launch_pad_scan found a
// nonzero value in the skipped float4s.  Indices past the singles fit 32
// bits whenever row4 is set.
struct ccm_pad_rows {
  long ns4 = 0;
  uint32_t row4 = 0, act4 = 0, nrows = 0;
  __device__ __forceinline__ bool skip(long i) const {
    return row4 != 0 && i >= ns4 && (uint32_t)(i - ns4) % row4 >= act4;
  }
  // float4 index of the q-th float4 that is not skipped
  __device__ __forceinline__ long active(long q) const {
    if (q < ns4) return q;
    const uint32_t t = (uint32_t)(q - ns4), r = t / act4;
    return ns4 + (long)r * row4 + (t - r * act4);
  }
  __host__ __device__ __forceinline__ long n_active() const { return ns4 + (long)nrows * act4; }
};

// the layout of an alignment of L columns (nsp singles, nvp floats in all);
// none where nvp is not that layout or its indices do not fit
__host__ __device__ __forceinline__ ccm_pad_rows ccm_pad_rows_make(int L, long nsp, long nvp) {
  ccm_pad_rows p;
  if ((nsp & 3) == 0 && (nvp & 3) == 0 && L > 0 && nsp >= (long)L * (CCM_NA - 1) &&
      nvp == nsp + (long)L * L * CCM_NA * CCM_NAP && (nvp >> 2) - (nsp >> 2) < (1L << 32)) {
    p.ns4 = nsp >> 2;
    p.row4 = (uint32_t)L * (CCM_NAP / 4);
    p.act4 = ((uint32_t)L * CCM_NA + 3) / 4;
    p.nrows = (uint32_t)L * CCM_NA;
  }
  return p;
}

// the layout of alignment B, checking its padding flag when scan is false
template <bool scan = false>
__device__ __forceinline__ ccm_pad_rows ccm_pad_rows_of(const ccm_bprob& B) {
  return (scan || !B.ctl->x_pad_nz) ? ccm_pad_rows_make(B.L, B.nsp, B.nvp) : ccm_pad_rows{};
}

// up to three dot products sharing the first operand: out[i] = <a, b_i>
__device__ __forceinline__ void dots_partial_block(const float* __restrict__ a,
                                                   const float* __restrict__ b0,
                                                   const float* __restrict__ b1,
                                                   const float* __restrict__ b2, long n, int nd,
                                                   double* __restrict__ part, uint32_t blk,
                                                   uint32_t nblk, const ccm_pad_rows pad = {}) {
  __shared__ double red[3][4];
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  // 16-byte loads over the first 4*(n/4) elements when every vector is
  // 16-byte aligned (else all elements take the scalar loop), then the
  // tail; a second operand equal to a is not read again
  const bool b1a = b1 == a;
  const bool al = (((uintptr_t)a | (uintptr_t)b0 | (uintptr_t)(nd > 1 ? b1 : a) |
                    (uintptr_t)(nd > 2 ? b2 : a)) & 15u) == 0;
  const long n4 = al ? n >> 2 : 0, stride = (long)nblk * 256;
  const float4* a4 = reinterpret_cast<const float4*>(a);
  const float4* b04 = reinterpret_cast<const float4*>(b0);
  const float4* b14 = reinterpret_cast<const float4*>(b1);
  const float4* b24 = reinterpret_cast<const float4*>(b2);
  for (long i = (long)blk * 256 + threadIdx.x; i < n4; i += stride) {
    if (pad.skip(i)) continue;
    const float4 av = a4[i], v0 = b04[i];
    s0 += (double)av.x * (double)v0.x + (double)av.y * (double)v0.y + (double)av.z * (double)v0.z +
          (double)av.w * (double)v0.w;
    if (nd > 1) {
      const float4 v1 = b1a ? av : b14[i];
      s1 += (double)av.x * (double)v1.x + (double)av.y * (double)v1.y + (double)av.z * (double)v1.z +
            (double)av.w * (double)v1.w;
    }
    if (nd > 2) {
      const float4 v2 = b24[i];
      s2 += (double)av.x * (double)v2.x + (double)av.y * (double)v2.y + (double)av.z * (double)v2.z +
            (double)av.w * (double)v2.w;
    }
  }
  for (long i = 4 * n4 + (long)blk * 256 + threadIdx.x; i < n; i += stride) {
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
    part[(long)d * nblk + blk] = red[d][0] + red[d][1] + red[d][2] + red[d][3];
  }
}

__global__ __launch_bounds__(256) void k_dots_partial(const float* __restrict__ a,
                                                      const float* __restrict__ b0,
                                                      const float* __restrict__ b1,
                                                      const float* __restrict__ b2, long n, int nd,
                                                      double* __restrict__ part) {
  dots_partial_block(a, b0, b1, b2, n, nd, part, blockIdx.x, gridDim.x);
}

// fixed-order sum of each dot's block partials: thread t adds partials t,
// t + 256, ... (all loads issued first), then a wave tree and the four
// wave sums in order
__device__ __forceinline__ void dots_final_block(const double* __restrict__ part, int nblk, int nd,
                                                 double* __restrict__ out) {
  __shared__ double sh[3][4];
  double s[3] = {0.0, 0.0, 0.0};
  for (int i0 = threadIdx.x; i0 < nblk; i0 += 256 * 4) {
    double v[3][4];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int i = i0 + u * 256;
        v[d][u] = (d < nd && i < nblk) ? part[(long)d * nblk + i] : 0.0;
      }
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int u = 0; u < 4; u++) s[d] += v[d][u];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    s[d] = ccm_wave_sum_f64(s[d]);
    if (lane == 0) sh[d][wv] = s[d];
  }
  __syncthreads();
  if (threadIdx.x < nd) {
    const int d = threadIdx.x;
    out[d] = ((sh[d][0] + sh[d][1]) + sh[d][2]) + sh[d][3];
  }
}

__global__ __launch_bounds__(256) void k_dots_final(const double* __restrict__ part, int nblk, int nd,
                                                    double* __restrict__ out) {
  dots_final_block(part, nblk, nd, out);
}

// x += alpha * s (one rounding)
__global__ void k_axpy(float* __restrict__ x, const float* __restrict__ s, float alpha, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = __builtin_fmaf(alpha, s[i], x[i]);
}

// s = -g + beta * s (one rounding)
__global__ void k_direction(float* __restrict__ s, const float* __restrict__ g, float beta, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = __builtin_fmaf(beta, s[i], -g[i]);
}

// s = -g
__global__ void k_negate(float* __restrict__ s, const float* __restrict__ g, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i] = -g[i];
}

// ---- batched conjugate gradient: every alignment of a group per launch ----

// <g,s>, <g,g>, <g,gprev> of each evaluated alignment: n_dots blocks each,
// the same partition and order as ccm_launch_dots over nvp elements (the
// skipped padding adds only zeros)
__global__ __launch_bounds__(256) void k_dots_partial_grouped(const ccm_bprob* __restrict__ probs,
                                                              const uint32_t* __restrict__ offs,
                                                              const int32_t* __restrict__ pidx, int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  dots_partial_block(B.g, B.s, B.g, B.gprev, B.nvp, 3, B.dpart, blockIdx.x - offs[gi], (uint32_t)B.n_dots,
                     ccm_pad_rows_of(B));
}

__global__ __launch_bounds__(256) void k_dots_final_grouped(const ccm_bprob* __restrict__ probs,
                                                            const int32_t* __restrict__ pidx) {
  const ccm_bprob& B = probs[pidx[blockIdx.x]];
  if (!B.ctl->eval_on) return;
  dots_final_block(B.dpart, B.n_dots, 3, B.ctl->dots);
}

// One element of the vector step (one rounding per multiply-add, as in
// k_axpy / k_direction / k_negate); gpi is read only by CCM_VEC_RESTART.
template <int MODE>
__device__ __forceinline__ void cg_vec_elem(float& xi, float& si, float gi, float gpi, float cpre, float ctr,
                                            float beta) {
  if (cpre != 0.0f) xi = __builtin_fmaf(cpre, si, xi);
  if constexpr (MODE == CCM_VEC_NEGATE)
    si = -gi;
  else if constexpr (MODE == CCM_VEC_DIRECTION)
    si = __builtin_fmaf(beta, si, -gi);
  else if constexpr (MODE == CCM_VEC_RESTART)
    si = -gpi;
  if (ctr != 0.0f) xi = __builtin_fmaf(ctr, si, xi);
}

// the vector step of one alignment with the step kind fixed (a branch on
// the kind inside the loop was compiled wrongly for one of the kinds)
template <int MODE>
__device__ __forceinline__ void cg_vec_block(const ccm_bprob& B, float cpre, float ctr, float beta,
                                             uint32_t blk, uint32_t nblk) {
  constexpr bool moves = MODE == CCM_VEC_NEGATE || MODE == CCM_VEC_DIRECTION;
  float* __restrict__ x = B.x;
  float* __restrict__ sd = B.s;
  float* __restrict__ gp = B.gprev;
  const float* __restrict__ g = B.g;
  const ccm_pad_rows pad = ccm_pad_rows_of(B);
  if (pad.row4 != 0 && (((uintptr_t)x | (uintptr_t)sd | (uintptr_t)gp | (uintptr_t)g) & 15u) == 0) {
    float4* __restrict__ x4 = reinterpret_cast<float4*>(x);
    float4* __restrict__ s4 = reinterpret_cast<float4*>(sd);
    float4* __restrict__ gp4 = reinterpret_cast<float4*>(gp);
    const float4* __restrict__ g4 = reinterpret_cast<const float4*>(g);
    const long nq = pad.n_active();
    for (long q = (long)blk * 256 + threadIdx.x; q < nq; q += (long)nblk * 256) {
      const long i = pad.active(q);
      float4 xv = x4[i], sv = s4[i];
      float4 gv = make_float4(0.f, 0.f, 0.f, 0.f), pv = gv;
      if constexpr (moves) gv = g4[i];
      if constexpr (MODE == CCM_VEC_RESTART) pv = gp4[i];
      cg_vec_elem<MODE>(xv.x, sv.x, gv.x, pv.x, cpre, ctr, beta);
      cg_vec_elem<MODE>(xv.y, sv.y, gv.y, pv.y, cpre, ctr, beta);
      cg_vec_elem<MODE>(xv.z, sv.z, gv.z, pv.z, cpre, ctr, beta);
      cg_vec_elem<MODE>(xv.w, sv.w, gv.w, pv.w, cpre, ctr, beta);
      x4[i] = xv;
      if constexpr (MODE != CCM_VEC_KEEP) s4[i] = sv;
      if constexpr (moves) gp4[i] = gv;
    }
    return;
  }
  const long n = B.nvp;
  for (long i = (long)blk * 256 + threadIdx.x; i < n; i += (long)nblk * 256) {
    float xi = x[i], si = sd[i];
    float gi2 = 0.0f, gpi = 0.0f;
    if constexpr (moves) gi2 = g[i];
    if constexpr (MODE == CCM_VEC_RESTART) gpi = gp[i];
    cg_vec_elem<MODE>(xi, si, gi2, gpi, cpre, ctr, beta);
    x[i] = xi;
    if constexpr (MODE != CCM_VEC_KEEP) sd[i] = si;
    if constexpr (moves) gp[i] = gi2;
  }
}

// The vector step of a round, as its control block says:
//   x += c_pre s;  s = per vec_mode (gprev = g where the mode moves on);
//   x += c_trial s
// with the same FP32 arithmetic as k_axpy / k_direction / k_negate.  With
// the padded layout, only the float4s that hold singles and pair entries
// a < 21 are visited (the padding of x is left as it is, which is what the
// step computes there from a zero direction).
__global__ __launch_bounds__(256) void k_cg_vec_grouped(const ccm_bprob* __restrict__ probs,
                                                        const uint32_t* __restrict__ offs,
                                                        const int32_t* __restrict__ pidx, int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  const ccm_cg_ctl* c = B.ctl;
  if (!c->vec_on) return;
  const float cpre = (float)c->c_pre, ctr = (float)c->c_trial, beta = (float)c->beta;
  const uint32_t blk = blockIdx.x - offs[gi], nblk = offs[gi + 1] - offs[gi];
  switch (c->vec_mode) {
    case CCM_VEC_NEGATE: cg_vec_block<CCM_VEC_NEGATE>(B, cpre, ctr, beta, blk, nblk); break;
    case CCM_VEC_DIRECTION: cg_vec_block<CCM_VEC_DIRECTION>(B, cpre, ctr, beta, blk, nblk); break;
    case CCM_VEC_RESTART: cg_vec_block<CCM_VEC_RESTART>(B, cpre, ctr, beta, blk, nblk); break;
    default: cg_vec_block<CCM_VEC_KEEP>(B, cpre, ctr, beta, blk, nblk); break;
  }
}

// ctl->x_pad_nz = 1 where x has a nonzero value in a float4 that the
// padded layout skips; blockIdx.y is the alignment
__device__ __forceinline__ void pad_scan_block(const float* __restrict__ x, const ccm_pad_rows& pad,
                                               int32_t* __restrict__ flag) {
  if (pad.row4 == 0 || pad.act4 >= pad.row4 || ((uintptr_t)x & 15u) != 0) return;
  const uint32_t skip4 = pad.row4 - pad.act4;
  const uint32_t n = pad.nrows * skip4;  // below the 2^32 the layout guarantees
  const float4* __restrict__ x4 = reinterpret_cast<const float4*>(x);
  bool nz = false;
  for (uint32_t t = blockIdx.x * 256 + threadIdx.x; t < n; t += gridDim.x * 256) {
    const uint32_t r = t / skip4, c = t - r * skip4;
    const float4 v = x4[pad.ns4 + (long)r * pad.row4 + pad.act4 + c];
    nz |= v.x != 0.0f || v.y != 0.0f || v.z != 0.0f || v.w != 0.0f;
  }
  if (nz) *flag = 1;
}

__global__ __launch_bounds__(256) void k_pad_scan(const ccm_bprob* __restrict__ probs) {
  const ccm_bprob& B = probs[blockIdx.y];
  pad_scan_block(B.x, ccm_pad_rows_of<true>(B), &B.ctl->x_pad_nz);
}

__global__ __launch_bounds__(256) void k_pad_scan_x(const float* __restrict__ x, ccm_pad_rows pad,
                                                    int32_t* __restrict__ flag) {
  pad_scan_block(x, pad, flag);
}

// ---- the single-alignment vector steps over the padded layout: only the
// float4s that hold singles or pair entries a < 21 (same arithmetic as
// k_axpy / k_direction / k_negate, which the padding leaves unchanged when
// its direction and gradient are zero)
__global__ __launch_bounds__(256) void k_dots_partial_pad(const float* __restrict__ a,
                                                          const float* __restrict__ b0,
                                                          const float* __restrict__ b1,
                                                          const float* __restrict__ b2, long n, int nd,
                                                          double* __restrict__ part, ccm_pad_rows pad) {
  dots_partial_block(a, b0, b1, b2, n, nd, part, blockIdx.x, gridDim.x, pad);
}

// MODE 0: x += alpha s;  1: s = -g + alpha s;  2: s = -g  (y is x or s, v
// is s or g).  The step is a template argument: with a three-way branch on
// a kernel argument the compiler stored uncomputed registers for mode 2.
template <int MODE>
__global__ __launch_bounds__(256) void k_vec_pad(float* __restrict__ y, const float* __restrict__ v,
                                                 float alpha, ccm_pad_rows pad) {
  float4* __restrict__ y4 = reinterpret_cast<float4*>(y);
  const float4* __restrict__ v4 = reinterpret_cast<const float4*>(v);
  const long nq = pad.n_active();
  for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < nq; q += (long)gridDim.x * 256) {
    const long i = pad.active(q);
    const float4 vv = v4[i];
    float4 yv;
    if constexpr (MODE == 0) {
      yv = y4[i];
      yv.x = __builtin_fmaf(alpha, vv.x, yv.x);
      yv.y = __builtin_fmaf(alpha, vv.y, yv.y);
      yv.z = __builtin_fmaf(alpha, vv.z, yv.z);
      yv.w = __builtin_fmaf(alpha, vv.w, yv.w);
    } else if constexpr (MODE == 1) {
      yv = y4[i];
      yv.x = __builtin_fmaf(alpha, yv.x, -vv.x);
      yv.y = __builtin_fmaf(alpha, yv.y, -vv.y);
      yv.z = __builtin_fmaf(alpha, yv.z, -vv.z);
      yv.w = __builtin_fmaf(alpha, yv.w, -vv.w);
    } else {
      yv = make_float4(-vv.x, -vv.y, -vv.z, -vv.w);
    }
    y4[i] = yv;
  }
}

// The sufficient-decrease test and the shrinking step of the host loop in
// cg_device.c (cg_decrease, cg_shrink), operation for operation: the Armijo
// test, or below the objective's relative resolution fres the approximate
// Armijo test on the slope (Hager and Zhang 2005); after a failure with no
// lower bracket, the minimiser of the quadratic through f0, dg0 and ft
// within [0.1, alpha_mul] alpha.
__device__ bool cg_decrease(double ft, double f0, double alpha, double dg0, double dgt, double ftol,
                            double fres) {
#pragma clang fp contract(off)
  if (ft <= f0 + ftol * alpha * dg0) return true;
  return ft <= f0 + fres * fabs(f0) && dgt <= (2.0 * ftol - 1.0) * dg0;
}

__device__ double cg_shrink(double ft, double f0, double alpha, double dg0, double alpha_mul) {
#pragma clang fp contract(off)
  const double den = 2.0 * ((ft - f0) - dg0 * alpha);
  const double aq = (-dg0 * alpha) * alpha / den;
  const double amin = 0.1 * alpha, amax = alpha_mul * alpha;
  return !(aq >= amin) ? amin : (aq > amax ? amax : aq);
}

// Line search and direction control of one alignment after an evaluation
// at x0 + moved s: the decisions of the host loop in cg_device.c, in the
// same order and FP64 arithmetic (f rounded to float as evaluate returns
// it), written as the next round's vector step.
__device__ void cg_control(ccm_cg_ctl* c, double* log, int log_cap, const ccm_cg_prm& p) {
  // no fused multiply-adds: every test rounds as the host's FP64 code does
#pragma clang fp contract(off)
  const double ft = (double)(float)c->fx;
  const double dgt = c->dots[0], ggn = c->dots[1], ggp = c->dots[2];
  c->nevals++;
  c->c_pre = 0.0;
  c->c_trial = 0.0;
  c->vec_mode = CCM_VEC_KEEP;
  c->vec_on = 1;
  c->eval_on = 1;
  if (isnan(ft)) {
    c->status = -2;  // CONJUGRADERR_DEVICE
    c->done = 1;
    c->vec_on = 0;
    c->eval_on = 0;
    return;
  }
  bool accept = false, exhausted = false;
  if (c->phase == CCM_CG_INIT) {
    c->f = ft;
    if (ggn == 0.0) {
      c->status = 1;  // CONJUGRAD_ALREADY_MINIMIZED
      c->done = 1;
      c->vec_on = 0;
      c->eval_on = 0;
      return;
    }
    c->status = 2;  // CONJUGRADERR_MAXIMUMITERATION until something else ends the run
    c->gg = ggn;
    c->vec_mode = CCM_VEC_NEGATE;
    c->dg0 = -ggn;
    c->alpha = 1.0 / sqrt(ggn);
    c->steepest = 1;
  } else if (c->phase == CCM_CG_FORCED) {
    accept = true;
  } else if (!cg_decrease(ft, c->f0, c->alpha, c->dg0, dgt, p.ftol, p.fres)) {
    // not enough decrease: shrink, bisecting once bracketed
    c->hi = c->alpha;
    c->alpha = c->lo > 0.0 ? 0.5 * (c->lo + c->hi) : cg_shrink(ft, c->f0, c->alpha, c->dg0, p.alpha_mul);
    c->ls++;
    if (c->ls > p.max_linesearch) exhausted = true;
  } else if (fabs(dgt) <= -p.wolfe * c->dg0 || c->ls == p.max_linesearch) {
    accept = true;
  } else {
    if (dgt > 0.0)
      c->hi = c->alpha;
    else
      c->lo = c->alpha;
    c->alpha = isinf(c->hi) ? c->alpha / p.alpha_mul : 0.5 * (c->lo + c->hi);
    c->ls++;
  }

  if (exhausted) {
    if (c->lo > 0.0) {
      // return to the longest step that passed the decrease test
      c->c_trial = c->lo - c->moved;
      c->moved = c->alpha = c->lo;
      c->phase = CCM_CG_FORCED;
      return;
    }
    c->c_pre = -c->moved;
    c->f = c->f0;
    if (c->steepest) {
      c->status = -1;  // CONJUGRADERR_MAXIMUMLINESEARCH
      c->done = 1;
      c->eval_on = 0;  // the restoring vector step still runs
      return;
    }
    c->vec_mode = CCM_VEC_RESTART;
    c->dg0 = -c->gg;
    c->alpha = 1.0 / sqrt(c->gg);
    c->steepest = 1;
  } else if (accept) {
    c->f = ft;
    if (c->it < log_cap) {
      double* r = log + 4L * c->it;
      r[0] = ft;
      r[1] = c->alpha;
      r[2] = (double)c->ls;
      r[3] = ggn;
    }
    c->it++;
    if (p.epsilon > 0.0 && p.k > 0 && c->it > p.k && c->it - 1 - p.k < log_cap) {
      const double before = log[4L * (c->it - 1 - p.k)];
      if ((before - ft) / fabs(ft) < p.epsilon) {
        c->status = 0;  // CONJUGRAD_SUCCESS
        c->done = 1;
        c->vec_on = 0;
        c->eval_on = 0;
        return;
      }
    }
    if (c->it >= p.max_iterations) {
      c->done = 1;
      c->vec_on = 0;
      c->eval_on = 0;
      return;
    }
    double beta = (ggn - ggp) / c->gg;
    if (beta < 0) beta = 0;
    double dgn = -ggn + beta * dgt;
    if (dgn >= 0) {
      c->vec_mode = CCM_VEC_NEGATE;
      dgn = -ggn;
      c->steepest = 1;
    } else {
      c->vec_mode = CCM_VEC_DIRECTION;
      c->steepest = beta == 0.0;
    }
    c->beta = beta;
    c->alpha = c->alpha * c->dg0 / dgn;
    c->dg0 = dgn;
    c->gg = ggn;
  } else if (c->phase != CCM_CG_INIT) {
    // next trial of the same search
    c->c_trial = c->alpha - c->moved;
    c->moved = c->alpha;
    return;
  }
  // a new line search from the current point along the (new) direction
  c->f0 = c->f;
  c->lo = 0.0;
  c->hi = __builtin_inf();
  c->ls = 1;
  c->c_trial = c->alpha;
  c->moved = c->alpha;
  c->phase = CCM_CG_SEARCH;
}

// one thread per alignment; *nactive = alignments with work left
__global__ __launch_bounds__(256) void k_cg_ctrl(ccm_bprob* __restrict__ probs, int nprob, ccm_cg_prm prm,
                                                 int32_t* __restrict__ nactive) {
  __shared__ int cnt;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < nprob; i += 256) {
    ccm_cg_ctl* c = probs[i].ctl;
    if (c->done) {
      c->vec_on = 0;
      c->eval_on = 0;
    } else if (c->eval_on) {
      cg_control(c, probs[i].log, probs[i].log_cap, prm);
    }
    if (!c->done || c->vec_on) atomicAdd(&cnt, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) *nactive = cnt;
}

extern "C" {

int ccm_dots_partial_len(void) { return 3 * VO_BLOCKS; }

// partial blocks of ccm_launch_dots over n elements (the batched dots use the same)
int ccm_dots_blocks(long n) {
  const long want = (n + 255) / 256;
  return (int)(want < VO_BLOCKS ? want : VO_BLOCKS);
}

// vector-step blocks of one alignment in the batched form
int ccm_cg_vec_blocks(long n) {
  const long want = (n + 255) / 256;
  return (int)(want < 2048 ? want : 2048);
}

int ccm_launch_dots_grouped(const ccm_bprob* d_probs, ccm_group g, ccm_group gfinal, hipStream_t s) {
  if (g.np <= 0 || g.total == 0 || gfinal.np <= 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_dots_partial_grouped, dim3(g.total), dim3(256), 0, s, d_probs, g.offs, g.pidx, g.np);
  CCM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dots_final_grouped, dim3(gfinal.np), dim3(256), 0, s, d_probs, gfinal.pidx);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_cg_vec_grouped(const ccm_bprob* d_probs, ccm_group g, hipStream_t s) {
  if (g.np <= 0 || g.total == 0) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_cg_vec_grouped, dim3(g.total), dim3(256), 0, s, d_probs, g.offs, g.pidx, g.np);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_pad_scan(const ccm_bprob* d_probs, int nprob, hipStream_t s) {
  if (nprob <= 0 || nprob > 65535 || !d_probs) return CCM_ERR_ARG;  // one grid row per alignment
  hipLaunchKernelGGL(k_pad_scan, dim3(256, (unsigned)nprob), dim3(256), 0, s, d_probs);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_cg_ctrl(ccm_bprob* d_probs, int nprob, const ccm_cg_prm* prm, int32_t* d_nactive,
                       hipStream_t s) {
  if (nprob <= 0 || !prm) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_cg_ctrl, dim3(1), dim3(256), 0, s, d_probs, nprob, *prm, d_nactive);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

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

// ---- single alignment, padded layout of ncol columns (n = nvar_padded)

static bool pad_layout(int ncol, long n, ccm_pad_rows* p) {
  if (ncol <= 0) return false;
  const long ns = (long)ncol * (CCM_NA - 1);
  const long nsp = ns + CCM_NAP - ns % CCM_NAP;
  *p = ccm_pad_rows_make(ncol, nsp, n);
  return p->row4 != 0;
}

static bool aligned16(const void* a, const void* b) { return ((((uintptr_t)a) | ((uintptr_t)b)) & 15u) == 0; }

static unsigned pad_grid(const ccm_pad_rows& p) {
  const long want = (p.n_active() + 255) / 256;
  return (unsigned)(want < 8192 ? want : 8192);
}

int ccm_launch_pad_scan_x(const float* x, int ncol, long n, int32_t* d_flag, hipStream_t s) {
  ccm_pad_rows p;
  if (!x || !d_flag || !pad_layout(ncol, n, &p)) return CCM_ERR_ARG;
  hipLaunchKernelGGL(k_pad_scan_x, dim3(1024), dim3(256), 0, s, x, p, d_flag);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_dots_padded(const float* a, const float* b0, const float* b1, const float* b2, long n,
                           int nd, int ncol, double* part, double* out, hipStream_t s) {
  ccm_pad_rows p;
  if (nd < 1 || nd > 3 || n <= 0) return CCM_ERR_ARG;
  if (!pad_layout(ncol, n, &p)) return ccm_launch_dots(a, b0, b1, b2, n, nd, part, out, s);
  const int nblk = ccm_dots_blocks(n);
  hipLaunchKernelGGL(k_dots_partial_pad, dim3(nblk), dim3(256), 0, s, a, b0, b1 ? b1 : b0, b2 ? b2 : b0, n,
                     nd, part, p);
  CCM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_dots_final, dim3(1), dim3(256), 0, s, part, nblk, nd, out);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_axpy_padded(float* x, const float* sdir, float alpha, int ncol, long n, hipStream_t s) {
  ccm_pad_rows p;
  if (!pad_layout(ncol, n, &p) || !aligned16(x, sdir)) return ccm_launch_axpy(x, sdir, alpha, n, s);
  hipLaunchKernelGGL(k_vec_pad<0>, dim3(pad_grid(p)), dim3(256), 0, s, x, sdir, alpha, p);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_direction_padded(float* sdir, const float* g, float beta, int ncol, long n, hipStream_t s) {
  ccm_pad_rows p;
  if (!pad_layout(ncol, n, &p) || !aligned16(sdir, g)) return ccm_launch_direction(sdir, g, beta, n, s);
  hipLaunchKernelGGL(k_vec_pad<1>, dim3(pad_grid(p)), dim3(256), 0, s, sdir, g, beta, p);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

int ccm_launch_negate_padded(float* sdir, const float* g, int ncol, long n, hipStream_t s) {
  ccm_pad_rows p;
  if (!pad_layout(ncol, n, &p) || !aligned16(sdir, g)) return ccm_launch_negate(sdir, g, n, s);
  hipLaunchKernelGGL(k_vec_pad<2>, dim3(pad_grid(p)), dim3(256), 0, s, sdir, g, 0.0f, p);
  CCM_LAUNCH_CHECK();
  return CCM_OK;
}

}  // extern "C"
