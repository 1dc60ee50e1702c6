// Device vector arithmetic of the conjugate-gradient loop.  All dot
// products accumulate FP32 products in FP64 with a fixed reduction order
// (per-block partials, then one block sums the partials), so repeated runs
// give bitwise-identical scalars.
#include "common.h"

#define VO_BLOCKS 1024

// up to three dot products sharing the first operand: out[i] = <a, b_i>
__device__ __forceinline__ void dots_partial_block(const float* __restrict__ a,
                                                   const float* __restrict__ b0,
                                                   const float* __restrict__ b1,
                                                   const float* __restrict__ b2, long n, int nd,
                                                   double* __restrict__ part, uint32_t blk,
                                                   uint32_t nblk) {
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

// ---- batched conjugate gradient: every alignment of a group per launch ----

// <g,s>, <g,g>, <g,gprev> of each evaluated alignment: n_dots blocks each,
// the same partition and order as ccm_launch_dots over nvp elements
__global__ __launch_bounds__(256) void k_dots_partial_grouped(const ccm_bprob* __restrict__ probs,
                                                              const uint32_t* __restrict__ offs,
                                                              const int32_t* __restrict__ pidx, int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  if (!B.ctl->eval_on) return;
  dots_partial_block(B.g, B.s, B.g, B.gprev, B.nvp, 3, B.dpart, blockIdx.x - offs[gi], (uint32_t)B.n_dots);
}

__global__ __launch_bounds__(256) void k_dots_final_grouped(const ccm_bprob* __restrict__ probs,
                                                            const int32_t* __restrict__ pidx) {
  const ccm_bprob& B = probs[pidx[blockIdx.x]];
  if (!B.ctl->eval_on) return;
  dots_final_block(B.dpart, B.n_dots, 3, B.ctl->dots);
}

// The vector step of a round, as its control block says:
//   x += c_pre s;  s = per vec_mode (gprev = g where the mode moves on);
//   x += c_trial s
// with the same FP32 arithmetic as k_axpy / k_direction / k_negate.
__global__ __launch_bounds__(256) void k_cg_vec_grouped(const ccm_bprob* __restrict__ probs,
                                                        const uint32_t* __restrict__ offs,
                                                        const int32_t* __restrict__ pidx, int np) {
  const int gi = ccm_group_find(offs, np, blockIdx.x);
  const ccm_bprob& B = probs[pidx[gi]];
  const ccm_cg_ctl* c = B.ctl;
  if (!c->vec_on) return;
  const int mode = c->vec_mode;
  const float cpre = (float)c->c_pre, ctr = (float)c->c_trial, beta = (float)c->beta;
  const uint32_t blk = blockIdx.x - offs[gi], nblk = offs[gi + 1] - offs[gi];
  float* __restrict__ x = B.x;
  float* __restrict__ sd = B.s;
  float* __restrict__ gp = B.gprev;
  const float* __restrict__ g = B.g;
  const long n = B.nvp;
  for (long i = (long)blk * 256 + threadIdx.x; i < n; i += (long)nblk * 256) {
    float xi = x[i], si = sd[i];
    if (cpre != 0.0f) xi += cpre * si;
    if (mode == CCM_VEC_NEGATE) {
      const float gi2 = g[i];
      si = -gi2;
      gp[i] = gi2;
    } else if (mode == CCM_VEC_DIRECTION) {
      const float gi2 = g[i];
      si = -gi2 + beta * si;
      gp[i] = gi2;
    } else if (mode == CCM_VEC_RESTART) {
      si = -gp[i];
    }
    if (ctr != 0.0f) xi += ctr * si;
    x[i] = xi;
    if (mode != CCM_VEC_KEEP) sd[i] = si;
  }
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
  } else if (!(ft <= c->f0 + p.ftol * c->alpha * c->dg0)) {
    // not enough decrease: shrink, bisecting once bracketed
    c->hi = c->alpha;
    c->alpha = c->lo > 0.0 ? 0.5 * (c->lo + c->hi) : c->alpha * p.alpha_mul;
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
    if (p.k > 0 && c->it > p.k && c->it - 1 - p.k < log_cap) {
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

}  // extern "C"
