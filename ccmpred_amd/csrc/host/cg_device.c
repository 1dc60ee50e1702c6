/*
 * Nonlinear conjugate gradient on device-resident vectors, with the call
 * shape of libconjugrad's conjugrad_gpu (ccmpred.c:526-534, 619-648).
 * libconjugrad is not vendored with the reference, so this is an own
 * design within the reference's parameter set:
 *
 *   s0 = -g0, alpha0 = 1/|g0|
 *   per iteration, line search along s from x (x moves in place), at most
 *   max_linesearch trials:
 *     x += (alpha - moved) s ; f_t, g_t = evaluate(x)
 *     sufficient decrease fails: hi = alpha, shrink to the minimiser of the
 *       quadratic through f0, <g0,s> and f_t, kept within
 *       [0.1 alpha, alpha_mul alpha] (or bisect [lo, hi] once a lower
 *       bound exists).  Sufficient decrease is the Armijo test
 *       f_t <= f0 + ftol alpha <g0,s>, or, below the objective's resolution
 *       (fres |f0|), the approximate Armijo test of Hager and Zhang:
 *       f_t <= f0 + fres |f0| and <g_t,s> <= (2 ftol - 1) <g0,s>; fres is
 *       ccm_hip_objective_resolution of evaluate_hip's instance, else 1e-6
 *     strong Wolfe (|<g_t,s>| <= wolfe |<g0,s>|), or sufficient decrease on
 *       the last trial: accept
 *     otherwise <g_t,s> > 0 sets hi, < 0 sets lo; extend (alpha /= alpha_mul)
 *       while unbracketed, else bisect
 *   if the last trial fails the decrease test the step returns to the longest
 *   one that passed it (one more evaluation); with none, x returns to the start, and a
 *   failed steepest-descent search ends the run with
 *   CONJUGRADERR_MAXIMUMLINESEARCH, otherwise the direction is reset to -g
 *   beta = max(0, (<g,g> - <g,g_prev>) / <g_prev,g_prev>)   (Polak-Ribiere+)
 *   s = -g + beta s unless its slope -<g,g> + beta <g,s_old> >= 0, then s = -g
 *   next alpha = alpha * <g_prev, s_prev> / <g, s>
 *   stop with CONJUGRAD_SUCCESS once (f_{it-k} - f_it) / |f_it| < epsilon
 *   (epsilon > 0; epsilon = 0 runs max_iterations: the approximate decrease
 *   test lets f rise within its resolution, which must not end the run).
 * All scalars, dot products and tests are FP64; vectors are FP32.
 */
#define __HIP_PLATFORM_AMD__
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "ccmpred_amd.h"
#include "ccmpred_amd_kernels.h"

#include "state.h"

void ccm_conjugrad_init(ccm_conjugrad_parameter *p)
{
	p->max_iterations = 250;
	p->epsilon = 0.01;
	p->k = 5;
	p->max_linesearch = 5;
	p->alpha_mul = 0.5;
	p->ftol = 1e-4;
	p->wolfe = 0.2;
	p->line_search = CCM_LINE_SEARCH_DEVICE;
}


/* Sufficient decrease (see above).  fres is the relative resolution below
 * which differences of f are not trusted: the BF16 forms quantise the
 * couplings, so their f moves in steps near the optimum while the
 * gradient, and with it the slope <g_t,s>, still resolves the search. */
#define CG_FRES_DEFAULT 1e-6
static int cg_decrease(double ft, double f0, double alpha, double dg0, double dgt, double ftol,
                       double fres)
{
	if (ft <= f0 + ftol * alpha * dg0)
		return 1;
	return ft <= f0 + fres * fabs(f0) && dgt <= (2.0 * ftol - 1.0) * dg0;
}

/* Step after a trial failed the decrease test with no lower bracket: the
 * minimiser of the quadratic through f0, dg0 = <g0,s> and ft (Nocedal and
 * Wright, Numerical Optimization, eq. 3.57) within [0.1, alpha_mul] alpha */
static double cg_shrink(double ft, double f0, double alpha, double dg0, double alpha_mul)
{
	const double den = 2.0 * ((ft - f0) - dg0 * alpha);
	const double aq = (-dg0 * alpha) * alpha / den;
	const double amin = 0.1 * alpha, amax = alpha_mul * alpha;
	return !(aq >= amin) ? amin : (aq > amax ? amax : aq);
}

typedef struct {
	hipStream_t s;
	double *d_part;
	double *d_out;
	double *h_out;
	long n;
	int ncol; /* > 0: skip the a >= 21 padding of this many columns' layout */
} cg_ctx;

static int dots(cg_ctx *c, const float *a, const float *b0, const float *b1, const float *b2,
                int nd, double *out)
{
	const int r = c->ncol ? ccm_launch_dots_padded(a, b0, b1, b2, c->n, nd, c->ncol, c->d_part,
	                                               c->d_out, c->s)
	                      : ccm_launch_dots(a, b0, b1, b2, c->n, nd, c->d_part, c->d_out, c->s);
	if (r != 0)
		return -1;
	if (hipMemcpyAsync(c->h_out, c->d_out, sizeof(double) * (size_t)nd, hipMemcpyDeviceToHost,
	                   c->s) != hipSuccess)
		return -1;
	if (hipStreamSynchronize(c->s) != hipSuccess)
		return -1;
	for (int i = 0; i < nd; i++)
		out[i] = c->h_out[i];
	return 0;
}

static int sync_stream(cg_ctx *c) { return hipStreamSynchronize(c->s) == hipSuccess ? 0 : -1; }

static int axpy(cg_ctx *c, float *x, const float *sdir, float alpha)
{
	return c->ncol ? ccm_launch_axpy_padded(x, sdir, alpha, c->ncol, c->n, c->s)
	               : ccm_launch_axpy(x, sdir, alpha, c->n, c->s);
}

static int negate(cg_ctx *c, float *sdir, const float *g)
{
	return c->ncol ? ccm_launch_negate_padded(sdir, g, c->ncol, c->n, c->s)
	               : ccm_launch_negate(sdir, g, c->n, c->s);
}

static int direction(cg_ctx *c, float *sdir, const float *g, float beta)
{
	return c->ncol ? ccm_launch_direction_padded(sdir, g, beta, c->ncol, c->n, c->s)
	               : ccm_launch_direction(sdir, g, beta, c->n, c->s);
}

/* The padded layout of evaluate_hip's alignment, when x starts with zeros
 * in its a >= 21 padding: the gradient there is 2 lambda_pair x
 * (evaluate_cpu.c:158-161), so x, g and s stay zero there for the whole
 * run and the vector steps and dots may skip it without changing a bit.
 * 0 (visit every entry) for any other callback, size or x. */
static int padded_columns(cg_ctx *c, const float *d_x, conjugrad_evaluate_t evaluate, void *instance)
{
	if (evaluate != evaluate_hip || !instance || ((uintptr_t)d_x & 15u) != 0)
		return 0;
	const userdata *ud = (const userdata *)instance;
	if (ud->ncol <= 0 || (long)ccm_nvar_padded(ud->ncol) != c->n)
		return 0;
	int32_t *flag = (int32_t *)c->d_out;
	if (hipMemsetAsync(flag, 0, sizeof(int32_t), c->s) != hipSuccess ||
	    ccm_launch_pad_scan_x(d_x, ud->ncol, c->n, flag, c->s) != 0 ||
	    hipMemcpyAsync(c->h_out, flag, sizeof(int32_t), hipMemcpyDeviceToHost, c->s) != hipSuccess ||
	    hipStreamSynchronize(c->s) != hipSuccess)
		return -1;
	return *(const int32_t *)c->h_out ? 0 : ud->ncol;
}

int ccm_conjugrad_gpu_ex(int n, conjugrad_float_t *d_x, double *fx, conjugrad_evaluate_t evaluate,
                         conjugrad_progress_t progress, void *instance,
                         const ccm_conjugrad_parameter *p, int *iterations, int *evaluations,
                         double *fhist)
{
	if (p->line_search == CCM_LINE_SEARCH_DEVICE && !progress && evaluate == evaluate_hip && instance &&
	    ((const userdata *)instance)->ncol > 0 &&
	    (long)ccm_nvar_padded(((const userdata *)instance)->ncol) == n) {
		int status = CONJUGRADERR_DEVICE;
		const int r = ccm_conjugrad_instance_device(instance, d_x, p, fx, iterations, evaluations,
		                                            fhist, &status);
		if (r <= 0)
			return r == 0 ? status : CONJUGRADERR_DEVICE;
		/* an instance the device-side CG does not take: the host loop */
	}
	cg_ctx c = {0};
	float *g = NULL, *gprev = NULL, *sdir = NULL;
	double *hist = NULL;
	int ret = CONJUGRADERR_DEVICE, it = 0, nevals = 0;
	double f = NAN;
	c.n = n;
	const double fres =
	    evaluate == evaluate_hip && instance ? ccm_hip_objective_resolution(instance) : CG_FRES_DEFAULT;
	if (hipStreamCreateWithFlags(&c.s, hipStreamNonBlocking) != hipSuccess)
		return CONJUGRADERR_DEVICE;
	if (hipMalloc((void **)&g, sizeof(float) * (size_t)n) != hipSuccess ||
	    hipMalloc((void **)&gprev, sizeof(float) * (size_t)n) != hipSuccess ||
	    hipMalloc((void **)&sdir, sizeof(float) * (size_t)n) != hipSuccess ||
	    hipMalloc((void **)&c.d_part, sizeof(double) * (size_t)ccm_dots_partial_len()) !=
	        hipSuccess ||
	    hipMalloc((void **)&c.d_out, sizeof(double) * 4) != hipSuccess ||
	    hipHostMalloc((void **)&c.h_out, sizeof(double) * 4, hipHostMallocDefault) != hipSuccess)
		goto done;
	hist = (double *)malloc(sizeof(double) * (size_t)(p->max_iterations + 1));
	if (!hist)
		goto done;
	/* zeroed, so that the padding of every vector is zero whichever
	 * launchers run */
	if (hipMemsetAsync(g, 0, sizeof(float) * (size_t)n, c.s) != hipSuccess ||
	    hipMemsetAsync(gprev, 0, sizeof(float) * (size_t)n, c.s) != hipSuccess ||
	    hipMemsetAsync(sdir, 0, sizeof(float) * (size_t)n, c.s) != hipSuccess)
		goto done;
	c.ncol = padded_columns(&c, d_x, evaluate, instance);
	if (c.ncol < 0)
		goto done;
	/* the callback may write g on a stream of its own: the zeroing above
	 * must be complete before it runs, whichever path padded_columns took */
	if (sync_stream(&c))
		goto done;

	f = evaluate(instance, d_x, g, n);
	nevals++;
	if (isnan(f))
		goto done;
	double gg;
	if (dots(&c, g, g, NULL, NULL, 1, &gg))
		goto done;
	if (gg == 0.0) {
		ret = CONJUGRAD_ALREADY_MINIMIZED;
		goto done;
	}
	if (negate(&c, sdir, g))
		goto done;
	double dg0 = -gg;
	double alpha = 1.0 / sqrt(gg);
	int steepest = 1;
	ret = CONJUGRADERR_MAXIMUMITERATION;

	while (it < p->max_iterations) {
		const double f0 = f;
		float *tmp = gprev;
		gprev = g;
		g = tmp;
		/* line search along s; x moves in place, `moved` is its offset */
		double moved = 0.0, ft = 0.0, lo = 0.0, hi = INFINITY;
		int accepted = 0, ls;
		for (ls = 1; ls <= p->max_linesearch; ls++) {
			if (axpy(&c, d_x, sdir, (float)(alpha - moved)) || sync_stream(&c)) {
				ret = CONJUGRADERR_DEVICE;
				goto done;
			}
			moved = alpha;
			ft = evaluate(instance, d_x, g, n);
			nevals++;
			if (isnan(ft)) {
				ret = CONJUGRADERR_DEVICE;
				goto done;
			}
			double dgt;
			if (dots(&c, g, sdir, NULL, NULL, 1, &dgt)) {
				ret = CONJUGRADERR_DEVICE;
				goto done;
			}
			if (!cg_decrease(ft, f0, alpha, dg0, dgt, p->ftol, fres)) {
				/* not enough decrease: shrink, bisecting once bracketed */
				hi = alpha;
				alpha = lo > 0.0 ? 0.5 * (lo + hi) : cg_shrink(ft, f0, alpha, dg0, p->alpha_mul);
				continue;
			}
			if (fabs(dgt) <= -p->wolfe * dg0 || ls == p->max_linesearch) {
				accepted = 1;
				break;
			}
			if (dgt > 0.0)
				hi = alpha; /* past the minimum along s */
			else
				lo = alpha; /* still descending steeply: extend */
			alpha = isinf(hi) ? alpha / p->alpha_mul : 0.5 * (lo + hi);
		}
		if (!accepted && lo > 0.0) {
			/* last trial failed the decrease test: return to the longest
			 * step that passed it */
			if (axpy(&c, d_x, sdir, (float)(lo - moved)) || sync_stream(&c)) {
				ret = CONJUGRADERR_DEVICE;
				goto done;
			}
			moved = alpha = lo;
			ft = evaluate(instance, d_x, g, n);
			nevals++;
			if (isnan(ft)) {
				ret = CONJUGRADERR_DEVICE;
				goto done;
			}
			accepted = 1;
		}
		if (!accepted) {
			if (axpy(&c, d_x, sdir, (float)(-moved)) || sync_stream(&c)) {
				ret = CONJUGRADERR_DEVICE;
				goto done;
			}
			tmp = gprev;
			gprev = g;
			g = tmp;
			f = f0;
			if (steepest) {
				ret = CONJUGRADERR_MAXIMUMLINESEARCH;
				goto done;
			}
			if (negate(&c, sdir, g)) {
				ret = CONJUGRADERR_DEVICE;
				goto done;
			}
			dg0 = -gg;
			alpha = 1.0 / sqrt(gg);
			steepest = 1;
			continue;
		}
		f = ft;
		hist[it] = f;
		if (fhist)
			fhist[it] = f;
		it++;

		double d3[3];
		if (dots(&c, g, sdir, g, gprev, 3, d3)) {
			ret = CONJUGRADERR_DEVICE;
			goto done;
		}
		const double dgs = d3[0], ggn = d3[1], ggp = d3[2];
		if (progress) {
			double xx;
			if (dots(&c, d_x, d_x, NULL, NULL, 1, &xx)) {
				ret = CONJUGRADERR_DEVICE;
				goto done;
			}
			progress(instance, d_x, g, (conjugrad_float_t)f, (conjugrad_float_t)sqrt(xx),
			         (conjugrad_float_t)sqrt(ggn), (conjugrad_float_t)alpha, n, it, ls);
		}
		if (p->epsilon > 0.0 && p->k > 0 && it > p->k) {
			const double before = hist[it - 1 - p->k];
			if ((before - f) / fabs(f) < p->epsilon) {
				ret = CONJUGRAD_SUCCESS;
				break;
			}
		}
		double beta = (ggn - ggp) / gg;
		if (beta < 0)
			beta = 0;
		/* slope of the new direction -g + beta s, from the dots above */
		double dgn = -ggn + beta * dgs;
		if (dgn >= 0) {
			if (negate(&c, sdir, g)) {
				ret = CONJUGRADERR_DEVICE;
				goto done;
			}
			dgn = -ggn;
			steepest = 1;
		} else {
			if (direction(&c, sdir, g, (float)beta)) {
				ret = CONJUGRADERR_DEVICE;
				goto done;
			}
			steepest = beta == 0.0;
		}
		alpha = alpha * dg0 / dgn;
		dg0 = dgn;
		gg = ggn;
	}
done:
	if (c.s)
		(void)hipStreamSynchronize(c.s);
	*fx = f;
	if (iterations)
		*iterations = it;
	if (evaluations)
		*evaluations = nevals;
	if (g)
		(void)hipFree(g);
	if (gprev)
		(void)hipFree(gprev);
	if (sdir)
		(void)hipFree(sdir);
	if (c.d_part)
		(void)hipFree(c.d_part);
	if (c.d_out)
		(void)hipFree(c.d_out);
	if (c.h_out)
		(void)hipHostFree(c.h_out);
	if (c.s)
		(void)hipStreamDestroy(c.s);
	free(hist);
	return ret;
}

int ccm_conjugrad_gpu(int n, conjugrad_float_t *d_x, conjugrad_float_t *fx,
                      conjugrad_evaluate_t evaluate, conjugrad_progress_t progress, void *instance,
                      const ccm_conjugrad_parameter *param)
{
	double f = 0;
	int r = ccm_conjugrad_gpu_ex(n, d_x, &f, evaluate, progress, instance, param, NULL, NULL,
	                             NULL);
	*fx = (conjugrad_float_t)f;
	return r;
}
