/*
 * TEST INFRASTRUCTURE ONLY -- part of the CPU oracle.  Nothing in the
 * product (ccmpred_amd/) may include, link or call this file.
 *
 * Type-generic restatement of CCMpred's serial pseudo-likelihood evaluator.
 * Included twice by ccm_oracle.c, once with REAL=float (the reference's
 * default CONJUGRAD_FLOAT=32 build) and once with REAL=double (the FP64
 * oracle every error figure is quoted against).
 *
 * Expects REAL, SFX (function-name suffix), EXPR(x), LOGR(x) and the
 * ORA_* layout helpers from ccm_oracle.h to be defined by the includer.
 */

#define ORA_CAT2(a, b) a##b
#define ORA_CAT(a, b) ORA_CAT2(a, b)
#define ORA_FN(name) ORA_CAT(name, SFX)

/*
 * f and tied gradient of the PLL objective.
 * Follows src/evaluate_cpu.c:16-166 step for step, including its FP32
 * operation order: per sequence, the logit table PC is seeded with the
 * singles (gap logit 0), the coupling rows W(x_nk, k, ., .) are added in k
 * order (:55-75), the log-partition sums exp over a=0..20 (:78-87), f and the
 * node gradient are accumulated per column (:96-110), then the pair
 * gradient rows (x_nk, k) receive -w at the observed (x_nj, j) and +w*P
 * (:112-126).  After all sequences the raw pair gradient is symmetrised
 * (:131-140), self-edge blocks are zeroed (:143-149) and the L2 terms are
 * added over the real singles and the whole padded pair block (:151-163).
 *
 * x, g: nvar_padded entries in the reference layout (ccmpred.h:11-17).
 * Returns f.
 */
REAL ORA_FN(ora_evaluate_)(const unsigned char *msa, const REAL *weights,
                           int ncol, int nrow, REAL lambda_single,
                           REAL lambda_pair, const REAL *x, REAL *g)
{
	const long L = ncol;
	const long nsingle = L * (ORA_NA - 1);
	const long nsp = ora_nsingle_padded(ncol);
	const long nvp = ora_nvar_padded(ncol);
	const long npair = nvp - nsp;
	const REAL *x1 = x;
	const REAL *x2 = x + nsp;
	REAL *g1 = g;
	REAL *g2 = g + nsp;

	REAL *raw = (REAL *)calloc((size_t)npair, sizeof(REAL));
	REAL *logit = (REAL *)malloc(sizeof(REAL) * ORA_NA * L);
	REAL *prob = (REAL *)malloc(sizeof(REAL) * ORA_NA * L);
	REAL *lz = (REAL *)malloc(sizeof(REAL) * L);

	REAL f = 0;
	memset(g, 0, sizeof(REAL) * (size_t)nvp);

	for (long n = 0; n < nrow; n++) {
		const REAL wn = weights[n];
		const unsigned char *seq = msa + n * L;

		/* seed with singles; the gap logit has no single term */
		for (long a = 0; a < ORA_NA - 1; a++)
			for (long s = 0; s < L; s++)
				logit[a * L + s] = x1[a * L + s];
		for (long s = 0; s < L; s++)
			logit[(ORA_NA - 1) * L + s] = 0;

		/* add the coupling row selected by each observed residue */
		for (long k = 0; k < L; k++) {
			const REAL *row = x2 + ORA_X2(seq[k], k, 0, 0, L);
			for (long c = 0; c < ORA_NA * L; c++)
				logit[c] += row[c];
		}

		/* log-partition per column, summed a = 0..20 */
		for (long s = 0; s < L; s++)
			lz[s] = 0;
		for (long a = 0; a < ORA_NA; a++)
			for (long s = 0; s < L; s++)
				lz[s] += EXPR(logit[a * L + s]);
		for (long s = 0; s < L; s++)
			lz[s] = LOGR(lz[s]);

		for (long a = 0; a < ORA_NA; a++)
			for (long s = 0; s < L; s++)
				prob[a * L + s] = EXPR(logit[a * L + s] - lz[s]);

		/* objective and node gradient */
		for (long k = 0; k < L; k++) {
			const int xk = seq[k];
			f += wn * (-logit[xk * L + k] + lz[k]);
			if (xk < ORA_NA - 1)
				g1[xk * L + k] -= wn;
			for (long a = 0; a < ORA_NA - 1; a++)
				g1[a * L + k] += wn * prob[a * L + k];
		}

		/* raw pair gradient rows (x_nk, k) */
		for (long k = 0; k < L; k++) {
			REAL *rg = raw + ORA_X2(seq[k], k, 0, 0, L);
			for (long j = 0; j < L; j++)
				rg[seq[j] * L + j] -= wn;
			for (long c = 0; c < ORA_NA * L; c++)
				rg[c] += wn * prob[c];
		}
	}

	/* g2 = raw + raw^T over the 21x21 blocks */
	for (long b = 0; b < ORA_NA; b++)
		for (long k = 0; k < L; k++)
			for (long a = 0; a < ORA_NA; a++)
				for (long j = 0; j < L; j++)
					g2[ORA_X2(b, k, a, j, L)] = raw[ORA_X2(b, k, a, j, L)] +
					                            raw[ORA_X2(a, j, b, k, L)];

	/* self-edge blocks carry no likelihood gradient */
	for (long b = 0; b < ORA_NA; b++)
		for (long k = 0; k < L; k++)
			for (long a = 0; a < ORA_NA; a++)
				g2[ORA_X2(b, k, a, k, L)] = 0;

	/* L2: real singles, then the whole padded pair block */
	REAL reg = 0;
	const REAL two = 2, half = (REAL)0.5;
	for (long v = 0; v < nsingle; v++) {
		reg += lambda_single * x[v] * x[v];
		g[v] += two * lambda_single * x[v];
	}
	for (long v = nsp; v < nvp; v++) {
		reg += half * lambda_pair * x[v] * x[v];
		g[v] += two * lambda_pair * x[v];
	}
	f += reg;

	free(raw);
	free(logit);
	free(prob);
	free(lz);
	return f;
}

/*
 * Per-(sequence, column) log-partition and per-sequence pseudo-log-likelihood
 * terms, used by tests that look at the forward pass alone.
 * lse_out: nrow x ncol, row-major.  Same arithmetic as the evaluator above.
 */
void ORA_FN(ora_forward_lse_)(const unsigned char *msa, int ncol, int nrow,
                              const REAL *x, REAL *lse_out)
{
	const long L = ncol;
	const long nsp = ora_nsingle_padded(ncol);
	const REAL *x1 = x;
	const REAL *x2 = x + nsp;
	REAL *logit = (REAL *)malloc(sizeof(REAL) * ORA_NA * L);

	for (long n = 0; n < nrow; n++) {
		const unsigned char *seq = msa + n * L;
		for (long a = 0; a < ORA_NA - 1; a++)
			for (long s = 0; s < L; s++)
				logit[a * L + s] = x1[a * L + s];
		for (long s = 0; s < L; s++)
			logit[(ORA_NA - 1) * L + s] = 0;
		for (long k = 0; k < L; k++) {
			const REAL *row = x2 + ORA_X2(seq[k], k, 0, 0, L);
			for (long c = 0; c < ORA_NA * L; c++)
				logit[c] += row[c];
		}
		for (long s = 0; s < L; s++) {
			REAL acc = 0;
			for (long a = 0; a < ORA_NA; a++)
				acc += EXPR(logit[a * L + s]);
			lse_out[n * L + s] = LOGR(acc);
		}
	}
	free(logit);
}

#undef ORA_FN
#undef ORA_CAT
#undef ORA_CAT2
