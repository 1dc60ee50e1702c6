/*
 * Host-side model set-up and contact-map output of the PLL path:
 * PSICOV reading (sequence.c), init_bias (ccmpred.c), contact scores and
 * APC (util.c), matrix and raw-coupling text I/O (io.c).
 */
#include <ctype.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccmpred_amd.h"

#define NA CCM_N_ALPHA
#define NAP CCM_N_ALPHA_PAD
#define X2I(b, k, a, j, L) ((((long)(b) * (L) + (k)) * NAP + (a)) * (L) + (j))

/* residue letter -> index, order ARNDCQEGHILKMFPSTWYV; others -> 20
 * (sequence.c:10-27) */
static unsigned char residue_index(unsigned char c)
{
	static const char order[] = "ARNDCQEGHILKMFPSTWYV";
	if (!isalpha(c))
		return 20;
	c = (unsigned char)toupper(c);
	const char *p = strchr(order, c);
	return p ? (unsigned char)(p - order) : 20;
}

/* sequence.c:48-80.  Lines are right-trimmed; shorter lines are gap-filled
 * (the reference reads stale buffer bytes there). */
unsigned char *ccm_read_msa(FILE *f, int *ncol, int *nrow)
{
	size_t cap = 0, len = 0;
	char *buf = NULL;
	int ch;
	/* slurp */
	while ((ch = fgetc(f)) != EOF) {
		if (len + 1 >= cap) {
			cap = cap ? cap * 2 : 65536;
			char *nb = (char *)realloc(buf, cap);
			if (!nb) {
				free(buf);
				return NULL;
			}
			buf = nb;
		}
		buf[len++] = (char)ch;
	}
	if (!buf)
		buf = (char *)calloc(1, 1);
	buf[len] = 0;
	if (strchr(buf, '>')) {
		free(buf);
		return NULL;
	}
	/* count lines and the longest trimmed line */
	int rows = 0, cols = 0;
	char *p = buf;
	while (*p) {
		char *e = strchr(p, '\n');
		size_t l = e ? (size_t)(e - p) : strlen(p);
		size_t t = l;
		while (t > 0 && isspace((unsigned char)p[t - 1]))
			t--;
		if (!(e == NULL && l == 0)) {
			rows++;
			if ((int)t > cols)
				cols = (int)t;
		}
		if (!e)
			break;
		p = e + 1;
	}
	unsigned char *out = (unsigned char *)malloc((size_t)rows * (size_t)(cols > 0 ? cols : 1));
	p = buf;
	for (int i = 0; i < rows; i++) {
		char *e = strchr(p, '\n');
		size_t l = e ? (size_t)(e - p) : strlen(p);
		size_t t = l;
		while (t > 0 && isspace((unsigned char)p[t - 1]))
			t--;
		for (int j = 0; j < cols; j++)
			out[(long)i * cols + j] = (size_t)j < t ? residue_index((unsigned char)p[j]) : 20;
		p = e ? e + 1 : p + l;
	}
	free(buf);
	*ncol = cols;
	*nrow = rows;
	return out;
}

void ccm_init_bias(float *x, const unsigned char *msa, int ncol, int nrow)
{
	memset(x, 0, sizeof(float) * (size_t)ccm_nvar_padded(ncol));
	for (int j = 0; j < ncol; j++) {
		int cnt[NA];
		for (int a = 0; a < NA; a++)
			cnt[a] = 1;
		for (long i = 0; i < nrow; i++)
			cnt[msa[i * ncol + j]]++;
		const int total = nrow + NA;
		float frac[NA];
		for (int a = 0; a < NA; a++)
			frac[a] = (float)cnt[a] / (float)total;
		const float gap = logf(frac[NA - 1]);
		for (int a = 0; a < NA - 1; a++)
			x[(long)a * ncol + j] = logf(frac[a]) - gap;
	}
}

void ccm_sum_submatrices(const float *x, float *out, int ncol)
{
	const long L = ncol;
	const float *x2 = x + ccm_nsingle_padded(ncol);
	for (long k = 0; k < L; k++) {
		for (long j = 0; j < L; j++) {
			float mean = 0;
			for (int a = 0; a < NA; a++)
				for (int b = 0; b < NA; b++)
					mean += x2[X2I(b, k, a, j, L)];
			mean /= (float)(NA * NA);
			float acc = 0;
			for (int a = 0; a < NA - 1; a++)
				for (int b = 0; b < NA - 1; b++) {
					const float d = x2[X2I(b, k, a, j, L)] - mean;
					acc += d * d;
				}
			out[k * L + j] = k == j ? 0.0f : (float)sqrt((double)acc);
		}
	}
}

void ccm_apc(float *mat, int ncol)
{
	const long L = ncol;
	float *colmean = (float *)calloc((size_t)L, sizeof(float));
	float total = 0;
	for (long i = 0; i < L; i++)
		for (long j = 0; j < L; j++) {
			const float v = mat[i * L + j];
			colmean[j] += v / (float)L;
			total += v;
		}
	total /= (float)(L * L);
	for (long i = 0; i < L; i++)
		for (long j = 0; j < L; j++)
			mat[i * L + j] -= (colmean[i] * colmean[j]) / total;
	float lo = INFINITY;
	for (long i = 0; i < L; i++)
		for (long j = i + 1; j < L; j++)
			if (mat[i * L + j] < lo)
				lo = mat[i * L + j];
	for (long i = 0; i < L; i++)
		for (long j = 0; j < L; j++)
			mat[i * L + j] = i == j ? 0.0f : mat[i * L + j] - lo;
	free(colmean);
}

void ccm_normalize(float *mat, int ncol)
{
	const long L = ncol;
	float lo = mat[1], hi = mat[1];
	for (long i = 0; i < L; i++)
		for (long j = 0; j < L; j++) {
			if (i == j)
				continue;
			const float v = mat[i * L + j];
			if (v < lo)
				lo = v;
			if (v > hi)
				hi = v;
		}
	const float range = hi - lo;
	for (long i = 0; i < L; i++)
		for (long j = 0; j < L; j++)
			mat[i * L + j] = i == j ? 0.0f : (mat[i * L + j] - lo) / range;
}

void ccm_write_matrix(FILE *out, const float *mat, int ncol, int nrow)
{
	for (long i = 0; i < nrow; i++) {
		for (long j = 0; j < ncol; j++)
			fprintf(out, "%.20e\t", mat[i * ncol + j]);
		fprintf(out, "\n");
	}
}

void ccm_write_raw(FILE *out, const float *x, int ncol)
{
	const long L = ncol;
	const float *x2 = x + ccm_nsingle_padded(ncol);
	for (long i = 0; i < L; i++) {
		for (int a = 0; a < NA - 1; a++)
			fprintf(out, "%g\t", x[(long)a * L + i]);
		fprintf(out, "\n");
	}
	for (long i = 0; i < L; i++)
		for (long j = i + 1; j < L; j++) {
			fprintf(out, "# %ld %ld\n", i, j);
			for (int a = 0; a < NA; a++) {
				for (int b = 0; b < NA; b++)
					fprintf(out, "%.20e\t", x2[X2I(b, j, a, i, L)]);
				fprintf(out, "\n");
			}
		}
}

int ccm_read_raw(const char *filename, int ncol, float *x)
{
	FILE *f = fopen(filename, "r");
	if (!f)
		return -1;
	const long L = ncol;
	float *x2 = x + ccm_nsingle_padded(ncol);
	char line[16384], expect[64];
	int rc = 0;
	for (long i = 0; i < L && rc == 0; i++) {
		if (!fgets(line, sizeof line, f)) {
			rc = -1;
			break;
		}
		char *tok = line;
		for (int a = 0; a < NA - 1; a++)
			x[(long)a * L + i] = (float)strtod(tok, &tok);
	}
	for (long i = 0; i < L && rc == 0; i++)
		for (long j = i + 1; j < L && rc == 0; j++) {
			if (!fgets(line, sizeof line, f)) {
				rc = -1;
				break;
			}
			snprintf(expect, sizeof expect, "# %ld %ld\n", i, j);
			if (strcmp(line, expect) != 0) {
				rc = -1;
				break;
			}
			for (int a = 0; a < NA; a++) {
				if (!fgets(line, sizeof line, f)) {
					rc = -1;
					break;
				}
				char *tok = line;
				for (int b = 0; b < NA; b++) {
					const float v = (float)strtod(tok, &tok);
					x2[X2I(b, j, a, i, L)] = v;
					x2[X2I(a, i, b, j, L)] = v;
				}
			}
		}
	fclose(f);
	return rc;
}
