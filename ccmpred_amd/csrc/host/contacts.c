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
	if (L < 2)
		return; /* no off-diagonal entry (the reference reads past the 1x1 map) */
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

/* ---- raw couplings as MessagePack ("ccm-1" map) ----
 * Each value takes the shortest MessagePack form for its size, as the
 * msgpack-c packer that io.c:105-196 calls does: positive fixint / uint8 /
 * uint16 / uint32 for non-negative ints, fixstr / str8, fixarray / array16 /
 * array32, fixmap / map16 / map32, float32 (0xca) for every coupling. */
static void mp_be(FILE *o, uint32_t v, int nbytes)
{
	for (int s = (nbytes - 1) * 8; s >= 0; s -= 8)
		fputc((int)((v >> s) & 0xFF), o);
}

static void mp_head(FILE *o, uint32_t n, int fix_base, uint32_t fix_limit, int tag16, int tag32)
{
	if (n < fix_limit) {
		fputc(fix_base | (int)n, o);
	} else if (n < 65536u) {
		fputc(tag16, o);
		mp_be(o, n, 2);
	} else {
		fputc(tag32, o);
		mp_be(o, n, 4);
	}
}

static void mp_map(FILE *o, uint32_t n) { mp_head(o, n, 0x80, 16, 0xde, 0xdf); }
static void mp_array(FILE *o, uint32_t n) { mp_head(o, n, 0x90, 16, 0xdc, 0xdd); }

static void mp_str(FILE *o, const char *s)
{
	const uint32_t n = (uint32_t)strlen(s);
	if (n < 32) {
		fputc(0xa0 | (int)n, o);
	} else if (n < 256) {
		fputc(0xd9, o);
		fputc((int)n, o);
	} else {
		mp_head(o, n, 0, 0, 0xda, 0xdb);
	}
	fwrite(s, 1, n, o);
}

/* non-negative ints only (column indices and ncol) */
static void mp_uint(FILE *o, uint32_t v)
{
	if (v < 128u) {
		fputc((int)v, o);
	} else if (v < 256u) {
		fputc(0xcc, o);
		fputc((int)v, o);
	} else if (v < 65536u) {
		fputc(0xcd, o);
		mp_be(o, v, 2);
	} else {
		fputc(0xce, o);
		mp_be(o, v, 4);
	}
}

static void mp_float(FILE *o, float f)
{
	uint32_t u;
	memcpy(&u, &f, 4);
	fputc(0xca, o);
	mp_be(o, u, 4);
}

int ccm_write_raw_msgpack(FILE *out, const float *x, int ncol)
{
	if (!out || !x || ncol < 1)
		return -1;
	const long L = ncol;
	const float *x2 = x + ccm_nsingle_padded(ncol);
	char key[32];
	mp_map(out, 4);
	mp_str(out, "format");
	mp_str(out, "ccm-1");
	mp_str(out, "ncol");
	mp_uint(out, (uint32_t)ncol);
	mp_str(out, "x_single");
	mp_array(out, (uint32_t)(L * (NA - 1)));
	for (long i = 0; i < L; i++)
		for (int a = 0; a < NA - 1; a++)
			mp_float(out, x[(long)a * L + i]);
	mp_str(out, "x_pair");
	mp_map(out, (uint32_t)(L * (L - 1) / 2));
	for (long i = 0; i < L; i++)
		for (long j = i + 1; j < L; j++) {
			snprintf(key, sizeof key, "%ld/%ld", i, j);
			mp_str(out, key);
			mp_map(out, 3);
			mp_str(out, "i");
			mp_uint(out, (uint32_t)i);
			mp_str(out, "j");
			mp_uint(out, (uint32_t)j);
			mp_str(out, "x");
			mp_array(out, NA * NA);
			for (int a = 0; a < NA; a++)
				for (int b = 0; b < NA; b++)
					mp_float(out, x2[X2I(b, j, a, i, L)]);
		}
	return ferror(out) ? -1 : 0;
}
