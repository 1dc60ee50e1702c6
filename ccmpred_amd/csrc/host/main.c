/*
 * ccmpred (MI355X): command-line driver for the PLL path.
 *
 * Flags follow the reference's CLI (ccmpred.c:183-339): same options, same
 * defaults (-n 250, -e 0.01, -k 5, -w 0.8, -l 0.2), same progress table and
 * output formats, same exit codes (1 usage/device, 2 MSA open, 3/4 output
 * open, 100 FASTA/a3m input), -b for the MessagePack raw file.  Addition:
 * -P selects the precision form (fp32 | bf16 | bf16f | bf16x2).  The
 * reference's CPU evaluators (-t) are not
 * offered: this driver evaluates on the GPU only.
 */
#define __HIP_PLATFORM_AMD__
#include <hip/hip_runtime_api.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "ccmpred_amd.h"

static int progress(void *instance, const conjugrad_float_t *x, const conjugrad_float_t *g,
                    const conjugrad_float_t fx, const conjugrad_float_t xnorm,
                    const conjugrad_float_t gnorm, const conjugrad_float_t step, int n, int k,
                    int ls)
{
	(void)instance;
	(void)x;
	(void)g;
	(void)n;
	printf("%-4d\t%-4d\t%-8g\t%-8g\t%-8.8g\t%-3.3g\n", k, ls, fx, xnorm, gnorm, step);
	fflush(stdout);
	return 0;
}

static void usage(const char *exe, int full)
{
	printf("Usage: %s [options] input.aln output.mat\n\n", exe);
	if (full) {
		printf("Options:\n");
		printf("\t-d DEVICE \tCalculate on HIP device number DEVICE [default: 0]\n");
		printf("\t-n NUMITER\tCompute a maximum of NUMITER operations [default: 250]\n");
		printf("\t-e EPSILON\tSet convergence criterion for minimum decrease in the last K "
		       "iterations to EPSILON [default: 0.01]\n");
		printf("\t-k LASTK  \tSet K parameter for convergence criterion to LASTK [default: 5]\n");
		printf("\n\t-i INIFILE\tRead initial weights from INIFILE\n");
		printf("\t-r RAWFILE\tStore raw prediction matrix in RAWFILE\n");
		printf("\t-b BRAWFLE\tStore raw prediction matrix in msgpack format in BRAWFLE\n");
		printf("\n\t-w IDTHRES\tSet sequence reweighting identity threshold to IDTHRES "
		       "[default: 0.8]\n");
		printf("\t-l LFACTOR\tSet pairwise regularization coefficients to LFACTOR * (L-1) "
		       "[default: 0.2]\n");
		printf("\t-A        \tDisable average product correction (APC)\n");
		printf("\t-R        \tRe-normalize output matrix to [0,1]\n");
		printf("\t-P PREC   \tPrecision form: fp32, bf16, bf16f (two-pass forward) or bf16x2 "
		       "[default: fp32]\n");
		printf("\t-h        \tDisplay help\n\n");
	}
	exit(1);
}

int main(int argc, char **argv)
{
	const char *rawfile = NULL, *msgpackfile = NULL, *initfile = NULL, *msafile = NULL, *matfile = NULL;
	int numiter = 250, use_apc = 1, use_norm = 0, device = 0, cg_k = 5;
	float lambda_single = 0.01f, lambda_pair_factor = 0.2f, theta = 0.8f;
	double cg_eps = 0.01;
	ccm_hip_config cfg;
	ccm_hip_default_config(&cfg);

	for (int i = 1; i < argc; i++) {
		const char *a = argv[i];
		if (a[0] == '-' && a[1] && !a[2]) {
			const char o = a[1];
			const char *val = NULL;
			if (strchr("dtnekirbwlP", o)) {
				if (i + 1 >= argc)
					usage(argv[0], 0);
				val = argv[++i];
			}
			switch (o) {
			case 'd': device = atoi(val); break;
			case 't':
				fprintf(stderr, "CPU evaluation (-t) is not available: this ccmpred evaluates on the GPU only\n");
				return 1;
			case 'n': numiter = atoi(val); break;
			case 'e': cg_eps = atof(val); break;
			case 'k': cg_k = atoi(val); break;
			case 'i': initfile = val; break;
			case 'r': rawfile = val; break;
			case 'b': msgpackfile = val; break;
			case 'w': theta = (float)atof(val); break;
			case 'l': lambda_pair_factor = (float)atof(val); break;
			case 'A': use_apc = 0; break;
			case 'R': use_norm = 1; break;
			case 'P':
				if (!strcmp(val, "fp32")) {
					cfg.precision = CCM_PRECISION_FP32;
				} else if (!strcmp(val, "bf16")) {
					cfg.precision = CCM_PRECISION_BF16;
					cfg.bf16_passes = 1;
				} else if (!strcmp(val, "bf16x2")) {
					cfg.precision = CCM_PRECISION_BF16;
					cfg.bf16_passes = 2;
					cfg.bf16_fwd_passes = 2;
				} else if (!strcmp(val, "bf16f")) {
					cfg.precision = CCM_PRECISION_BF16;
					cfg.bf16_passes = 1;
					cfg.bf16_fwd_passes = 2;
				} else {
					usage(argv[0], 0);
				}
				break;
			case 'h':
			case '?': usage(argv[0], 1); break;
			default: fprintf(stderr, "\nERROR: Unknown argument\n\n"); return 1;
			}
		} else if (!msafile) {
			msafile = a;
		} else if (!matfile) {
			matfile = a;
		} else {
			usage(argv[0], 0);
		}
	}
	if (!msafile || !matfile)
		usage(argv[0], 0);

	FILE *mf = fopen(msafile, "r");
	if (!mf) {
		printf("Cannot open %s!\n\n", msafile);
		return 2;
	}
	int ncol = 0, nrow = 0;
	unsigned char *msa = ccm_read_msa(mf, &ncol, &nrow);
	fclose(mf);
	if (!msa) {
		printf("ERROR: The alignment file seems to be in a3m or fasta format!\n"
		       "Please reformat the alignment to PSICOV format!\n\n");
		return 100;
	}

	int ndev = 0;
	if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
		printf("No HIP devices available\n");
		return 1;
	}
	if (device < 0 || device >= ndev) {
		printf("Error: %d is not a valid device number. Please choose a number between 0 and %d\n",
		       device, ndev - 1);
		return 1;
	}
	cfg.device = device;
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, device) == hipSuccess)
		printf("Found %d HIP devices, using device #%d: %s\n", ndev, device, prop.name);

	const long nvp = ccm_nvar_padded(ncol);
	const int nsingle = ncol * (CCM_N_ALPHA - 1);
	float *x = (float *)malloc(sizeof(float) * (size_t)nvp);
	if (!x) {
		fprintf(stderr, "\nERROR: Not enough memory to allocate variables!\n\n");
		return 1;
	}
	userdata ud;
	memset(&ud, 0, sizeof ud);
	ud.msa = msa;
	ud.ncol = ncol;
	ud.nrow = nrow;
	ud.nsingle = nsingle;
	ud.nvar = nsingle + ncol * ncol * CCM_N_ALPHA * CCM_N_ALPHA;
	ud.lambda_single = lambda_single;
	ud.lambda_pair = lambda_pair_factor * (float)(ncol - 1);
	ud.reweighting_threshold = theta;
	ud.weights = (float *)malloc(sizeof(float) * (size_t)nrow);

	if (initfile) {
		memset(x, 0, sizeof(float) * (size_t)nvp);
		if (ccm_read_raw(initfile, ncol, x) != 0) {
			fprintf(stderr, "\nERROR: malformed raw file %s\n\n", initfile);
			return 1;
		}
	} else {
		ccm_init_bias(x, msa, ncol, nrow);
	}

	ccm_conjugrad_parameter param;
	ccm_conjugrad_init(&param);
	param.max_iterations = numiter;
	param.epsilon = cg_eps;
	param.k = cg_k;

	if (init_hip_ex(&ud, &cfg) != 0) {
		fprintf(stderr, "\nERROR: device initialisation failed\n\n");
		return 1;
	}
	printf("\nWill optimize %d %ld-bit variables\n\n", ud.nvar, (long)sizeof(conjugrad_float_t) * 8);
	printf("iter\teval\tf(x)    \t║x║     \t║g║     \tstep\n");

	float *d_x = NULL;
	if (hipMalloc((void **)&d_x, sizeof(float) * (size_t)nvp) != hipSuccess ||
	    hipMemcpy(d_x, x, sizeof(float) * (size_t)nvp, hipMemcpyHostToDevice) != hipSuccess) {
		fprintf(stderr, "\nERROR: device memory\n\n");
		return 1;
	}
	conjugrad_float_t fx = 0;
	int ret = ccm_conjugrad_gpu((int)nvp, d_x, &fx, evaluate_hip, progress, &ud, &param);
	if (hipMemcpy(x, d_x, sizeof(float) * (size_t)nvp, hipMemcpyDeviceToHost) != hipSuccess) {
		fprintf(stderr, "\nERROR: device memory\n\n");
		return 1;
	}
	(void)hipFree(d_x);

	printf("\n%s with status code %d - ", ret < 0 ? "Exit" : "Done", ret);
	if (ret == CONJUGRAD_SUCCESS)
		printf("Success!\n");
	else if (ret == CONJUGRAD_ALREADY_MINIMIZED)
		printf("Already minimized!\n");
	else if (ret == CONJUGRADERR_MAXIMUMITERATION)
		printf("Maximum number of iterations reached.\n");
	else
		printf("Unknown status code!\n");
	printf("\nFinal fx = %f\n\n", fx);

	FILE *out = fopen(matfile, "w");
	if (!out) {
		printf("Cannot open %s for writing!\n\n", matfile);
		return 3;
	}
	if (rawfile) {
		printf("Writing raw output to %s\n", rawfile);
		FILE *rf = fopen(rawfile, "w");
		if (!rf) {
			printf("Cannot open %s for writing!\n\n", rawfile);
			return 4;
		}
		ccm_write_raw(rf, x, ncol);
		fclose(rf);
	}
	if (msgpackfile) {
		printf("Writing msgpack raw output to %s\n", msgpackfile);
		FILE *bf = fopen(msgpackfile, "w");
		if (!bf) {
			printf("Cannot open %s for writing!\n\n", msgpackfile);
			return 4;
		}
		const int wrc = ccm_write_raw_msgpack(bf, x, ncol);
		fclose(bf);
		if (wrc != 0) {
			fprintf(stderr, "\nERROR: writing %s failed\n\n", msgpackfile);
			return 4;
		}
	}
	float *mat = (float *)malloc(sizeof(float) * (size_t)ncol * ncol);
	ccm_sum_submatrices(x, mat, ncol);
	if (use_apc)
		ccm_apc(mat, ncol);
	if (use_norm)
		ccm_normalize(mat, ncol);
	ccm_write_matrix(out, mat, ncol, ncol);
	fclose(out);

	destroy_hip(&ud);
	free(mat);
	free(x);
	free(ud.weights);
	free(msa);
	printf("Output can be found in %s\n", matfile);
	return 0;
}
