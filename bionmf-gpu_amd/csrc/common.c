/*
 * common.c -- drop-in for the reference's src/common.c: process globals, messages, matrix
 * limits, command-line options, connectivity difference and seed (single process).
 */
#include "common.h"

#include <errno.h>
#include <inttypes.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <unistd.h>

/* Default alignment, in items (include/common_defaults.h:100-103: 64 bytes of real). */
#define DEFAULT_MEMORY_ALIGNMENT ((index_t)(64 / sizeof(real)))

index_t process_id = 0, num_processes = 1, num_act_processes = 1;
index_t memory_alignment = DEFAULT_MEMORY_ALIGNMENT;
size_t matrix_max_num_items = SIZE_MAX / (2 * sizeof(real));
index_t matrix_max_pitch = IDX_MAX, matrix_max_non_padded_dim = IDX_MAX;
index_t N = 0, M = 0, K = 0;
index_t NpP = 0, MpP = 0, bN = 0, bM = 0;
index_t Mp = 0, Kp = 0, MpPp = 0;
index_t *classification = NULL, *last_classification = NULL;
real *W = NULL, *H = NULL, *Vcol = NULL, *Vrow = NULL;
char const *const file_extension[3] = {".txt", ".dat", ".native.dat"};

/* ---------------------------------------------------------------- messages */

static int vprint(FILE *f, bool newline_first, int errnum, char const *fmt, va_list ap)
{
	if (f == stderr)
		fflush(stdout);
	if (newline_first && num_processes > 1)
		fprintf(f, "\n[P%" PRI_IDX "] ", process_id);
	int r = vfprintf(f, fmt, ap);
	if (errnum)
		fprintf(f, ": %s\n", strerror(errnum));
	return r < 0 ? EXIT_FAILURE : EXIT_SUCCESS;
}

#define DEFINE_PRINTER(name, stream, first, use_errnum)                                      \
	int name(bool all_processes, PRINTER_ERRNUM_##use_errnum char const *fmt, ...)       \
	{                                                                                    \
		(void)all_processes;                                                         \
		va_list ap;                                                                  \
		va_start(ap, fmt);                                                           \
		int r = vprint(stream, first, PRINTER_ERRVAL_##use_errnum, fmt, ap);         \
		va_end(ap);                                                                  \
		return r;                                                                    \
	}
#define PRINTER_ERRNUM_0
#define PRINTER_ERRNUM_1 int errnum,
#define PRINTER_ERRVAL_0 0
#define PRINTER_ERRVAL_1 errnum

DEFINE_PRINTER(print_message, stdout, true, 0)
DEFINE_PRINTER(append_printed_message, stdout, false, 0)
DEFINE_PRINTER(print_error, stderr, true, 0)
DEFINE_PRINTER(append_printed_error, stderr, false, 0)
DEFINE_PRINTER(print_errnum, stderr, true, 1)
DEFINE_PRINTER(append_printed_errnum, stderr, false, 1)

int flush_output(bool permanently)
{
	(void)permanently;
	const int a = fflush(stdout), b = fflush(stderr);
	return (a || b) ? EXIT_FAILURE : EXIT_SUCCESS;
}

/* ---------------------------------------------------------------- help */

int help_file_formats(void)
{
	print_message(false,
		      "Input/output file formats:\n"
		      "  ASCII text (default): tab-separated values, one matrix row per line; optional\n"
		      "    column headers (first line) and row labels (first column); numeric ones need\n"
		      "    -c / -r.\n"
		      "  Non-native binary (-b 0 / -e 0): file signature, uint32 dimensions, little-endian\n"
		      "    FP64 data.\n"
		      "  Native binary (-b 1 / -e 1): native-endian dimensions and 'real' data, no signature.\n");
	return EXIT_SUCCESS;
}

int help_nmf(void)
{
	print_message(false,
		      "NMF options:\n"
		      "  -k <rank>      factorization rank (default %" PRI_IDX ")\n"
		      "  -i <iters>     maximum number of iterations (default %" PRI_IDX ")\n"
		      "  -j <iters>     iterations between convergence tests (default %" PRI_IDX ")\n"
		      "  -t <tests>     stop after this many unchanged tests, plus one (default %" PRI_IDX ")\n"
		      "  -z <device>    GPU device (default %" PRI_IDX ")\n",
		      DEFAULT_K, DEFAULT_NITERS, DEFAULT_NITER_CONV, DEFAULT_STOP_THRESHOLD, DEFAULT_GPU_DEVICE);
	return EXIT_SUCCESS;
}

int print_nmf_gpu_help(char const *execname)
{
	print_message(false,
		      "Usage: %s <filename> [-b <native>] [-c] [-r] [-e <native>] [-k <rank>] [-i <iters>]\n"
		      "          [-j <iters>] [-t <tests>] [-z <device>] [-h]\n\n",
		      execname ? execname : "NMF_GPU");
	help_file_formats();
	return help_nmf();
}

/* ---------------------------------------------------------------- limits */

int set_matrix_limits(index_t data_alignment, index_t max_dimension, size_t max_nitems)
{
	const index_t align = data_alignment ? data_alignment : DEFAULT_MEMORY_ALIGNMENT;
	const index_t max_dim = (max_dimension >= align) ? max_dimension : IDX_MAX;
	const size_t default_items = SIZE_MAX / (2 * sizeof(real));
	size_t items = (max_nitems >= (size_t)align) ? (max_nitems < default_items ? max_nitems : default_items)
						     : default_items;
	items -= items % align;
	size_t pitch = items / 2 < (size_t)max_dim ? items / 2 : (size_t)max_dim;
	pitch -= pitch % align;
	const index_t two_or_align = align > 2 ? align : 2;
	const size_t npd = items / two_or_align < (size_t)max_dim ? items / two_or_align : (size_t)max_dim;
	memory_alignment = align;
	matrix_max_num_items = items;
	matrix_max_pitch = (index_t)pitch;
	matrix_max_non_padded_dim = (index_t)npd;
	return EXIT_SUCCESS;
}

int set_default_matrix_limits(void)
{
	return set_matrix_limits(0, 0, 0);
}

index_t get_padding(index_t dim)
{
	const index_t r = dim % memory_alignment;
	return r ? dim + (memory_alignment - r) : dim;
}

size_t next_power_2(size_t x)
{
	if (x <= 1)
		return 1;
	size_t p = 1;
	while (p < x && p <= SIZE_MAX / 2)
		p <<= 1;
	return p;
}

size_t prev_power_2(size_t x)
{
	if (x <= 1)
		return 1;
	size_t p = 1;
	while (p <= x / 2)
		p <<= 1;
	return p;
}

/* ---------------------------------------------------------------- options */

/* Parses a decimal integer option value in [lo, hi]; returns 0 on success. */
static int parse_index(char const *s, intmax_t lo, intmax_t hi, index_t *out)
{
	errno = 0;
	char *end = NULL;
	const intmax_t v = strtoimax(s, &end, 10);
	if (!s[0] || *end || errno || v < lo || v > hi)
		return 1;
	*out = (index_t)v;
	return 0;
}

int check_arguments(int argc, char const *restrict *restrict argv, bool *restrict help_requested,
		    struct input_arguments *restrict arguments)
{
	if (argc <= 0 || !argv || !help_requested || !arguments) {
		print_errnum(false, EFAULT, "check_arguments");
		return EXIT_FAILURE;
	}
	struct input_arguments a = {0};
	a.input_file_fmt = ASCII_TEXT_FMT;
	a.output_file_fmt = ASCII_TEXT_FMT;
	a.k = DEFAULT_K;
	a.kp = get_padding(DEFAULT_K);
	a.nIters = DEFAULT_NITERS;
	a.niter_test_conv = DEFAULT_NITER_CONV;
	a.stop_threshold = DEFAULT_STOP_THRESHOLD;
	a.gpu_device = DEFAULT_GPU_DEVICE;

	/* optind = 0 makes glibc's getopt re-initialise completely (argument permutation state
	 * included), so a previous parse in the process cannot leak into this one. */
	optarg = NULL;
	optind = 0;
	opterr = 0;
	optopt = 0;
	int opt;
	index_t v;
	while ((opt = getopt(argc, (char *const *)argv, ":B:b:CcE:e:HhI:i:J:j:K:k:RrT:t:Z:z:")) != -1) {
		switch (opt) {
		case 'B':
		case 'b':
			if (parse_index(optarg, 0, INTMAX_MAX, &v)) {
				print_error(false, "Error: Invalid binary mode for input file: '%s'\n"
						   "It must be a non-negative integer value.\n", optarg);
				return EXIT_FAILURE;
			}
			a.input_file_fmt = v ? NATIVE_BINARY_FMT : NON_NATIVE_BINARY_FMT;
			break;
		case 'C':
		case 'c':
			a.numeric_hdrs = true;
			break;
		case 'E':
		case 'e':
			if (parse_index(optarg, 0, INTMAX_MAX, &v)) {
				print_error(false, "Error: Invalid binary mode for output file(s): '%s'\n"
						   "It must be a non-negative integer value.\n", optarg);
				return EXIT_FAILURE;
			}
			a.output_file_fmt = v ? NATIVE_BINARY_FMT : NON_NATIVE_BINARY_FMT;
			break;
		case 'H':
		case 'h':
			*help_requested = true;
			return EXIT_SUCCESS;
		case 'I':
		case 'i':
			if (parse_index(optarg, 1, IDX_MAX, &a.nIters)) {
				print_error(false, "Error: Invalid number of iterations: '%s'\n"
						   "It must be a positive integer value less than or equal to %" PRI_IDX ".\n",
					    optarg, IDX_MAX);
				return EXIT_FAILURE;
			}
			break;
		case 'J':
		case 'j':
			if (parse_index(optarg, 1, IDX_MAX, &a.niter_test_conv)) {
				print_error(false, "Error: Invalid number of iterations for convergence test: '%s'\n"
						   "It must be a positive integer value less than or equal to %" PRI_IDX ".\n",
					    optarg, IDX_MAX);
				return EXIT_FAILURE;
			}
			break;
		case 'K':
		case 'k':
			if (parse_index(optarg, 2, matrix_max_pitch, &a.k)) {
				print_error(false, "Error: invalid factorization rank: '%s'\n"
						   "It must be an integer value in the range [2 .. %" PRI_IDX "].\n",
					    optarg, matrix_max_pitch);
				return EXIT_FAILURE;
			}
			a.kp = get_padding(a.k);
			break;
		case 'R':
		case 'r':
			a.numeric_lbls = true;
			break;
		case 'T':
		case 't':
			if (parse_index(optarg, 1, IDX_MAX, &a.stop_threshold)) {
				print_error(false, "Error: invalid stopping threshold '%s'\n"
						   "It must be a positive integer value less than or equal to %" PRI_IDX ".\n",
					    optarg, IDX_MAX);
				return EXIT_FAILURE;
			}
			break;
		case 'Z':
		case 'z':
			if (parse_index(optarg, 0, IDX_MAX, &a.gpu_device)) {
				print_error(false, "Error: invalid basis device ID number '%s'\n"
						   "It must be a non-negative integer value less than or equal to %" PRI_IDX
						   ".\n", optarg, IDX_MAX);
				return EXIT_FAILURE;
			}
			break;
		case ':':
			print_error(false, "Error: option -%c requires an argument.\nSee help (option '-h').\n", optopt);
			return EXIT_FAILURE;
		default:
			break;  /* unknown options are ignored, as in the reference */
		}
	}
	if (optind >= argc) {
		print_error(false, "Error: No filename. Not enough arguments.\nSee help (option '-h').\n");
		return EXIT_FAILURE;
	}
	a.filename = argv[optind];
	a.idx_other_args = (index_t)(optind + 1);
	optarg = NULL;
	optind = 1;
	opterr = optopt = 0;
	*arguments = a;
	*help_requested = false;
	return EXIT_SUCCESS;
}

/* ---------------------------------------------------------------- connectivity */

/*
 * Number of pairs (i < j) whose co-membership changed.  Counted through the contingency
 * table of (class, previous class) pairs when the labels are small (the usual case, labels
 * < k), which gives the same integer as the pairwise loop of the reference:
 *   #{i<j: [c_i==c_j] != [l_i==l_j]} = sum_a C(r_a,2) + sum_b C(s_b,2) - 2 sum_ab C(n_ab,2).
 * Labels >= 64 fall back to the pairwise loop.
 */
size_t get_difference(index_t const *cls, index_t const *last, index_t m)
{
	enum { L = 64 };
	index_t maxl = 0;
	for (index_t i = 0; i < m; i++) {
		if (cls[i] > maxl)
			maxl = cls[i];
		if (last[i] > maxl)
			maxl = last[i];
	}
	if (maxl < L) {
		size_t tab[L * L];
		size_t rows[L] = {0}, cols[L] = {0};
		memset(tab, 0, sizeof(size_t) * (maxl + 1) * L);
		for (index_t i = 0; i < m; i++)
			tab[(size_t)cls[i] * L + last[i]]++;
		size_t both = 0, sr = 0, sc = 0;
		for (index_t a = 0; a <= maxl; a++)
			for (index_t b = 0; b <= maxl; b++) {
				const size_t v = tab[(size_t)a * L + b];
				both += v * (v ? v - 1 : 0) / 2;
				rows[a] += v;
				cols[b] += v;
			}
		for (index_t a = 0; a <= maxl; a++) {
			sr += rows[a] * (rows[a] ? rows[a] - 1 : 0) / 2;
			sc += cols[a] * (cols[a] ? cols[a] - 1 : 0) / 2;
		}
		return sr + sc - 2 * both;
	}
	size_t diff = 0;
	for (index_t i = 0; i + 1 < m; i++)
		for (index_t j = i + 1; j < m; j++)
			diff += (size_t)((cls[i] == cls[j]) != (last[i] == last[j]));
	return diff;
}

/* ---------------------------------------------------------------- seed */

/* src/common.c:1218-1258: FIXED_SEED in a NMFGPU_FIXED_INIT build, else /dev/urandom (the
 * clock if that fails).  Addition: a decimal NMFGPU_SEED in the environment is returned
 * instead, so that a driver built on this library (the reference's own bin/NMF_GPU included)
 * can be run reproducibly. */
index_t get_seed(void)
{
	const char *env = getenv("NMFGPU_SEED");
	if (env && *env)
		return (index_t)strtoul(env, NULL, 10);
	index_t seed = FIXED_SEED;
#ifndef NMFGPU_FIXED_INIT
	FILE *f = fopen("/dev/urandom", "r");
	if (!f || fread(&seed, sizeof seed, 1, f) != 1) {
		struct timeval tv;
		gettimeofday(&tv, NULL);
		seed = (index_t)tv.tv_usec;
	}
	if (f)
		fclose(f);
#endif
	return seed;
}
