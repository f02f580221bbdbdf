/*
 * NMF_GPU.c -- command-line driver: factorizes the matrix in a file, writes W and H.
 *
 * Behaves as the reference's bin/NMF_GPU (src/NMF_GPU.c:1034-1269): the same options
 * (check_arguments, src/common.c:847-1119), the same messages, and the same outputs:
 *   <file>_W<ext>   W, N x K, row labels of the input, headers Factor_0 .. Factor_{K-1}
 *   <file>_H<ext>   H transposed back, K x M, row labels Factor_i, the input's headers
 * (src/NMF_GPU.c:931-1017), plus, as an addition, the class vector of the final H in
 *   <file>_classes.txt   one line per input column: [header <tab>] class index.
 * The loop is nmfk_job_run (the reference's nmf(), src/NMF_GPU.c:312-924) on one job,
 * started from W then H uniform (0,1] drawn from get_seed() (or the NMFGPU_SEED
 * environment variable, an addition for reproducible runs).
 */
#include "common.h"
#include "index_type.h"
#include "nmfk.h"
#include "nmfk_io.h"
#include "real_type.h"

#include <hip/hip_runtime_api.h>

#include <inttypes.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int write_outputs(const char *filename, int fmt, const float *W, const float *H, uint32_t n, uint32_t m,
			 uint32_t k, const uint32_t *cls, const nmfk_tags *in)
{
	static const char *const ext[3] = {".txt", ".dat", ".native.dat"};
	const int tagged = in->name || in->headers || in->labels;
	char **factors = NULL;
	int e = 0;
	if (tagged) {
		factors = (char **)calloc(k, sizeof(char *));
		for (uint32_t i = 0; factors && i < k; i++) {
			factors[i] = (char *)malloc(32);
			if (factors[i])
				snprintf(factors[i], 32, "Factor_%" PRIu32, i);
			else
				e = 1;
		}
		if (!factors || e) {
			print_error(true, "Error initializing the output labels.\n");
			e = 1;
		}
	}
	const size_t len = strlen(filename) + 32;
	char *out = (char *)malloc(len);
	if (!out) {
		print_error(true, "Error allocating memory for the output file names.\n");
		e = 1;
	}
	if (!e) {
		nmfk_tags tw = {0};
		if (tagged) {
			tw.name = in->name;
			tw.headers = factors, tw.nheaders = k;
			tw.labels = in->labels, tw.nlabels = in->labels ? n : 0;
		}
		snprintf(out, len, "%s_W%s", filename, ext[fmt]);
		print_message(false, "Saving output file (W: %" PRIu32 " x %" PRIu32 ") to %s...\n", n, k, out);
		if (nmfk_matrix_save(out, fmt, W, n, k, k, 0, tagged ? &tw : NULL)) {
			print_error(true, "Error writing matrix W: %s\n", nmfk_last_error());
			e = 1;
		}
	}
	if (!e) {
		/* H is M x K on the device (the reference's transposed storage); the file holds K x M. */
		nmfk_tags th = {0};
		if (tagged) {
			/* native files skip every transformation: H is written as stored, M x K */
			const int asis = fmt == NMFK_FMT_NATIVE;
			th.name = in->name;
			th.headers = asis ? factors : in->headers;
			th.nheaders = asis ? k : (in->headers ? m : 0);
			th.labels = asis ? in->headers : factors;
			th.nlabels = asis ? (in->headers ? m : 0) : k;
		}
		snprintf(out, len, "%s_H%s", filename, ext[fmt]);
		if (nmfk_matrix_save(out, fmt, H, m, k, k, fmt != NMFK_FMT_NATIVE, tagged ? &th : NULL)) {
			print_error(true, "Error writing matrix H: %s\n", nmfk_last_error());
			e = 1;
		}
	}
	if (!e && cls) {
		snprintf(out, len, "%s_classes.txt", filename);
		FILE *f = fopen(out, "w");
		int bad = !f;
		for (uint32_t j = 0; f && j < m; j++) {
			if (in->headers)
				bad |= fprintf(f, "%s\t", in->headers[j] ? in->headers[j] : "") < 0;
			bad |= fprintf(f, "%" PRIu32 "\n", cls[j]) < 0;
		}
		if (f)
			bad |= fclose(f) != 0;
		if (bad) {
			print_error(true, "Error writing the class vector to %s.\n", out);
			e = 1;
		}
	}
	free(out);
	for (uint32_t i = 0; factors && i < k; i++)
		free(factors[i]);
	free(factors);
	return e ? EXIT_FAILURE : EXIT_SUCCESS;
}

int main(int argc, char *argv[])
{
	process_id = 0;
	num_act_processes = num_processes = 1;
	set_default_matrix_limits();

	bool help = false;
	struct input_arguments a;
	if (check_arguments(argc, (char const *restrict *restrict)argv, &help, &a) != EXIT_SUCCESS)
		return EXIT_FAILURE;
	if (help)
		return print_nmf_gpu_help(argv[0]);

	const uint32_t k = a.k, nIters = a.nIters, T = a.niter_test_conv, S = a.stop_threshold;
	print_message(false, "\t<<< NMF-mGPU: Non-negative Matrix Factorization on GPU >>>\n"
			     "\t\t\t\tSingle-GPU version\n");
	if (nmfk_set_device((int)a.gpu_device)) {
		print_error(true, "Error selecting GPU device %" PRI_IDX ": %s\n", a.gpu_device, nmfk_last_error());
		return EXIT_FAILURE;
	}

	float *V = NULL;
	uint32_t n = 0, m = 0;
	nmfk_tags tags;
	const int infmt = a.input_file_fmt == NON_NATIVE_BINARY_FMT ? NMFK_FMT_BINARY
			  : a.input_file_fmt == NATIVE_BINARY_FMT  ? NMFK_FMT_NATIVE
								   : NMFK_FMT_TEXT;
	const int outfmt = a.output_file_fmt == NON_NATIVE_BINARY_FMT ? NMFK_FMT_BINARY
			   : a.output_file_fmt == NATIVE_BINARY_FMT  ? NMFK_FMT_NATIVE
								     : NMFK_FMT_TEXT;
	if (nmfk_matrix_load(a.filename, infmt, a.numeric_hdrs, a.numeric_lbls, &V, &n, &m, &tags)) {
		print_error(true, "Error reading input file: %s\n", nmfk_last_error());
		return EXIT_FAILURE;
	}
	if (k > (n < m ? n : m)) {
		print_error(true, "Error: invalid factorization rank: K=%" PRIu32 ".\nIt must not be greater "
				  "than any of matrix dimensions.\n", k);
		free(V), nmfk_tags_free(&tags);
		return EXIT_FAILURE;
	}

	int status = EXIT_FAILURE;
	float *W = (float *)malloc((size_t)n * k * sizeof(float));
	float *H = (float *)malloc((size_t)m * k * sizeof(float));
	uint32_t *cls = (uint32_t *)malloc((size_t)m * sizeof(uint32_t));
	/* A V that does not fit the device is streamed from host memory in row blocks, as the
	 * reference switches to blockwise processing (src/GPU_setup.c:902-1270).
	 * NMFGPU_STREAM_ROWS (environment) forces streaming with that many rows per block (0 =
	 * the library's default). */
	const char *env_rows = getenv("NMFGPU_STREAM_ROWS");
	bool streamed = env_rows && *env_rows;
	uint32_t stream_rows = streamed ? (uint32_t)strtoul(env_rows, NULL, 10) : 0;
	if (!streamed) {
		size_t free_b = 0, total_b = 0;
		if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && nmfk_job_bytes(n, m, k) > free_b / 10 * 9)
			streamed = true;
	}
	nmfk_job *job = NULL;
	/* A streamed V is copied once into pinned memory, which the device reads directly at the
	 * PCIe rate; if that allocation fails it is staged from the loaded matrix instead. */
	float *Vpin = NULL;
	if (streamed && W && H && cls) {
		Vpin = (float *)nmfk_host_alloc((size_t)n * m * sizeof(float));
		if (Vpin)
			memcpy(Vpin, V, (size_t)n * m * sizeof(float));
	}
	if (W && H && cls)
		job = streamed ? nmfk_job_create_streamed(n, m, k, Vpin ? Vpin : V, m, stream_rows, Vpin != NULL)
			       : nmfk_job_create(n, m, k);
	if (job && streamed) {
		uint32_t br = 0, nb = 0;
		nmfk_job_streamed_info(job, &br, &nb, NULL);
		print_message(false, "Input matrix streamed from host memory in %" PRIu32 " blocks of %" PRIu32
				     " rows.\n", nb, br);
	}
	if (!job) {
		print_error(true, "Error setting up the factorization (N=%" PRIu32 ", M=%" PRIu32 ", K=%" PRIu32 "): %s\n", n,
			    m, k, (W && H && cls) ? nmfk_last_error() : "out of host memory");
		goto done;
	}
	/* the reference's precision: FP32 arithmetic at every rank (streamed jobs always are) */
	if (!streamed && nmfk_job_kernel_path(job) == 3)
		nmfk_job_set_kernel_path(job, 2);
	/* get_seed() returns NMFGPU_SEED from the environment when set, for reproducible runs. */
	const index_t seed = get_seed();
	if ((!streamed && nmfk_job_set_V(job, V, m)) || nmfk_job_init_random(job, seed)) {
		print_error(true, "Error initializing the factorization: %s\n", nmfk_last_error());
		goto done;
	}
	print_message(false, "Starting NMF( K=%" PRIu32 " )...\n", k);
	nmfk_options opt = {k, nIters, T, S};
	nmfk_result r;
	if (nmfk_job_run(job, &opt, &r)) {
		print_error(true, "Error in NMF: %s\n", nmfk_last_error());
		goto done;
	}
	if (r.converged)
		print_message(false, "NMF: Algorithm converged in %" PRIu32 " iterations.\n", r.iterations);
	else
		print_message(false, "NMF: %" PRIu32 " iterations performed.\n", r.iterations);
	double dv = 0.0, dvwh = 0.0, kl = 0.0;
	if (nmfk_job_dot_kl(job, &dv, &dvwh, &kl) || nmfk_job_get_W(job, W, k) || nmfk_job_get_H(job, H, k) ||
	    nmfk_job_get_classification(job, cls)) {
		print_error(true, "Error reading the results: %s\n", nmfk_last_error());
		goto done;
	}
	print_message(false, "Distance between V and W*H: %g\n", (double)sqrtf((float)dvwh) / (double)sqrtf((float)dv));
	print_message(false, "Generalized KL divergence D(V||W*H): %g\n", kl);
	status = write_outputs(a.filename, outfmt, W, H, n, m, k, nIters >= T ? cls : NULL, &tags);
	if (status == EXIT_SUCCESS)
		print_message(false, "Done.\n");
done:
	nmfk_job_destroy(job);
	nmfk_host_free(Vpin);
	free(W), free(H), free(cls), free(V);
	nmfk_tags_free(&tags);
	return status;
}
