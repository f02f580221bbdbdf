/*
 * GPU_setup.c -- drop-in for the reference's src/GPU_setup.c on HIP / MI355X.
 *
 * initialize_GPU selects the device and fixes the padded rank; setup_GPU creates the job
 * that holds every device buffer of the factorization and points the reference's device
 * globals into it (src/GPU_setup.c:1289-1610 allocate the same set: d_Vrow, d_Vcol, d_W,
 * d_H, d_WH, d_Aux, d_accum, d_classification).
 */
#include "GPU_setup.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "common.h"
#include "matrix_operations.h"
#include "nmfk_internal.h"

index_t computeCapability = 0, computeCapability_minor = 0;
index_t maxThreadsPerBlock = 1024, multiProcessorCount = 0, maxThreadsPerMultiProcessor = 0;
index_t maxGridSizeX = 0, maxGridSizeY = 0;
index_t threadsPerBlock = 256, threadsPerBlock_pitch = 256, maxBlockHeight_pitch = 0;
bool mappedHostMemory = false;
size_t gpu_max_num_items = 0;
block_t block_N, block_M;

hipEvent_t event_Vrow = NULL, event_Vcol = NULL, event_W = NULL, event_H = NULL, event_reduction = NULL;
hipStream_t stream_Vrow = NULL, stream_Vcol = NULL, stream_W = NULL, stream_H = NULL;
hipStream_t *streams_NMF = NULL;
index_t num_streams_NMF = 0;

real *h_accum = NULL;

real *d_Vrow = NULL, *d_Vcol = NULL, *d_H = NULL, *d_W = NULL, *d_WH = NULL, *d_Aux = NULL, *d_accum = NULL;
index_t *d_classification = NULL, *d_last_classification = NULL;
real const *d_scalars = NULL, *d_zero = NULL, *d_one = NULL;

nmfk_job *nmfk_dropin_job = NULL;
uint64_t nmfk_dropin_seed = 0;
uint32_t nmfk_dropin_rng_calls = 0;
int nmfk_dropin_rng_ready = 0;

static int g_device = -1;

int check_cuda_status_st(hipError_t cuda_status)
{
	hipError_t e2 = hipDeviceSynchronize();
	hipError_t e3 = hipGetLastError();
	hipError_t e = cuda_status != hipSuccess ? cuda_status : (e2 != hipSuccess ? e2 : e3);
	if (e != hipSuccess) {
		print_error(true, "HIP error: %s\n", hipGetErrorString(e));
		return EXIT_FAILURE;
	}
	return EXIT_SUCCESS;
}

int check_cuda_status(void)
{
	return check_cuda_status_st(hipSuccess);
}

size_t initialize_GPU(index_t dev_id, index_t factorization_rank)
{
	hipDeviceProp_t prop;
	if (hipSetDevice((int)dev_id) != hipSuccess || hipGetDeviceProperties(&prop, (int)dev_id) != hipSuccess) {
		print_error(true, "Error: could not initialize HIP device %" PRI_IDX ".\n", dev_id);
		return 0;
	}
	g_device = (int)dev_id;
	computeCapability = (index_t)prop.major;
	computeCapability_minor = (index_t)prop.minor;
	maxThreadsPerBlock = (index_t)prop.maxThreadsPerBlock;
	multiProcessorCount = (index_t)prop.multiProcessorCount;
	maxThreadsPerMultiProcessor = (index_t)prop.maxThreadsPerMultiProcessor;
	maxGridSizeX = (index_t)prop.maxGridSize[0];
	maxGridSizeY = (index_t)prop.maxGridSize[1];
	threadsPerBlock = 256;
	/* Rows of V and Vt are padded to 4 items (16-byte vector loads). */
	set_matrix_limits(4, 0, 0);
	if (factorization_rank < 1 || factorization_rank > NMFK_KMAX) {
		print_error(true, "Error: factorization rank %" PRI_IDX " outside [1 .. %d].\n", factorization_rank,
			    NMFK_KMAX);
		return 0;
	}
	K = factorization_rank;
	Kp = nmfk_kt(factorization_rank);
	size_t free_b = 0, total_b = 0;
	if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
		print_error(true, "Error: hipMemGetInfo failed.\n");
		return 0;
	}
	gpu_max_num_items = total_b / sizeof(real);
	init_kernel_params(Kp);
	return free_b;
}

static int create_event(hipEvent_t *ev)
{
	return hipEventCreateWithFlags(ev, hipEventDisableTiming) == hipSuccess ? 0 : 1;
}

int setup_GPU(size_t mem_size, bool do_classf)
{
	(void)do_classf;
	if (g_device < 0) {
		print_error(true, "Error: setup_GPU() called before initialize_GPU().\n");
		return EXIT_FAILURE;
	}
	if (K > N || K > M) {
		/* src/NMF_GPU.c:1130 */
		print_error(true, "Error: invalid factorization rank: K=%" PRI_IDX " must be <= min(N=%" PRI_IDX
				  ", M=%" PRI_IDX ").\n", K, N, M);
		return EXIT_FAILURE;
	}
	/* When V does not fit in <mem_size>, the input is processed in blocks of rows streamed
	 * from the host matrix Vrow (pitch Mp), as the reference's setup_GPU() switches to
	 * blockwise mode (src/GPU_setup.c:902-1270): the largest block (halving from N rows) whose
	 * buffers fit.  NMFGPU_STREAM_ROWS (environment) forces this mode with that many rows per
	 * block. */
	const size_t need = nmfk_job_bytes(N, M, K);
	const char *env_rows = getenv("NMFGPU_STREAM_ROWS");
	const bool forced = env_rows && *env_rows;
	const bool streamed = forced || (mem_size && need > mem_size);
	if (streamed) {
		const index_t ldv = (M + 3u) & ~3u;
		if (!Vrow || Mp != ldv) {
			print_error(true, "Error: blockwise processing needs the host input matrix Vrow with pitch Mp=%" PRI_IDX
					  " (set: %s, Mp=%" PRI_IDX ").\n", ldv, Vrow ? "yes" : "no", Mp);
			return EXIT_FAILURE;
		}
		uint32_t rows = forced ? (uint32_t)strtoul(env_rows, NULL, 10) : 0;
		if (!forced) {
			rows = (N + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS * NMFK_FIN_ROWS;
			while (rows > NMFK_FIN_ROWS && nmfk_streamed_bytes(N, M, K, rows) > mem_size)
				rows = (rows / 2 + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS * NMFK_FIN_ROWS;
			if (nmfk_streamed_bytes(N, M, K, rows) > mem_size) {
				print_error(true, "Error: not enough device memory: %zu bytes required for blocks of %u rows, "
						  "%zu available.\n", nmfk_streamed_bytes(N, M, K, rows), rows, mem_size);
				return EXIT_FAILURE;
			}
		}
		nmfk_dropin_job = nmfk_job_create_streamed(N, M, K, Vrow, Mp, rows, 1 /* getHostMemory is pinned */);
	} else {
		nmfk_dropin_job = nmfk_job_create(N, M, K);
	}
	if (!nmfk_dropin_job) {
		print_error(true, "Error: %s\n", nmfk_last_error());
		return EXIT_FAILURE;
	}
	nmfk_job *j = nmfk_dropin_job;
	/* The reference computes in FP32 throughout: the drop-in keeps every rank on FP32
	 * arithmetic (BF16 contractions are opt-in through the job API). */
	if (!streamed && j->planH.path == NMFK_PATH_MFMA_BF16 && nmfk_job_set_kernel_path(j, NMFK_PATH_MFMA32)) {
		print_error(true, "Error: %s\n", nmfk_last_error());
		return EXIT_FAILURE;
	}
	NpP = N, MpP = M, bN = 0, bM = 0;
	Mp = (index_t)j->ldv;
	MpPp = Mp;
	/* The job's padded rank (which depends on N and M, unknown at initialize_GPU) is the one
	 * the host matrices W and H take; the caller sizes them after this call. */
	Kp = j->kt;
	init_kernel_params(Kp);
	/* Streamed: the caller's first upload of block_N.BL[0] rows lands in slot 0 (the job
	 * copies every block itself); resident: one copy of V, as in the reference when V fits. */
	d_Vrow = streamed ? j->Vb[0] : j->V;
	d_Vcol = d_Vrow;
	d_W = j->W;
	d_H = j->H;
	d_WH = NULL;
	d_Aux = j->part;
	d_accum = j->denW;
	d_classification = j->cls;
	d_last_classification = j->last;
	stream_Vrow = stream_Vcol = stream_W = stream_H = j->stream;
	streams_NMF = (hipStream_t *)malloc(sizeof(hipStream_t));
	if (!streams_NMF)
		return EXIT_FAILURE;
	streams_NMF[0] = j->stream;
	num_streams_NMF = 1;
	memset(&block_N, 0, sizeof block_N);
	memset(&block_M, 0, sizeof block_M);
	if (streamed && j->nblocks > 1) {
		block_N.BL[0] = block_N.BLp[0] = j->brows, block_N.num_steps[0] = N / j->brows;
		block_N.BL[1] = block_N.BLp[1] = N % j->brows, block_N.num_steps[1] = (N % j->brows) != 0;
		print_message(false, "Input matrix processed in blocks of %" PRIu32 " rows (%" PRIu32 " blocks).\n", j->brows,
			      j->nblocks);
	} else {
		block_N.BL[0] = N, block_N.BLp[0] = N, block_N.num_steps[0] = 1;
	}
	block_M.BL[0] = M, block_M.BLp[0] = Mp, block_M.num_steps[0] = 1;
	if (create_event(&event_Vrow) || create_event(&event_Vcol) || create_event(&event_W) || create_event(&event_H) ||
	    create_event(&event_reduction)) {
		print_error(true, "Error: could not create HIP events.\n");
		return EXIT_FAILURE;
	}
	static const real scal[2] = {0.0f, 1.0f};
	real *ds = NULL;
	if (hipMalloc((void **)&ds, sizeof scal) != hipSuccess ||
	    hipMemcpy(ds, scal, sizeof scal, hipMemcpyHostToDevice) != hipSuccess) {
		print_error(true, "Error: could not allocate device scalars.\n");
		return EXIT_FAILURE;
	}
	d_scalars = ds, d_zero = ds, d_one = ds + 1;
	return EXIT_SUCCESS;
}

int shutdown_GPU(void)
{
	return hipDeviceSynchronize() == hipSuccess ? EXIT_SUCCESS : EXIT_FAILURE;
}

int finalize_GPU_device(void)
{
	int st = EXIT_SUCCESS;
	hipEvent_t *evs[5] = {&event_Vrow, &event_Vcol, &event_W, &event_H, &event_reduction};
	for (int i = 0; i < 5; i++)
		if (*evs[i]) {
			hipEventDestroy(*evs[i]);
			*evs[i] = NULL;
		}
	if (d_scalars)
		hipFree((void *)d_scalars);
	d_scalars = d_zero = d_one = NULL;
	if (nmfk_dropin_job)
		nmfk_job_destroy(nmfk_dropin_job);
	nmfk_dropin_job = NULL;
	free(streams_NMF);
	streams_NMF = NULL;
	num_streams_NMF = 0;
	stream_Vrow = stream_Vcol = stream_W = stream_H = NULL;
	d_Vrow = d_Vcol = d_W = d_H = d_WH = d_Aux = d_accum = NULL;
	d_classification = d_last_classification = NULL;
	if (shutdown_GPU() != EXIT_SUCCESS)
		st = EXIT_FAILURE;
	return st;
}

void *getHostMemory(size_t size, bool wc, bool clear)
{
	void *p = NULL;
	unsigned flags = wc ? hipHostMallocWriteCombined : hipHostMallocDefault;
	if (hipHostMalloc(&p, size ? size : 1, flags) != hipSuccess) {
		print_error(true, "Error: could not allocate %zu bytes of pinned host memory.\n", size);
		return NULL;
	}
	if (clear)
		memset(p, 0, size);
	return p;
}

int freeHostMemory(void *pHost, char const *pHost_name)
{
	if (pHost && hipHostFree(pHost) != hipSuccess) {
		print_error(true, "Error: could not free pinned host memory '%s'.\n", pHost_name ? pHost_name : "");
		return EXIT_FAILURE;
	}
	return EXIT_SUCCESS;
}

int init_randomGenerator(void)
{
	nmfk_dropin_rng_ready = 1;
	nmfk_dropin_rng_calls = 0;
	return EXIT_SUCCESS;
}

int set_randomGenerator_seed(unsigned long long seed)
{
	if (!nmfk_dropin_rng_ready) {
		print_error(true, "Error: random generator not initialized.\n");
		return EXIT_FAILURE;
	}
	nmfk_dropin_seed = (uint64_t)seed;
	nmfk_dropin_rng_calls = 0;
	return EXIT_SUCCESS;
}

int init_GPU_random(unsigned long long seed)
{
	if (init_randomGenerator() != EXIT_SUCCESS)
		return EXIT_FAILURE;
	return set_randomGenerator_seed(seed);
}

int finalize_randomGenerator(void)
{
	nmfk_dropin_rng_ready = 0;
	return EXIT_SUCCESS;
}

int sync_GPU(hipStream_t stream)
{
	hipError_t e = hipStreamSynchronize(stream);
	if (e != hipSuccess) {
		print_error(true, "Error: hipStreamSynchronize: %s\n", hipGetErrorString(e));
		return EXIT_FAILURE;
	}
	return EXIT_SUCCESS;
}
