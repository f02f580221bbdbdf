/*
 * GPU_setup.h -- drop-in for the reference's include/GPU_setup.h on HIP / MI355X.
 *
 * Same entry points and globals as include/GPU_setup.h:160-433 of NMF-mGPU, with
 * hipStream_t / hipEvent_t in place of the CUDA types.  What changes underneath:
 *   - the whole input matrix is resident (MI355X holds 288 GB), so the block plan is one
 *     block per dimension (block_N.BL[0] = N, block_M.BL[0] = M, num_steps = 1) and d_Vcol
 *     aliases d_Vrow, as the reference does when V fits (src/GPU_setup.c:1289-1610);
 *   - all work of one factorization runs on one in-order stream; the stream and event
 *     globals are kept for callers and all name that stream;
 *   - d_WH is never allocated (W.H^T is formed in registers) and d_Aux holds the partial
 *     contractions.
 * The cuBLAS/cuRAND handles and helpers of the reference (cublas_handle,
 * curand_generator, getCublasErrorString, getCurandErrorString, check_cublas_status_st)
 * have no counterpart: no vendor BLAS or RNG is used.
 */
#ifndef NMFGPU_GPU_SETUP_H
#define NMFGPU_GPU_SETUP_H

#include <stdbool.h>
#include <stddef.h>
#include <hip/hip_runtime_api.h>

#include "index_type.h"
#include "real_type.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Block configuration of V (include/GPU_setup.h:170-174). */
typedef struct {
	index_t BL[2];
	index_t BLp[2];
	index_t num_steps[2];
} block_t;

extern index_t computeCapability, computeCapability_minor;  /* gfx major / minor (9, 5) */
extern index_t maxThreadsPerBlock, multiProcessorCount, maxThreadsPerMultiProcessor;
extern index_t maxGridSizeX, maxGridSizeY;
extern index_t threadsPerBlock, threadsPerBlock_pitch, maxBlockHeight_pitch;
extern bool mappedHostMemory;
extern size_t gpu_max_num_items;
extern block_t block_N, block_M;

extern hipEvent_t event_Vrow, event_Vcol, event_W, event_H, event_reduction;
extern hipStream_t stream_Vrow, stream_Vcol, stream_W, stream_H;
extern hipStream_t *streams_NMF;
extern index_t num_streams_NMF;

extern real *h_accum;

extern real *d_Vrow, *d_Vcol, *d_H, *d_W, *d_WH, *d_Aux, *d_accum;
extern index_t *d_classification, *d_last_classification;
extern real const *d_scalars, *d_zero, *d_one;

/* Waits for all device work and checks <cuda_status> (include/GPU_setup.h:305-325).
 * Return EXIT_SUCCESS or EXIT_FAILURE. */
int check_cuda_status_st(hipError_t cuda_status);
int check_cuda_status(void);

/* Selects device <dev_id>, sets memory_alignment and the padded rank Kp for rank
 * <factorization_rank>.  Returns the free device memory in bytes, or 0 on failure
 * (include/GPU_setup.h:329-343). */
size_t initialize_GPU(index_t dev_id, index_t factorization_rank);

/* Checks N, M, K, allocates every device buffer, streams and events (include/GPU_setup.h:345-356). */
int setup_GPU(size_t mem_size, bool do_classf);

int shutdown_GPU(void);
int finalize_GPU_device(void);

/* Pinned host memory (include/GPU_setup.h:370-393).  Must be freed before finalize. */
void *getHostMemory(size_t size, bool wc, bool clear);
int freeHostMemory(void *pHost, char const *pHost_name);

/* Random generator of W0/H0 (include/GPU_setup.h:395-425): Philox4x32-10 keyed by the seed. */
int init_randomGenerator(void);
int set_randomGenerator_seed(unsigned long long seed);
int init_GPU_random(unsigned long long seed);
int finalize_randomGenerator(void);

/* Blocks until all work on <stream> has completed. */
int sync_GPU(hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* NMFGPU_GPU_SETUP_H */
