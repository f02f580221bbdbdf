/*
 * NMF_routines.h -- drop-in for the reference's include/NMF_routines.h (release build:
 * no debug, no profiling, GPU random generator), on HIP / MI355X.
 *
 * update_H / update_W perform one half-step each, exactly as src/NMF_routines.c:704-1000
 * and :1344-1636 define it (column sums of the other factor, then the multiplicative
 * update), but as two streaming gfx950 kernels instead of GEMM / divide / GEMM / mul-div.
 */
#ifndef NMFGPU_NMF_ROUTINES_H
#define NMFGPU_NMF_ROUTINES_H

#include <stdbool.h>
#include <hip/hip_runtime_api.h>

#include "index_type.h"
#include "real_type.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Cursor state of the blockwise plan (include/NMF_routines.h:160-170); with V resident the
 * plan has one block and these stay at their initial values. */
extern index_t pBLN, pBLM;
extern int stepN, stepM;
extern index_t psNMF_N, psNMF_M;
extern index_t colIdx, rowIdx;

/* Seeds the W0/H0 generator (include/NMF_routines.h:181-188). */
int init_random(index_t seed);

/* Releases the generator (include/NMF_routines.h:192-198). */
int destroy_random(void);

/* d_A[height][width] (pitch <padding>) = uniform (0,1] from the seeded generator; successive
 * calls draw successive streams (W first, then H).  If <event_A> is non-NULL the operation
 * is recorded on it (include/NMF_routines.h:200-232). */
int set_random_values(real *d_A, index_t height, index_t width, index_t padding,
		      char const *matrix_name_dA, hipStream_t stream_A, hipEvent_t *event_A);

/* H half-step on d_H with d_W and V (include/NMF_routines.h:234-250). */
int update_H(void);

/* W half-step on d_W with the new d_H and V (include/NMF_routines.h:252-268). */
int update_W(void);

/* ld_classification = argmax of every row of d_H, downloaded to lh_classification
 * (include/NMF_routines.h:270-276).  Blocks until done. */
int get_classification(index_t *ld_classification, index_t *lh_classification);

/* dot_V = <V,V>, dot_VWH = <V-WH,V-WH> (include/NMF_routines.h:278-287); accumulated in FP64. */
int dot_product_VWH(real *dot_V, real *dot_VWH);

#ifdef __cplusplus
}
#endif

#endif /* NMFGPU_NMF_ROUTINES_H */
