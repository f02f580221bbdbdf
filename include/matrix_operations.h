/*
 * matrix_operations.h -- drop-in for the reference's include/matrix_operations.h (release
 * build), on HIP / MI355X.  Same entry points, argument meaning and return convention
 * (EXIT_SUCCESS / EXIT_FAILURE, message on stderr) as include/matrix_operations.h:100-406.
 * Every operation is asynchronous on its stream unless stated otherwise.
 */
#ifndef NMFGPU_MATRIX_OPERATIONS_H
#define NMFGPU_MATRIX_OPERATIONS_H

#include <stdbool.h>
#include <stddef.h>
#include <hip/hip_runtime_api.h>

#include "index_type.h"
#include "real_type.h"

#ifdef __cplusplus
extern "C" {
#endif

struct matrix_tags_t;

/* Largest height for <pitch> regardless of free memory (include/matrix_operations.h:113-118). */
index_t gpu_max_height(index_t pitch);

/* Largest number of items of a device matrix (include/matrix_operations.h:122-126). */
size_t gpu_max_nitems(void);

/* Kernel launch parameters for rank pitch <pitch> (include/matrix_operations.h:130-133). */
void init_kernel_params(index_t pitch);

/* Synchronously downloads and prints part of a device matrix; <mt> (row/column tags) may be
 * NULL (include/matrix_operations.h:137-153). */
int show_device_matrix(void const *dMatrix, index_t nrows, index_t ncols, index_t pitch, bool real_data,
		       bool transpose, bool all_processes, struct matrix_tags_t const *mt);

/* d_A = uniform (0,1] values, width <= padding (include/matrix_operations.h:157-177). */
int matrix_random(real *d_A, index_t height, index_t width, index_t padding, char const *matrix_name,
		  hipStream_t stream_A, hipEvent_t *event_A);

/* d_accum_A[j] = sum_i d_A[i][j] over <height> rows of pitch <pitch>; d_Tmp is scratch of at
 * least (height / 4) * pitch items, ignored when height <= 2.  Recorded on event_reduction
 * (include/matrix_operations.h:181-208). */
int matrix_to_row(real const *d_A, index_t height, index_t pitch, char const *matrix_name, real *d_Tmp,
		  real *d_accum_A, hipStream_t stream_AccA);

/* d_A = d_B ./ d_A (div_operator) or d_B - d_A, over height x pitch items; waits for
 * <event_B> and records it afterwards (include/matrix_operations.h:212-239). */
int matrix_div_sub(real *d_A, real const *d_B, index_t height, index_t pitch, char const *matrix_name_A,
		   char const *matrix_name_B, bool div_operator, hipStream_t stream_A, hipEvent_t event_B);

/* d_A[i][j] = (d_A[i][j] * d_Aux[i][j]) / d_accum_B[j]; waits for event_reduction
 * (include/matrix_operations.h:243-265). */
int matrix_mul_div(real *d_A, real const *d_Aux, real const *d_accum_B, index_t height, index_t pitch,
		   char const *matrix_name_A, hipStream_t stream_A);

/* d_A = max(d_A, R_MIN) with NaN -> R_MIN (include/matrix_operations.h:269-287). */
int matrix_adjust(real *d_A, index_t height, index_t pitch, char const *matrix_name_A, hipStream_t stream_A,
		  hipEvent_t *event_A);

/* d_Idx[i] = column of the maximum of row i (first <width> columns; reference tie rule;
 * width <= 192) (include/matrix_operations.h:291-313). */
int matrix_idx_max(real const *d_A, index_t width, index_t pitch, index_t height, char const *matrix_name_A,
		   hipStream_t stream_A, index_t *d_Idx);

/* Host -> device, height x pitch items (include/matrix_operations.h:317-342). */
int upload_matrix(real const *A, index_t height, index_t pitch, real *d_A, char const *matrix_name_A,
		  hipStream_t stream_A, hipEvent_t *event_A);

/* Host -> device of the block pA[strow..strow+height][stcol..stcol+block_pitch] of a row-major
 * matrix of pitch <pitch> (include/matrix_operations.h:346-377). */
int upload_matrix_partial(real const *pA, index_t height, index_t pitch, index_t strow, index_t stcol,
			  char const *matrix_name_A, index_t block_pitch, real *d_A, hipStream_t stream_A,
			  hipEvent_t event_A);

/* Device -> host of nitems items of data_size bytes; blocks until done
 * (include/matrix_operations.h:381-401). */
int download_matrix(void *A, size_t nitems, size_t data_size, void const *d_A, char const *matrix_name_dA,
		    hipStream_t stream_A);

#ifdef __cplusplus
}
#endif

#endif /* NMFGPU_MATRIX_OPERATIONS_H */
