/*
 * matrix_operations.c -- drop-in for the reference's src/matrix_operations.c on HIP /
 * MI355X: transfers and the elementwise / reduction operations of
 * include/matrix_operations.h, each one gfx950 kernel launch on the given stream.
 */
#include "matrix_operations.h"

#include <stdio.h>
#include <stdlib.h>

#include "GPU_setup.h"
#include "common.h"
#include "nmfk_internal.h"

static int fail(const char *who, const char *name, hipError_t e)
{
	print_error(true, "Error in %s(%s): %s\n", who, name ? name : "", hipGetErrorString(e));
	return EXIT_FAILURE;
}

/* The job's factors keep columns K..Kp-1 at zero; elementwise operations on them touch
 * only the first K columns. */
static index_t op_width(real const *d_A, index_t pitch)
{
	nmfk_job *j = nmfk_dropin_job;
	if (j && (d_A == j->W || d_A == j->H))
		return K < pitch ? K : pitch;
	return pitch;
}

static void mark_if_V(real const *d_A)
{
	if (nmfk_dropin_job && d_A == nmfk_dropin_job->V)
		nmfk_job_mark_V_changed(nmfk_dropin_job);
	if (nmfk_dropin_job && nmfk_dropin_job->streamed && d_A == nmfk_dropin_job->Vb[0])
		nmfk_dropin_job->slot0_written = 1;
}

index_t gpu_max_height(index_t pitch)
{
	if (!pitch)
		return 0;
	size_t h = gpu_max_num_items / pitch;
	return h > IDX_MAX ? IDX_MAX : (index_t)h;
}

size_t gpu_max_nitems(void)
{
	return gpu_max_num_items;
}

void init_kernel_params(index_t pitch)
{
	threadsPerBlock_pitch = 256;
	maxBlockHeight_pitch = pitch ? threadsPerBlock_pitch / pitch : 0;
}

int show_device_matrix(void const *dMatrix, index_t nrows, index_t ncols, index_t pitch, bool real_data,
		       bool transpose, bool all_processes, struct matrix_tags_t const *mt)
{
	(void)mt;
	(void)all_processes;
	const index_t show_r = nrows < 9 ? nrows : 9, show_c = ncols < 15 ? ncols : 15;
	const size_t items = (size_t)show_r * pitch;
	void *h = malloc(items * 4 + 4);
	if (!h)
		return EXIT_FAILURE;
	hipError_t e = hipMemcpy(h, dMatrix, items * 4, hipMemcpyDeviceToHost);
	if (e != hipSuccess) {
		free(h);
		return fail("show_device_matrix", "", e);
	}
	print_message(false, "%s%" PRI_IDX " x %" PRI_IDX " (pitch %" PRI_IDX ")%s:\n", transpose ? "transposed " : "",
		      nrows, ncols, pitch, (show_r < nrows || show_c < ncols) ? ", first rows/columns" : "");
	for (index_t r = 0; r < show_r; r++) {
		for (index_t c = 0; c < show_c; c++) {
			const size_t p = (size_t)r * pitch + c;
			if (real_data)
				append_printed_message(false, " %g", (double)((real *)h)[p]);
			else
				append_printed_message(false, " %" PRI_IDX, ((index_t *)h)[p]);
		}
		append_printed_message(false, "\n");
	}
	free(h);
	return EXIT_SUCCESS;
}

int matrix_random(real *d_A, index_t height, index_t width, index_t padding, char const *matrix_name,
		  hipStream_t stream_A, hipEvent_t *event_A)
{
	if (!nmfk_dropin_rng_ready) {
		print_error(true, "Error in matrix_random(%s): random generator not initialized.\n",
			    matrix_name ? matrix_name : "");
		return EXIT_FAILURE;
	}
	if (width > padding) {
		print_error(true, "Error in matrix_random(%s): width > padding.\n", matrix_name ? matrix_name : "");
		return EXIT_FAILURE;
	}
	/* successive calls draw successive streams of the seeded generator (W first, then H) */
	const uint32_t tag = nmfk_dropin_rng_calls++;
	int e = nmfk_launch_uniform(d_A, height, width, padding, nmfk_dropin_seed, tag, stream_A);
	if (!e && event_A)
		e = hipEventRecord(*event_A, stream_A);
	return e ? fail("matrix_random", matrix_name, (hipError_t)e) : EXIT_SUCCESS;
}

int matrix_to_row(real const *d_A, index_t height, index_t pitch, char const *matrix_name, real *d_Tmp,
		  real *d_accum_A, hipStream_t stream_AccA)
{
	const size_t tmp_items = (d_Tmp && height > 2) ? (size_t)(height / 4) * pitch : 0;
	int e = nmfk_launch_colsum_any(d_A, height, pitch, pitch, d_Tmp, tmp_items, d_accum_A, stream_AccA);
	if (!e && event_reduction)
		e = hipEventRecord(event_reduction, stream_AccA);
	return e ? fail("matrix_to_row", matrix_name, (hipError_t)e) : EXIT_SUCCESS;
}

int matrix_div_sub(real *d_A, real const *d_B, index_t height, index_t pitch, char const *matrix_name_A,
		   char const *matrix_name_B, bool div_operator, hipStream_t stream_A, hipEvent_t event_B)
{
	(void)matrix_name_B;
	int e = event_B ? hipStreamWaitEvent(stream_A, event_B, 0) : 0;
	if (!e)
		e = nmfk_launch_div_sub(d_A, d_B, (size_t)height * pitch, div_operator, stream_A);
	if (!e && event_B)
		e = hipEventRecord(event_B, stream_A);
	return e ? fail("matrix_div_sub", matrix_name_A, (hipError_t)e) : EXIT_SUCCESS;
}

int matrix_mul_div(real *d_A, real const *d_Aux, real const *d_accum_B, index_t height, index_t pitch,
		   char const *matrix_name_A, hipStream_t stream_A)
{
	int e = event_reduction ? hipStreamWaitEvent(stream_A, event_reduction, 0) : 0;
	if (!e)
		e = nmfk_launch_mul_div(d_A, d_Aux, d_accum_B, height, op_width(d_A, pitch), pitch, stream_A);
	return e ? fail("matrix_mul_div", matrix_name_A, (hipError_t)e) : EXIT_SUCCESS;
}

int matrix_adjust(real *d_A, index_t height, index_t pitch, char const *matrix_name_A, hipStream_t stream_A,
		  hipEvent_t *event_A)
{
	int e = event_A ? hipStreamWaitEvent(stream_A, *event_A, 0) : 0;
	if (!e)
		e = nmfk_launch_adjust(d_A, height, op_width(d_A, pitch), pitch, stream_A);
	if (!e && event_A)
		e = hipEventRecord(*event_A, stream_A);
	if (!e && nmfk_dropin_job && d_A == nmfk_dropin_job->W)
		nmfk_dropin_job->colW_valid = 0;
	return e ? fail("matrix_adjust", matrix_name_A, (hipError_t)e) : EXIT_SUCCESS;
}

int matrix_idx_max(real const *d_A, index_t width, index_t pitch, index_t height, char const *matrix_name_A,
		   hipStream_t stream_A, index_t *d_Idx)
{
	int e = nmfk_launch_idx_max(d_A, height, width, pitch, d_Idx, stream_A);
	return e ? fail("matrix_idx_max", matrix_name_A, (hipError_t)e) : EXIT_SUCCESS;
}

int upload_matrix(real const *A, index_t height, index_t pitch, real *d_A, char const *matrix_name_A,
		  hipStream_t stream_A, hipEvent_t *event_A)
{
	int e = hipMemcpyAsync(d_A, A, (size_t)height * pitch * sizeof(real), hipMemcpyHostToDevice, stream_A);
	if (!e && event_A)
		e = hipEventRecord(*event_A, stream_A);
	mark_if_V(d_A);
	return e ? fail("upload_matrix", matrix_name_A, (hipError_t)e) : EXIT_SUCCESS;
}

int upload_matrix_partial(real const *pA, index_t height, index_t pitch, index_t strow, index_t stcol,
			  char const *matrix_name_A, index_t block_pitch, real *d_A, hipStream_t stream_A,
			  hipEvent_t event_A)
{
	if (stcol >= pitch) {
		print_error(true, "Error in upload_matrix_partial(%s): stcol >= pitch.\n", matrix_name_A ? matrix_name_A : "");
		return EXIT_FAILURE;
	}
	const index_t width = (stcol + block_pitch <= pitch) ? block_pitch : pitch - stcol;
	int e = event_A ? hipStreamWaitEvent(stream_A, event_A, 0) : 0;
	if (!e)
		e = hipMemcpy2DAsync(d_A, (size_t)block_pitch * sizeof(real), pA + (size_t)strow * pitch + stcol,
				     (size_t)pitch * sizeof(real), (size_t)width * sizeof(real), height,
				     hipMemcpyHostToDevice, stream_A);
	if (!e && event_A)
		e = hipEventRecord(event_A, stream_A);
	mark_if_V(d_A);
	return e ? fail("upload_matrix_partial", matrix_name_A, (hipError_t)e) : EXIT_SUCCESS;
}

int download_matrix(void *A, size_t nitems, size_t data_size, void const *d_A, char const *matrix_name_dA,
		    hipStream_t stream_A)
{
	int e = hipMemcpyAsync(A, d_A, nitems * data_size, hipMemcpyDeviceToHost, stream_A);
	if (!e)
		e = hipStreamSynchronize(stream_A);
	return e ? fail("download_matrix", matrix_name_dA, (hipError_t)e) : EXIT_SUCCESS;
}
