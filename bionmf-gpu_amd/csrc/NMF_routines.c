/*
 * NMF_routines.c -- drop-in for the reference's src/NMF_routines.c on HIP / MI355X.
 *
 * update_H / update_W keep the reference's definition of a half-step
 * (src/NMF_routines.c:704-1000, :1344-1636): column sums of the other factor, then
 * A = A .* (contraction of V ./ (W H^T)) ./ column sums, on the factorization that
 * setup_GPU() created.  The work is two streaming kernels per half-step instead of the
 * reference's GEMM / div / GEMM / mul-div over a materialised n x m WH.
 */
#include "NMF_routines.h"

#include <stdlib.h>

#include "GPU_setup.h"
#include "common.h"
#include "matrix_operations.h"
#include "nmfk_internal.h"

index_t pBLN = 0, pBLM = 0;
int stepN = 1, stepM = 1;
index_t psNMF_N = 0, psNMF_M = 0;
index_t colIdx = 0, rowIdx = 0;

static int need_job(const char *who)
{
	if (!nmfk_dropin_job) {
		print_error(true, "Error: %s() called before setup_GPU().\n", who);
		return 1;
	}
	return 0;
}

int init_random(index_t seed)
{
	return init_GPU_random((unsigned long long)seed);
}

int destroy_random(void)
{
	return finalize_randomGenerator();
}

int set_random_values(real *d_A, index_t height, index_t width, index_t padding, char const *matrix_name_dA,
		      hipStream_t stream_A, hipEvent_t *event_A)
{
	return matrix_random(d_A, height, width, padding, matrix_name_dA, stream_A, event_A);
}

int update_H(void)
{
	if (need_job("update_H"))
		return EXIT_FAILURE;
	if (nmfk_job_update_H(nmfk_dropin_job)) {
		print_error(true, "Error in update_H(): %s\n", nmfk_last_error());
		return EXIT_FAILURE;
	}
	if (hipEventRecord(event_H, stream_H) != hipSuccess)
		return EXIT_FAILURE;
	return EXIT_SUCCESS;
}

int update_W(void)
{
	if (need_job("update_W"))
		return EXIT_FAILURE;
	if (nmfk_job_update_W(nmfk_dropin_job)) {
		print_error(true, "Error in update_W(): %s\n", nmfk_last_error());
		return EXIT_FAILURE;
	}
	if (hipEventRecord(event_W, stream_W) != hipSuccess)
		return EXIT_FAILURE;
	return EXIT_SUCCESS;
}

int get_classification(index_t *ld_classification, index_t *lh_classification)
{
	if (need_job("get_classification"))
		return EXIT_FAILURE;
	if (matrix_idx_max(d_H, K, Kp, M, "d_H", stream_H, ld_classification) != EXIT_SUCCESS)
		return EXIT_FAILURE;
	if (download_matrix(lh_classification, M, sizeof(index_t), ld_classification, "d_classification", stream_H) !=
	    EXIT_SUCCESS)
		return EXIT_FAILURE;
	return EXIT_SUCCESS;
}

int dot_product_VWH(real *dot_V, real *dot_VWH)
{
	if (need_job("dot_product_VWH"))
		return EXIT_FAILURE;
	double dv = 0.0, dvwh = 0.0;
	if (nmfk_job_refresh_Vt(nmfk_dropin_job) || nmfk_job_dot_kl(nmfk_dropin_job, &dv, &dvwh, NULL)) {
		print_error(true, "Error in dot_product_VWH(): %s\n", nmfk_last_error());
		return EXIT_FAILURE;
	}
	*dot_V = (real)dv;
	*dot_VWH = (real)dvwh;
	return EXIT_SUCCESS;
}
