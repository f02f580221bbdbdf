/*
 * nmf_oracle.h -- CPU oracle for the KL-divergence multiplicative-update loop of
 * NMF-mGPU (bioinfo-cnb/bionmf-gpu).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the
 * smoke() entry of __graft_entry__.py and the cpu_baseline leg of bench.py may
 * load it.  The product library (libnmfgpu_amd.so) never links or calls it.
 *
 * It restates, in plain C99, the reference update order and control flow:
 *   update_H / update_W           src/NMF_routines.c:376-525,704-1000 (H half)
 *                                 src/NMF_routines.c:1014-1164,1344-1636 (W half)
 *   elementwise ops               src/GPU_kernels.cu:1690 (A = B / A),
 *                                 src/GPU_kernels.cu:2029-2033 (A = (A*Aux)/acc),
 *                                 src/GPU_kernels.cu:2338-2345 (R_MIN clamp)
 *   argmax and its tie rules      src/GPU_kernels.cu:2592-2640,2761-2830,3227,
 *                                 src/matrix_operations.c:190-203
 *   connectivity difference       src/common.c:1195-1209
 *   loop / adjust cadence / stop  src/NMF_GPU.c:522-795
 *   Frobenius distance            src/NMF_routines.c:1758-2794, src/NMF_GPU.c:861-891
 *
 * Every routine exists in an FP64 form (suffix _f64: the accuracy oracle) and
 * an FP32 form (suffix _f32: a scalar C rendering of the reference update
 * sequence, also used as the CPU throughput baseline).
 *
 * Seeded initial values come from Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11).  The reference draws
 * W0/H0 from cuRAND (src/matrix_operations.c:643-664), whose streams are not
 * available here; both this oracle and the GPU library use the generator below
 * so that parity starts from the same W0 and H0.
 */
#ifndef NMF_ORACLE_H
#define NMF_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- generator (identical in the GPU library) ---------------- */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Uniform value in (0,1] for logical element <idx> of stream <tag>. */
float orc_uniform01(uint64_t seed, uint32_t tag, uint64_t idx);

/* A[i*pitch + j] = orc_uniform01(seed, tag, i*width + j), 0 <= j < width;
 * padding columns are set to 0. */
void orc_uniform_fill(float *A, uint32_t height, uint32_t width, uint32_t pitch,
		      uint64_t seed, uint32_t tag);

/* Standard normal for logical element <idx> of stream <tag> (Box-Muller, FP64). */
double orc_normal(uint64_t seed, uint32_t tag, uint64_t idx);

/* V_ij = exp(mu + sigma * z_ij), rounded to float. */
void orc_gen_lognormal(float *V, uint32_t n, uint32_t m, uint32_t pitch,
		       uint64_t seed, double mu, double sigma);

/* Planted low rank: V = (W* . H*) (.) exp(sigma * z), W* n x kstar, H* kstar x m
 * with Gamma(1,1) = Exp(1) entries; all intermediate arithmetic in FP64. */
void orc_gen_planted(float *V, uint32_t n, uint32_t m, uint32_t pitch, uint32_t kstar,
		     uint64_t seed, double sigma);

/* ---------------- integer parts ---------------- */

/* Number of pairs i<j whose co-membership differs (src/common.c:1195-1209). */
size_t orc_get_difference(const uint32_t *cls, const uint32_t *last, uint32_t m);

/* Stop rule of src/NMF_GPU.c:695-708 applied to a sequence of differences.
 * Returns the number of tests performed before convergence (1-based), or 0
 * if the sequence ends without convergence. */
uint32_t orc_stop_rule(const size_t *diffs, uint32_t ntests, uint32_t stop_threshold);

/* ---------------- per-precision routines ---------------- */

#define ORC_DECLARE(REAL, SFX)                                                            \
void orc_colsum_##SFX(const REAL *A, uint32_t height, uint32_t width, uint32_t pitch,      \
		      REAL *acc);                                                          \
void orc_update_H_##SFX(const REAL *V, uint32_t n, uint32_t m, uint32_t ldv,              \
			const REAL *W, REAL *H, uint32_t k, uint32_t ldk);                 \
void orc_update_W_##SFX(const REAL *V, uint32_t n, uint32_t m, uint32_t ldv,              \
			REAL *W, const REAL *H, uint32_t k, uint32_t ldk);                 \
void orc_adjust_##SFX(REAL *A, uint32_t height, uint32_t width, uint32_t pitch);           \
void orc_idx_max_##SFX(const REAL *A, uint32_t height, uint32_t width, uint32_t pitch,     \
		       uint32_t *idx);                                                     \
void orc_mul_div_##SFX(REAL *A, const REAL *Aux, const REAL *acc, uint32_t height,        \
		       uint32_t width, uint32_t pitch);                                    \
void orc_div_sub_##SFX(REAL *A, const REAL *B, uint32_t height, uint32_t width,           \
		       uint32_t pitch, int div_operator);                                  \
void orc_dot_VWH_##SFX(const REAL *V, uint32_t n, uint32_t m, uint32_t ldv,               \
		       const REAL *W, const REAL *H, uint32_t k, uint32_t ldk,             \
		       double *dot_V, double *dot_VWH);                                    \
double orc_kl_div_##SFX(const REAL *V, uint32_t n, uint32_t m, uint32_t ldv,              \
			const REAL *W, const REAL *H, uint32_t k, uint32_t ldk);           \
int orc_nmf_##SFX(const REAL *V, uint32_t n, uint32_t m, uint32_t ldv,                    \
		  REAL *W, REAL *H, uint32_t k, uint32_t ldk,                              \
		  uint32_t nIters, uint32_t niter_test_conv, uint32_t stop_threshold,      \
		  uint32_t *cls_out, uint32_t *iters_out, int *converged_out,              \
		  uint32_t *cls_trace, size_t *diff_trace, uint32_t max_trace,             \
		  uint32_t *ntests_out);

ORC_DECLARE(double, f64)
ORC_DECLARE(float, f32)

#undef ORC_DECLARE

#ifdef __cplusplus
}
#endif

#endif /* NMF_ORACLE_H */
