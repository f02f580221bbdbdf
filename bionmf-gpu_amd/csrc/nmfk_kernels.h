/*
 * nmfk_kernels.h -- C prototypes of the gfx950 kernel launchers (nmfk_kernels.hip).
 *
 * Every launcher is asynchronous on <stream>, takes plain device pointers and sizes,
 * and returns 0 or a hipError_t value.  Matrices are row-major FP32:
 *   V  n x ldv   (ldv >= m, ldv % 4 == 0)        Vt  m x ldvt (ldvt >= n, ldvt % 4 == 0)
 *   W  n x kt    H  m x kt   (H stored transposed, as the reference stores it)
 * kt is the job's padded rank (nmfk_job_kt(n, m, k)); columns k..kt-1 of W and H stay zero.
 */
#ifndef NMFK_KERNELS_H
#define NMFK_KERNELS_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Device-side convergence state of one factorization (src/NMF_GPU.c:522-711). */
typedef struct nmfk_state {
	uint32_t inc;         /* unchanged tests in a row */
	uint32_t converged;   /* 1: the stop rule was met; 2: test budget spent without converging;
	                         nonzero makes every later launch of the job a no-op */
	uint32_t ntests;      /* tests performed */
	uint32_t stop_test;   /* 1-based test index that converged (0 = none) */
	uint64_t last_diff;   /* connectivity difference of the most recent test */
	uint32_t stop_threshold;
	uint32_t max_tests;   /* tests after which an unconverged job is done (UINT32_MAX: no limit) */
} nmfk_state;

/* Largest rank handled by the kernels. */
#define NMFK_KMAX 64

/* Padded rank of a rank alone (4 or 8 up to k = 8, then 16, 32, 64). */
uint32_t nmfk_kt(uint32_t k);
/* Padded rank a job of n x m at rank k uses for storage and kernel selection: nmfk_kt(k),
 * except that ranks up to 8 pad to 16 (FP32 matrix cores) unless n or m is below 32. */
uint32_t nmfk_job_kt(uint32_t n, uint32_t m, uint32_t k);

/* Rows of A whose column sums one partial row covers (k_finish / k_colsum_parts). */
#define NMFK_FIN_ROWS 64

/* Kernel paths of the partial kernel. */
#define NMFK_PATH_AUTO 0     /* VALU for kt <= 8, FP32 MFMA otherwise */
#define NMFK_PATH_VALU 1     /* FP32 FMA on the vector ALUs */
#define NMFK_PATH_MFMA32 2   /* FP32 MFMA 16x16x4 (kt % 16 == 0) */
#define NMFK_PATH_MFMA_BF16 3 /* BF16 MFMA 16x16x32, FP32 accumulation (kt % 32 == 0); automatic for kt 64 */

/* Launch plan of one half-step: grid of the partial kernel. */
typedef struct nmfk_hplan {
	uint32_t nab;     /* a-blocks (grid.x) */
	uint32_t S;       /* b-splits (grid.y) */
	uint32_t bchunk;  /* rows of U per b-split */
	uint32_t ap;      /* a's per lane */
	uint32_t nw;      /* waves per workgroup */
	uint32_t path;    /* NMFK_PATH_* actually used */
	uint32_t lb;      /* VALU: lane groups of a wave taking different rows of U (1, 2, 4, 8) */
	uint32_t ke;      /* FP32 MFMA at kt 16: columns of the rank GEMM1 sums (4, 8 or 16; the
			     planner gives kt, nmfk_plan_set_rank narrows it to the job's rank) */
} nmfk_hplan;

/* Launch-shape tunables of the planner (process-wide; jobs planned afterwards use them). */
#define NMFK_TUNE_WGS_VALU 0     /* workgroups the VALU partial kernel aims for */
#define NMFK_TUNE_WGS_MFMA 1     /* workgroups the MFMA partial kernel aims for */
#define NMFK_TUNE_MIN_STEPS 2    /* minimum row steps per wave before splitting further */
#define NMFK_TUNE_MAX_SPLITS 3   /* most splits of the contraction range (partials to reduce) */
#define NMFK_TUNE_MFMA_NA 4      /* MFMA workgroup: waves side by side along the updated factor */
#define NMFK_TUNE_MFMA_NB 5      /* MFMA workgroup: wave groups splitting the rows of U */
#define NMFK_TUNE_WGS_BF16 6     /* workgroups the BF16 partial kernel aims for */
#define NMFK_TUNE_FUSED_WAVES 7  /* waves the one-pass kernel aims for (workgroups x waves each) */
#define NMFK_TUNE_COUNT 8
int nmfk_tune_set(int key, int value);
int nmfk_tune_get(int key);

/* Chooses the partial-kernel grid for updating a p x kt factor against q rows of U. */
void nmfk_plan_half(uint32_t p, uint32_t q, uint32_t kt, nmfk_hplan *pl);
/* Narrows a plan's GEMM1 width to rank k (FP32 matrix cores at kt 16 with k <= 8). */
void nmfk_plan_set_rank(nmfk_hplan *pl, uint32_t kt, uint32_t k);
void nmfk_plan_half_path(uint32_t p, uint32_t q, uint32_t kt, int path, nmfk_hplan *pl);
/* Same, aiming for <target> workgroups instead of the process tunable (0 = tunable). */
void nmfk_plan_half_target(uint32_t p, uint32_t q, uint32_t kt, int path, uint32_t target, nmfk_hplan *pl);

/*
 * Half-step, first kernel: part[s][a][c] = sum over b in split s of
 *   B[b][c] * U[b][a] / (sum_c' A[a][c'] * B[b][c'])
 * H half: U = V (q = n, p = m), A = H, B = W.   W half: U = Vt (q = m, p = n), A = W, B = H.
 * Workgroup (0,0) also writes den[c] = sum_g colB[g][c] (the column sums of B).
 */
int nmfk_launch_partial(const float *U, size_t ldu, const float *A, const float *B,
			uint32_t p, uint32_t q, uint32_t kt, const nmfk_hplan *pl,
			float *part, const float *colB, uint32_t ncolB, float *den,
			const nmfk_state *st, hipStream_t stream);

/*
 * Half-step, second kernel:
 *   A[a][c] = (A[a][c] * sum_s part[s][a][c]) / den[c]     for c < k
 * then optionally the R_MIN clamp, then colA[blk][c] = column sums of the new A over
 * each block of NMFK_FIN_ROWS rows (the next half-step's denominators).
 */
/* One pass over V per iteration (k_fused_mfma): eligible shapes and their launch plan.
 * Returns 0 and the workgroups G and waves per workgroup when the pass can run the job
 * (kt 16, m <= 1024), else nonzero. */
int nmfk_fused_plan(uint32_t n, uint32_t m, uint32_t kt, uint32_t *G, uint32_t *nwv);
/* The pass: new W rows (upd_w) from W and H, then the next H half-step's sums, one slice
 * of part (G x m x kt) and one row of colW (G x kt) per workgroup; clamp on test passes. */
int nmfk_launch_fused(const float *V, size_t ldv, float *W, const float *H, uint32_t n, uint32_t m, uint32_t k,
		      uint32_t kt, uint32_t G, uint32_t nwv, float *part, float *colW, int upd_w, int clamp,
		      const nmfk_state *st, hipStream_t stream);
/* H = H * (sum of the G slices) / (sum of the G colW rows), columns < k. */
int nmfk_launch_fused_finish(const float *part, uint32_t G, const float *colW, float *H, uint32_t m, uint32_t k,
			     uint32_t kt, const nmfk_state *st, hipStream_t stream);
int nmfk_launch_finish(const float *part, uint32_t S, float *A, uint32_t p, uint32_t k, uint32_t kt,
		       const float *den, float *colA, int clamp, const nmfk_state *st,
		       hipStream_t stream);

/* The same, also writing the BF16 copy of the new A: Ab (p x kt, row-major); kt % 32 == 0
 * when Ab is given. */
int nmfk_launch_finish_shadow(const float *part, uint32_t S, float *A, uint32_t p, uint32_t k, uint32_t kt,
			      const float *den, float *colA, int clamp, const nmfk_state *st, uint16_t *Ab,
			      hipStream_t stream);
/* The BF16 copy of A alone (initial factors). */
int nmfk_launch_shadow(const float *A, uint32_t p, uint32_t kt, uint16_t *Ab, hipStream_t stream);
/* Partial kernel with both contractions in BF16 (plan path NMFK_PATH_MFMA_BF16): B is read
 * from its row-major BF16 copy Bb (q x kt). */
int nmfk_launch_partial_bf16(const float *U, size_t ldu, const float *A, const uint16_t *Bb, uint32_t p, uint32_t q,
			     uint32_t kt, const nmfk_hplan *pl, float *part, const float *colB, uint32_t ncolB,
			     float *den, const nmfk_state *st, hipStream_t stream);

/* colA[blk][c] = sum of A[a][c] over the rows of block blk (same order as k_finish). */
int nmfk_launch_colsum_parts(const float *A, uint32_t p, uint32_t kt, float *colA, hipStream_t stream);

/* den[c] = sum_g parts[g][c] in the fixed order the partial kernel uses. */
int nmfk_launch_den(const float *parts, uint32_t nparts, uint32_t kt, float *den, hipStream_t stream);

/*
 * Convergence test (src/NMF_GPU.c:611-708): clamps H (m x k, pitch kt) in place, writes
 * cls[j] = argmax with the reference tie rule, counts the connectivity difference against
 * last[] exactly (contingency table), copies cls into last and applies the stop rule.
 */
int nmfk_launch_test(float *H, uint32_t m, uint32_t k, uint32_t kt, uint32_t *cls, uint32_t *last,
		     nmfk_state *st, hipStream_t stream);

/* Resets a state block and zeroes last[] (the reference's initial all-zeros vector). */
int nmfk_launch_state_reset(nmfk_state *st, uint32_t stop_threshold, uint32_t max_tests, uint32_t *last,
			    uint32_t m, hipStream_t stream);

/*
 * Batched launches: one grid advances many jobs.  Each job is described on the device by an
 * nmfk_bjob; a launch walks a table of tiles (job, bx, by), each tile being exactly the
 * workgroup (bx, by) of that job's own single-job grid, so a job computes the same bits
 * alone and in a batch.  Workgroups loop over tiles t = blockIdx.x, + gridDim.x, ...
 * (tiles per workgroup = ceil(ntiles / grid)); tiles of a job whose state is nonzero
 * return at once.
 */
typedef struct nmfk_bjob {
	const float *V, *Vt;
	float *W, *H, *part;
	float *colW, *colH, *denW, *denH;
	nmfk_state *st;
	uint32_t *cls, *last;
	uint64_t ldv, ldvt;
	uint32_t n, m, k, kt;
	uint32_t ncolW, ncolH;
	nmfk_hplan planH, planW;
	uint32_t live;        /* 0: empty slot, every tile of it returns without reading the job */
	uint32_t pad_;
} nmfk_bjob;

typedef struct nmfk_tile {
	uint32_t job, bx, by, pad_;
} nmfk_tile;

/* Largest padded rank a batched launch handles (its LDS is sized for it). */
#define NMFK_BATCH_KTMAX 32

/* half: 0 = H half (U = V, A = H, B = W), 1 = W half (U = Vt, A = W, B = H).  cls: the body
 * class of every tile's half-step: 0 = vector ALUs (kt 4/8) without lane groups, 1 = vector
 * ALUs with lane groups (plan lb > 1), 2 and 3 = matrix cores at kt 16 and 32 in 1 x 4 waves,
 * 4 and 5 = the same in 2 x 2 waves (plan ap 2); tiles of other classes are skipped. */
int nmfk_launch_batch_partial(const nmfk_bjob *jobs, const nmfk_tile *tiles, uint32_t ntiles, int half,
			      int cls, uint32_t grid, hipStream_t stream);
int nmfk_launch_batch_finish(const nmfk_bjob *jobs, const nmfk_tile *tiles, uint32_t ntiles, int half,
			     int clamp, uint32_t grid, hipStream_t stream);
/* Test: rows tiles (job, row block of NMFK_TEST_ROWS), then one final workgroup per job. */
#define NMFK_TEST_ROWS 32
int nmfk_launch_batch_test(const nmfk_bjob *jobs, const nmfk_tile *row_tiles, uint32_t nrow_tiles,
			   uint32_t njobs, uint32_t grid, hipStream_t stream);

/* out[t] = *jobs[t].st (zeros for empty slots), t < njobs. */
int nmfk_launch_batch_gather_states(const nmfk_bjob *jobs, uint32_t njobs, nmfk_state *out, hipStream_t stream);

/* C[i][j] += (cls[i] == cls[j]) for an m x m int32 consensus matrix. */
int nmfk_launch_consensus_add(const uint32_t *cls, uint32_t m, int32_t *C, hipStream_t stream);

/* A[i][j] = uniform (0,1] from Philox4x32-10 (seed, tag, i*width + j); padding columns = 0. */
int nmfk_launch_uniform(float *A, uint32_t height, uint32_t width, uint32_t pitch,
			uint64_t seed, uint32_t tag, hipStream_t stream);

/* Synthetic V generators (benchmark inputs). */
int nmfk_launch_gen_lognormal(float *V, uint32_t n, uint32_t m, uint32_t ldv, uint64_t seed,
			      double mu, double sigma, hipStream_t stream);
int nmfk_launch_gen_planted(float *V, uint32_t n, uint32_t m, uint32_t ldv, uint32_t kstar,
			    uint64_t seed, double sigma, double *scratch, hipStream_t stream);

/* Vt[j][i] = V[i][j]. */
int nmfk_launch_transpose(const float *V, uint32_t n, uint32_t m, size_t ldv, float *Vt, size_t ldvt,
			  hipStream_t stream);

/* Per-block FP64 partial sums {sum V^2, sum (V-WH)^2, sum V log(V/WH) - V + WH}. */
uint32_t nmfk_dot_kl_blocks(uint32_t n, uint32_t m);
int nmfk_launch_dot_kl(const float *V, size_t ldv, const float *W, const float *H, uint32_t n,
		       uint32_t m, uint32_t k, uint32_t kt, double *parts, hipStream_t stream);

/* Elementwise / row ops of the reference's matrix_operations.h. */
int nmfk_launch_div_sub(float *A, const float *B, size_t nitems, int div_operator, hipStream_t stream);
int nmfk_launch_mul_div(float *A, const float *Aux, const float *acc, uint32_t height, uint32_t width,
			uint32_t pitch, hipStream_t stream);
int nmfk_launch_adjust(float *A, uint32_t height, uint32_t width, uint32_t pitch, hipStream_t stream);
int nmfk_launch_idx_max(const float *A, uint32_t height, uint32_t width, uint32_t pitch, uint32_t *idx,
			hipStream_t stream);
/* acc[c] = sum of A[i][c] over <height> rows, c < width; tmp (tmp_items floats, may be
 * NULL) holds per-split partial rows. */
int nmfk_launch_colsum_any(const float *A, uint32_t height, uint32_t width, uint32_t pitch,
			   float *tmp, size_t tmp_items, float *acc, hipStream_t stream);

/* Streaming probe: kind 0 reads nfloats from src (dst receives nothing unless the data
 * sums to a sentinel), kind 1 copies src to dst; <blocks> workgroups of 256 threads. */
int nmfk_launch_stream(int kind, const float *src, float *dst, size_t nfloats, unsigned blocks, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* NMFK_KERNELS_H */
