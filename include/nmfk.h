/*
 * nmfk.h -- job-handle C-ABI of the MI355X KL-NMF library (libnmfgpu_amd.so).
 *
 * The reference keeps one factorization in process globals (include/common.h:222-245,
 * include/GPU_setup.h:236-248) and drives it from nmf() (src/NMF_GPU.c:312-924).  This
 * header exposes the same loop on an explicit handle, so that several factorizations can
 * live on one GPU at once (rank sweeps, the serving loop).  The drop-in headers
 * NMF_routines.h, matrix_operations.h and GPU_setup.h are implemented on top of it.
 *
 * Conventions: plain pointers and sizes; host matrices are row-major FP32 with an explicit
 * leading dimension; H is m x k (transposed, as in the reference).  Functions return 0 on
 * success and a nonzero code on failure, with a message retrievable by nmfk_last_error().
 * A handle is used from one host thread at a time.
 */
#ifndef NMFK_H
#define NMFK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nmfk_job nmfk_job;

/* Options of the loop, as the reference's -k/-i/-j/-t flags (src/common.c:847-1119,
 * defaults include/common_defaults.h:78-96). */
typedef struct nmfk_options {
	uint32_t k;              /* factorization rank, 1..64 */
	uint32_t nIters;         /* maximum iterations (default 2000) */
	uint32_t niter_test_conv;/* iterations between convergence tests (default 10) */
	uint32_t stop_threshold; /* unchanged tests needed, minus one (default 40) */
} nmfk_options;

/* Outcome of nmfk_job_run. */
typedef struct nmfk_result {
	uint32_t iterations;     /* as the reference prints: tests*T when converged, else nIters */
	uint32_t converged;      /* 1 if the stop rule was met */
	uint32_t ntests;         /* convergence tests performed */
	uint32_t pad_;
	uint64_t last_diff;      /* connectivity difference of the last test */
	double seconds;          /* wall time of the iteration loop (device work included) */
} nmfk_result;

/* Fills <opt> with the reference defaults (k=2, 2000, 10, 40). */
void nmfk_default_options(nmfk_options *opt);

/* Last error message of the calling thread ("" if none). */
const char *nmfk_last_error(void);

/* Selects the device used by subsequent calls of this thread (hipSetDevice). */
int nmfk_set_device(int dev);

/* Creates a job for an n x m input and rank k; allocates every device buffer it needs. */
nmfk_job *nmfk_job_create(uint32_t n, uint32_t m, uint32_t k);
void nmfk_job_destroy(nmfk_job *job);

/* Device bytes a job of this shape allocates (used for admission by HBM budget). */
size_t nmfk_job_bytes(uint32_t n, uint32_t m, uint32_t k);

/* Input matrix V: from host memory (row-major, leading dimension ldv >= m), from device
 * memory, or generated on the device (synthetic benchmark inputs).  All values must be
 * non-negative; the transposed copy the W half-step streams is built here. */
int nmfk_job_set_V(nmfk_job *job, const float *V, size_t ldv);
int nmfk_job_set_V_device(nmfk_job *job, const float *dV, size_t ldv);
int nmfk_job_gen_V_lognormal(nmfk_job *job, uint64_t seed, double mu, double sigma);
int nmfk_job_gen_V_planted(nmfk_job *job, uint64_t seed, uint32_t kstar, double sigma);
/* Copies V back to the host (ldv >= m). */
int nmfk_job_get_V(nmfk_job *job, float *V, size_t ldv);

/* Tells the job that its device V (nmfk_job_info.dV) was written by the caller; the
 * transposed copy is rebuilt before the next W half-step. */
int nmfk_job_mark_V_changed(nmfk_job *job);

/* Initial W (stream tag 0) and H (tag 1): uniform (0,1] from Philox4x32-10(seed). */
int nmfk_job_init_random(nmfk_job *job, uint64_t seed);
/* Initial W (n x k, ldw) and H (m x k, ldh) from host memory. */
int nmfk_job_set_WH(nmfk_job *job, const float *W, size_t ldw, const float *H, size_t ldh);
int nmfk_job_get_W(nmfk_job *job, float *W, size_t ldw);
int nmfk_job_get_H(nmfk_job *job, float *H, size_t ldh);

/* The whole loop of src/NMF_GPU.c:522-795 on the device. */
int nmfk_job_run(nmfk_job *job, const nmfk_options *opt, nmfk_result *res);

/* nmfk_job_run, also returning the connectivity difference the device counted at every
 * convergence test (the values src/NMF_GPU.c:677-708 feeds to the stop rule): diffs[t] is
 * that of test t + 1, for t < min(tests performed, cap); *ntrace = tests observed. */
int nmfk_job_run_traced(nmfk_job *job, const nmfk_options *opt, nmfk_result *res, uint64_t *diffs,
			uint32_t cap, uint32_t *ntrace);

/* Single half-steps with the reference's structure (column sum, then the update). */
int nmfk_job_update_H(nmfk_job *job);
int nmfk_job_update_W(nmfk_job *job);
/* R_MIN clamp of H and W (src/NMF_GPU.c:611-638). */
int nmfk_job_adjust(nmfk_job *job);

/* Class vector: argmax of every row of the current H with the reference tie rule. */
int nmfk_job_get_classification(nmfk_job *job, uint32_t *cls);

/* FP64 sums: <V,V>, <V-WH,V-WH> and the generalized KL divergence D(V||WH). */
int nmfk_job_dot_kl(nmfk_job *job, double *dot_V, double *dot_VWH, double *kl);

/* nmfk_job_dot_kl on caller-owned device buffers, allocating nothing.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  V is n x ldv
 * (ldv >= m), W n x kt and H m x kt with kt = nmfk_job_kt(n, m, k) and columns k..kt-1
 * zero.  Writes {<V,V>, <V-WH,V-WH>, D(V||WH)} as 3 doubles to device memory <out>,
 * asynchronously on <stream> (a hipStream_t; NULL = the null stream): the same bits
 * nmfk_job_dot_kl returns for the same buffers.  The FP64 partial sums go to a static device array of
 * nmfk_dot_kl_noalloc_max_bytes() bytes (1 MiB) per device, held until the process ends;
 * 24 bytes per 256 columns x 64 row splits, so m <= 174592; larger m is refused.  Calls
 * that overlap on one device share that array and race: serialise them. */
int nmfk_dot_kl_noalloc(const float *dV, size_t ldv, const float *dW, const float *dH, uint32_t n, uint32_t m,
			uint32_t k, double *d_out, void *stream);
size_t nmfk_dot_kl_noalloc_max_bytes(void);

/* Synchronizes the job's stream. */
int nmfk_job_sync(nmfk_job *job);

/* Kernel path of the half-step contractions: 0 = automatic (vector ALUs for k <= 8 when V
 * has fewer than 32 rows or columns, FP32 matrix cores up to k = 32, BF16 matrix cores
 * above), 1 = vector ALUs, 2 = FP32 matrix cores (k padded to 16), 3 = BF16 matrix cores (k
 * padded to 32; FP32 accumulation).  The job's padded rank (nmfk_job_kt) is fixed when it is
 * created, and a path its padding does not fit falls back to the next simpler one. */
int nmfk_job_set_kernel_path(nmfk_job *job, int path);
/* Padded rank of a job of n x m at rank k (the row pitch of W and H): 4 or 8 for k <= 8
 * when n or m is below 32 (the vector-ALU kernel's widths), else 16 up to k = 16, then 32
 * or 64.  Padding small ranks to 16 puts them on the FP32 matrix cores, which run them
 * 1.1-2x faster than the vector ALUs at every width but the narrowest. */
uint32_t nmfk_job_kt(uint32_t n, uint32_t m, uint32_t k);
int nmfk_job_kernel_path(const nmfk_job *job);

/* Iteration schedule of nmfk_job_run.  NMFK_SCHED_TWO_PASS: one pass over V per half-step
 * (V for the H half, its transposed copy for the W half), the schedule batches and the
 * drop-in half-steps use.  NMFK_SCHED_ONE_PASS: one pass over V per iteration, each row
 * block giving its new W rows and then its share of the next H half-step's sums (FP32
 * matrix cores; padded rank 16, i.e. k <= 16 on V of 32 rows and columns or more, and
 * m <= 1024 on the default kernel path); it reads V half as often but does the same
 * matrix-core work, on which both are bound at these ranks.  NMFK_SCHED_AUTO (default): two passes.  Both keep the reference's update order,
 * clamp and stop rule; their sums are added in different orders, so they agree to rounding,
 * not bit for bit. */
#define NMFK_SCHED_AUTO 0
#define NMFK_SCHED_TWO_PASS 1
#define NMFK_SCHED_ONE_PASS 2
int nmfk_job_set_schedule(nmfk_job *job, int sched);
/* The schedule nmfk_job_run will use (NMFK_SCHED_TWO_PASS or NMFK_SCHED_ONE_PASS; a requested
 * one pass reads as two where the job's kernel path no longer allows it). */
int nmfk_job_schedule(const nmfk_job *job);

/* Launch-shape tunables of the kernel planner, process-wide; they apply to jobs created
 * afterwards.  Keys: 0 = workgroups the vector-ALU partial kernel aims for (default 512),
 * 1 = workgroups the FP32 matrix-core partial kernel aims for (default 1024), 2 = minimum
 * row steps per wave before the contraction range is split further (default 8), 3 = most
 * splits of the contraction range (default 64), 4 and 5 = matrix-core workgroup shape:
 * waves side by side along the updated factor x wave groups splitting the contraction
 * range (default 1 x 4; also 2x2, 4x1, 2x4, 4x2), 6 = workgroups the BF16 matrix-core
 * partial kernel aims for (default 256), 7 = waves the one-pass kernel aims for (workgroups x
 * waves per workgroup, default 2048).  With keys 1-5 at their defaults, an FP32 matrix-core
 * job planned on its own (not replanned to a workgroup target) gets 2 x 2 workgroups over a
 * contraction range of 1000-4095 rows where that leaves at least 240 workgroups, a split
 * count balanced over the CUs (at most the default's), and 4 minimum row steps instead of 8
 * where 8 leave fewer than 240 workgroups; changing any of them turns all three off. */
int nmfk_tune_set(int key, int value);
int nmfk_tune_get(int key);

/* A job of rank k on the V of <src>, without copying it: src must outlive the job, and
 * its V must not change while the job uses it (nmfk_job_set_V* on the new job fail). */
nmfk_job *nmfk_job_create_shared(nmfk_job *src, uint32_t k);

/* A job whose V stays in host memory: the out-of-core counterpart of the reference's
 * blockwise mode for inputs that do not fit the device (src/GPU_setup.c:902-1270, block
 * loops of src/NMF_routines.c:543-1000).  V (n x m, row-major, leading dimension ldv) is
 * copied to the device in row blocks of <block_rows> rows (a multiple of 64; 0 = as many
 * as fit in 1 GiB of slots) through two slots, once per iteration: each block serves the W
 * half-step of its rows and then its share of the next H half-step's sums.  pinned != 0
 * declares that V lies in memory from nmfk_host_alloc (copied directly); otherwise it is
 * staged through pinned buffers.  V is read, never written, and must outlive the job.
 * Runs follow the reference's update order and stop rule like nmfk_job_run, in FP32 (kernel
 * paths VALU or MFMA32); W, H, the class vector and nmfk_job_dot_kl work as for any job,
 * and so do nmfk_job_update_H / nmfk_job_update_W (one pass over V each, as the reference's
 * blockwise half-steps), while the V accessors, batching and timing refuse such a job. */
nmfk_job *nmfk_job_create_streamed(uint32_t n, uint32_t m, uint32_t k, const float *V, size_t ldv,
				   uint32_t block_rows, int pinned);
/* Device bytes nmfk_job_create_streamed allocates for these arguments (0: invalid). */
size_t nmfk_streamed_bytes(uint32_t n, uint32_t m, uint32_t k, uint32_t block_rows);
/* Rows per block, number of blocks, and bytes of V copied to the device by the last run. */
int nmfk_job_streamed_info(const nmfk_job *job, uint32_t *block_rows, uint32_t *nblocks, double *bytes_copied);

/* Pinned host memory for a streamed job's V (hipHostMalloc), and its release. */
void *nmfk_host_alloc(size_t bytes);
void nmfk_host_free(void *p);

/* Re-plans both half-steps aiming for <target_wgs> partial-kernel workgroups (0 = the
 * process tunables); the number of splits of the contraction range, and with it the
 * summation order, follows the plan, so jobs compared bit for bit must share plans. */
int nmfk_job_replan(nmfk_job *job, uint32_t target_wgs);
/* {nab, S, bchunk, path} of the H half, then of the W half. */
int nmfk_job_plan_get(const nmfk_job *job, uint32_t out[8]);

/* ---------------------------------------------------------------------------------------
 * Batched execution (continuous batching).  A batch advances up to <capacity> jobs of
 * uneven shape (n, m, k <= 32 on the default kernel paths) with one launch per phase of
 * the iteration; each job's results are bitwise those of nmfk_job_run on it.  Jobs join
 * with nmfk_batch_add (W and H already initialised) and leave through nmfk_batch_poll
 * once the stop rule is met or nIters is reached; after poll returns a job, the batch no
 * longer touches it.  A job must not be destroyed while it is in a batch (added and not yet
 * returned by poll), unless the batch itself has been destroyed first.  All jobs of a batch
 * share niter_test_conv.  Not thread-safe.
 * ------------------------------------------------------------------------------------- */
typedef struct nmfk_batch nmfk_batch;

/* Launch tuning of a batch (nmfk_batch_autotune picks the first two fields). */
typedef struct nmfk_batch_tuning {
	uint32_t grid_partial;    /* workgroups of the partial launches (0: one per tile, else each
	                             workgroup runs ceil(tiles / grid) tiles of possibly different jobs) */
	uint32_t grid_finish;     /* same for the finish and test launches */
	uint32_t order;           /* tile order: 0 job by job, 1 split by split across jobs */
	uint32_t chunks_per_poll; /* test intervals enqueued per nmfk_batch_step in nmfk_batch_run-style loops */
	uint32_t compact_pct;     /* rebuild the tile tables when finished jobs hold this % of tiles */
} nmfk_batch_tuning;

void nmfk_batch_default_tuning(nmfk_batch_tuning *t);
nmfk_batch *nmfk_batch_create(uint32_t niter_test_conv, uint32_t capacity);
void nmfk_batch_destroy(nmfk_batch *b);
int nmfk_batch_set_tuning(nmfk_batch *b, const nmfk_batch_tuning *t);
int nmfk_batch_get_tuning(const nmfk_batch *b, nmfk_batch_tuning *t);
/* Adds a job; resets its stop-rule state.  Returns its slot (>= 0) or -1. */
int nmfk_batch_add(nmfk_batch *b, nmfk_job *job, const nmfk_options *opt);
/* Jobs added and not yet returned by nmfk_batch_poll. */
uint32_t nmfk_batch_active(const nmfk_batch *b);
/* Enqueues <chunks> test intervals (niter_test_conv iterations + test each) for every
 * active job; asynchronous. */
int nmfk_batch_step(nmfk_batch *b, uint32_t chunks);
/* Waits for the enqueued chunks, runs the leftover nIters % niter_test_conv iterations of
 * jobs that reached nIters unconverged, and returns (up to <cap>) finished jobs with their
 * results (seconds = time since nmfk_batch_add). */
int nmfk_batch_poll(nmfk_batch *b, nmfk_job **done, nmfk_result *res, uint32_t cap, uint32_t *ndone);
/* Autotuner: times <iters> batched iterations of the current jobs for each candidate
 * launch shape and keeps the fastest (W and H are restored afterwards). */
int nmfk_batch_autotune(nmfk_batch *b, uint32_t iters, nmfk_batch_tuning *best, double *best_ms_per_iter);

/* ---------------------------------------------------------------------------------------
 * Rank sweep x restarts over one V with consensus matrices.
 * ------------------------------------------------------------------------------------- */
#define NMFK_SWEEP_BATCHED 0
#define NMFK_SWEEP_ONE_AT_A_TIME 1

typedef struct nmfk_sweep_config {
	const uint32_t *ks;       /* ranks */
	uint32_t nk;
	uint32_t restarts;        /* restarts per rank; restart r of rank k uses nmfk_sweep_seed(base_seed, k, r) */
	uint64_t base_seed;
	uint32_t nIters, niter_test_conv, stop_threshold;
	uint32_t mode;            /* NMFK_SWEEP_BATCHED or NMFK_SWEEP_ONE_AT_A_TIME */
	uint32_t plan_target;     /* partial-kernel workgroups each job's plan aims for (0: process tunables,
	                             or, with autotune, chosen by the plan search) */
	uint32_t autotune;        /* batched mode: tune the launch shape first (and, when plan_target is 0,
	                             the per-job plan among 8/16/32/64 workgroups and the tunables) */
	uint32_t autotune_iters;
} nmfk_sweep_config;

typedef struct nmfk_sweep_out {
	nmfk_result *res;         /* nk * restarts (may be NULL), index q * restarts + r */
	uint32_t *cls;            /* nk * restarts * m final class vectors (may be NULL) */
	int32_t *consensus;       /* nk * m * m consensus matrices (may be NULL) */
	float *W_H_out;           /* per job W (n x k) then H (m x k), jobs in index order (may be NULL) */
	double seconds;           /* wall time of the factorizations (job set-up and autotuning excluded) */
	double tune_seconds;
	nmfk_batch_tuning tuning; /* launch tuning used (batched mode) */
	uint32_t plan_target;     /* per-job plan target used */
} nmfk_sweep_out;

void nmfk_sweep_default_config(nmfk_sweep_config *c);
/* Seed of restart r at rank k: splitmix64 finaliser of base ^ (k << 32 | r). */
uint64_t nmfk_sweep_seed(uint64_t base, uint32_t k, uint32_t restart);
int nmfk_sweep_run(nmfk_job *src, const nmfk_sweep_config *cfg, nmfk_sweep_out *out);

/* ---------------------------------------------------------------------------------------
 * Serving loop: requests arrive on a wall-clock schedule; each is an input matrix
 * V ~ lognormal(mu, sigma^2) generated on the device from its seed and factorized at rank
 * k from <restarts> starting points (restart r: nmfk_sweep_seed(seed, k, r)).  Requests
 * are admitted in arrival order once their memory fits in the HBM budget and the batch
 * has slots, advance together with every other admitted request, and retire when all
 * their restarts have stopped.
 * ------------------------------------------------------------------------------------- */
typedef struct nmfk_request {
	uint32_t n, m, k, restarts;   /* 1 <= k <= min(n, m, 32) */
	uint64_t seed;
	double arrival;               /* seconds after the start of nmfk_serve_run, non-decreasing */
} nmfk_request;

typedef struct nmfk_request_result {
	double admitted, finished;    /* seconds after the start */
	double latency;               /* finished - arrival */
	uint32_t iterations_max;      /* over restarts */
	uint32_t iterations_sum;
	uint32_t converged;           /* restarts that met the stop rule */
	uint32_t best_restart;        /* restart with the smallest KL divergence */
	double kl_best;               /* its generalized KL divergence D(V || WH), FP64 */
} nmfk_request_result;

typedef struct nmfk_serve_config {
	size_t hbm_budget;            /* bytes of device memory requests may hold at once */
	uint32_t capacity;            /* jobs (restarts) advanced at once */
	uint32_t nIters, niter_test_conv, stop_threshold;
	uint32_t chunks_per_poll;
	uint32_t plan_target;         /* partial-kernel workgroups each job's plan aims for (0: the
	                                 single-job tunables); many jobs share the device, so fewer
	                                 splits per job than alone (default 64) */
	uint32_t progress_every;      /* print a progress line to stderr every N retired requests (0: never) */
	double mu, sigma;             /* V = exp(N(mu, sigma^2)) */
} nmfk_serve_config;

typedef struct nmfk_serve_stats {
	double seconds;               /* start to last retirement */
	uint32_t requests, peak_jobs;
	double requests_per_s, factorizations_per_s;
	double p50, p99, mean_latency;/* request latency, seconds */
	size_t peak_bytes;            /* largest admitted total of nmfk_request_bytes */
	/* Where the serving thread's time went, seconds: enqueueing chunks (including tile-table
	 * rebuilds and graph re-captures when jobs joined or left), admitting requests (job set-up
	 * and input generation), waiting for the device in nmfk_batch_poll, and retiring finished
	 * restarts (their FP64 KL pass and copies). */
	double step_s, admit_s, wait_s, retire_s;
	uint32_t chunks;              /* nmfk_batch_step calls that enqueued work */
	uint32_t pad_;
} nmfk_serve_stats;

void nmfk_serve_default_config(nmfk_serve_config *cfg);
/* Device bytes one request holds (its V once, plus every restart's buffers). */
size_t nmfk_request_bytes(uint32_t n, uint32_t m, uint32_t k, uint32_t restarts);
int nmfk_serve_run(const nmfk_serve_config *cfg, const nmfk_request *req, uint32_t nreq,
		   nmfk_request_result *out, nmfk_serve_stats *stats);

/* Outcome of one restart of a served request (nmfk_serve_run_detail). */
typedef struct nmfk_restart_result {
	uint32_t iterations, converged, ntests, pad_;   /* as nmfk_result */
	uint64_t last_diff;
	double kl;                    /* D(V || WH) of the final factors, FP64 */
	uint32_t *cls;                /* caller's buffer of m entries for the final class vector, or NULL */
	float *W, *H;                 /* caller's buffers (n x k, m x k) for the final factors, or NULL */
} nmfk_restart_result;

/* nmfk_serve_run that also fills one record per restart: request i's restart r is
 * restarts[sum of the restarts of requests 0..i-1 + r].  The cls, W and H pointers of each
 * record are set by the caller (NULL: not copied). */
int nmfk_serve_run_detail(const nmfk_serve_config *cfg, const nmfk_request *req, uint32_t nreq,
			  nmfk_request_result *out, nmfk_serve_stats *stats, nmfk_restart_result *restarts);

/* Raw device pointers and shape (for the drop-in layer and for measurements). */
typedef struct nmfk_job_info {
	uint32_t n, m, k, kt;
	size_t ldv, ldvt;
	float *dV, *dVt, *dW, *dH;
	void *stream;            /* hipStream_t */
} nmfk_job_info;
int nmfk_job_info_get(const nmfk_job *job, nmfk_job_info *info);

/* Times <iters> full iterations (both half-steps, no tests) with HIP events on the job's
 * stream; returns the average ms per iteration and per kernel kind. */
typedef struct nmfk_timing {
	double ms_per_iter;
	double ms_partial_H, ms_finish_H, ms_partial_W, ms_finish_W;
} nmfk_timing;
int nmfk_job_time_iterations(nmfk_job *job, uint32_t iters, nmfk_timing *t);

/* Times <iters> iterations of the one-pass schedule (finish of H, then the pass over V; no
 * tests) with HIP events on the job's stream, after one untimed pass; W and H advance. */
typedef struct nmfk_timing_one_pass {
	double ms_per_iter;
	double ms_pass;          /* the pass over V (new W rows and the next H half's sums) */
	double ms_finish;        /* the H update from the pass's sums */
	uint32_t workgroups, waves_per_workgroup;
} nmfk_timing_one_pass;
int nmfk_job_time_one_pass(nmfk_job *job, uint32_t iters, nmfk_timing_one_pass *t);

/* Measured streaming ceiling of the current device: a plain float4 read kernel and a
 * copy kernel over a fresh <bytes> buffer, <iters> timed launches of <blocks> workgroups
 * each.  Bandwidths count bytes read (read) and read + written (copy). */
int nmfk_measure_stream(size_t bytes, uint32_t iters, uint32_t blocks, double *read_GBps, double *copy_GBps);

#ifdef __cplusplus
}
#endif

#endif /* NMFK_H */
