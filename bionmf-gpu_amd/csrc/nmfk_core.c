/*
 * nmfk_core.c -- host side (C99) of one KL-NMF factorization on one MI355X.
 *
 * Replaces the buffer set-up of src/GPU_setup.c:1289-1610 and the loop of
 * src/NMF_GPU.c:312-924 with a job handle.  One iteration is four launches on the job's
 * stream (H half: partial + finish, W half: partial + finish); every niter_test_conv
 * iterations the test kernel clamps, classifies, counts the connectivity difference and
 * applies the stop rule on the device.  A whole test interval is captured once as a
 * hipGraph and replayed; the host reads the 32-byte state of interval c while interval
 * c+1 is already queued, and every kernel returns at once after the state says
 * "converged", so at most one no-op interval is ever queued past convergence.
 */
#include "nmfk_internal.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static __thread char g_err[512];

void nmfk_set_error(const char *fmt, ...)
{
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof g_err, fmt, ap);
	va_end(ap);
}

const char *nmfk_last_error(void)
{
	return g_err;
}

int nmfk_hip_fail(int err, const char *what)
{
	if (err == 0)
		return 0;
	nmfk_set_error("%s: %s", what, hipGetErrorString((hipError_t)err));
	return NMFK_EHIP;
}

double nmfk_now(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void nmfk_default_options(nmfk_options *opt)
{
	opt->k = 2;
	opt->nIters = 2000;
	opt->niter_test_conv = 10;
	opt->stop_threshold = 40;
}

int nmfk_set_device(int dev)
{
	return nmfk_hip_fail(hipSetDevice(dev), "hipSetDevice");
}

static size_t round_up(size_t x, size_t a)
{
	return (x + a - 1) / a * a;
}

/* Sizes of every device buffer of a job, shared by create() and bytes(). */
typedef struct job_layout {
	uint32_t kt;
	size_t ldv, ldvt;
	nmfk_hplan planH, planW;
	size_t n_V, n_Vt, n_W, n_H, n_part, n_colW, n_colH, n_den, n_dot;
	size_t n_shadow;               /* uint16 elements of the two BF16 copies (0: none) */
} job_layout;

static void layout_of(uint32_t n, uint32_t m, uint32_t k, job_layout *L)
{
	L->kt = nmfk_job_kt(n, m, k);
	L->ldv = round_up(m, 4);
	/* Vt rows start on 256-byte boundaries, so that each wave's 64-column strip of a row is
	 * two whole cache lines (at 50000 x 2000, k = 64, the BF16 W partial takes 88.5 instead of
	 * 92 us).  V keeps the 4-float padding the drop-in's host matrices share. */
	L->ldvt = round_up(n, 64);
	nmfk_plan_half(m, n, L->kt, &L->planH);
	nmfk_plan_half(n, m, L->kt, &L->planW);
	nmfk_plan_set_rank(&L->planH, L->kt, k);
	nmfk_plan_set_rank(&L->planW, L->kt, k);
	L->n_V = (size_t)n * L->ldv;
	L->n_Vt = (size_t)m * L->ldvt;
	L->n_W = (size_t)n * L->kt;
	L->n_H = (size_t)m * L->kt;
	/* partial buffer: large enough for every kernel path a job may be switched to, planned on
	 * its own or replanned to the path's default workgroup target (which skips the lone-job
	 * rules, and may split more) */
	L->n_part = 0;
	for (int path = NMFK_PATH_VALU; path <= NMFK_PATH_MFMA_BF16; path++) {
		const int key = path == NMFK_PATH_VALU ? NMFK_TUNE_WGS_VALU
				: path == NMFK_PATH_MFMA_BF16 ? NMFK_TUNE_WGS_BF16 : NMFK_TUNE_WGS_MFMA;
		for (int t = 0; t < 2; t++) {
			const uint32_t target = t ? (uint32_t)nmfk_tune_get(key) : 0u;
			nmfk_hplan h, w;
			nmfk_plan_half_target(m, n, L->kt, path, target, &h);
			nmfk_plan_half_target(n, m, L->kt, path, target, &w);
			size_t pH = (size_t)h.S * m * L->kt, pW = (size_t)w.S * n * L->kt;
			if (pH > L->n_part)
				L->n_part = pH;
			if (pW > L->n_part)
				L->n_part = pW;
		}
	}
	L->n_colW = (size_t)((n + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS) * L->kt;
	L->n_colH = (size_t)((m + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS) * L->kt;
	L->n_den = 2 * (size_t)L->kt;
	L->n_shadow = L->kt % 32 == 0 ? ((size_t)n + m) * L->kt : 0;
	L->n_dot = (size_t)nmfk_dot_kl_blocks(n, m) * 3;
}

/* Device bytes of a job's buffers, each rounded as the serving arena places it; with_V = 0
 * for a job sharing another job's V and Vt. */
static size_t layout_bytes(const job_layout *L, uint32_t m, int with_V)
{
	const size_t sz[] = {4 * L->n_W, 4 * L->n_H, 4 * L->n_part, 4 * L->n_colW, 4 * L->n_colH, 4 * L->n_den,
			     8 * L->n_dot, 4 * (size_t)m, 4 * (size_t)m, sizeof(nmfk_state)};
	size_t b = 0;
	for (size_t i = 0; i < sizeof sz / sizeof sz[0]; i++)
		b += nmfk_arena_round(sz[i]);
	if (L->n_shadow)
		b += nmfk_arena_round(2 * L->n_shadow);
	if (with_V)
		b += nmfk_arena_round(4 * L->n_V) + nmfk_arena_round(4 * L->n_Vt);
	return b;
}

size_t nmfk_job_bytes(uint32_t n, uint32_t m, uint32_t k)
{
	job_layout L;
	layout_of(n, m, k, &L);
	return layout_bytes(&L, m, 1);
}

size_t nmfk_job_bytes_shared(uint32_t n, uint32_t m, uint32_t k)
{
	job_layout L;
	layout_of(n, m, k, &L);
	return layout_bytes(&L, m, 0);
}

/* A device buffer of job j: from its arena when it has one, else hipMalloc. */
static int dalloc_job(nmfk_job *j, void **p, size_t bytes, const char *what)
{
	*p = NULL;
	if (bytes == 0)
		bytes = 4;
	if (j->arena) {
		*p = nmfk_arena_alloc(j->arena, bytes);
		if (!*p) {
			nmfk_set_error("%s: %zu bytes do not fit in the device arena", what, bytes);
			return NMFK_EBUDGET;
		}
		return 0;
	}
	return nmfk_hip_fail(hipMalloc(p, bytes), what);
}

static void dfree_job(nmfk_job *j, void *p)
{
	if (!p)
		return;
	if (j->arena)
		nmfk_arena_free(j->arena, p);
	else
		hipFree(p);
}

/* The pinned state slots and events nmfk_job_run uses; arena jobs (the serving loop's, which
 * run batched) get them on first use. */
static int run_resources(nmfk_job *j)
{
	int e = 0;
	if (!j->h_st)
		e = nmfk_hip_fail(hipHostMalloc((void **)&j->h_st, 2 * sizeof(nmfk_state), hipHostMallocDefault),
				  "hipHostMalloc(state)");
	for (int i = 0; i < 2 && !e; i++)
		if (!j->ev[i])
			e = nmfk_hip_fail(hipEventCreateWithFlags(&j->ev[i], hipEventDisableTiming), "hipEventCreate");
	return e;
}

/* Jobs whose device V can be written: not shared (nmfk_job_create_shared), not streamed. */
static int v_writable(const nmfk_job *j)
{
	if (j->streamed) {
		nmfk_set_error("a streamed job's V stays in host memory (nmfk_job_create_streamed)");
		return 0;
	}
	if (!j->owns_V) {
		nmfk_set_error("V of a job created by nmfk_job_create_shared belongs to its source job");
		return 0;
	}
	return 1;
}

/* Operations that need V on the device. */
static int needs_device_V(const nmfk_job *j, const char *what)
{
	if (j->streamed) {
		nmfk_set_error("%s: not available for a streamed job (V stays in host memory)", what);
		return NMFK_EINVAL;
	}
	return 0;
}

static int finish_V(nmfk_job *j)
{
	j->colW_valid = 0;
	j->vt_stale = 0;
	int e = nmfk_launch_transpose(j->V, j->n, j->m, j->ldv, j->Vt, j->ldvt, j->stream);
	if (e)
		return nmfk_hip_fail(e, "transpose V");
	return nmfk_hip_fail(hipStreamSynchronize(j->stream), "set V");
}

int nmfk_job_mark_V_changed(nmfk_job *j)
{
	j->vt_stale = !j->streamed;   /* a streamed job has no device V or Vt */
	return 0;
}

int nmfk_job_refresh_Vt(nmfk_job *j)
{
	if (!j->vt_stale || j->streamed)
		return 0;
	j->vt_stale = 0;
	return nmfk_hip_fail(nmfk_launch_transpose(j->V, j->n, j->m, j->ldv, j->Vt, j->ldvt, j->stream),
			     "transpose V");
}

/* Host <-> device copies of <rows> rows of <width> floats between pitches (in floats).
 * Unpadded rows on both sides are one contiguous transfer; otherwise the rows are repacked
 * on the host, so that a tall matrix of narrow rows is never moved row by row.  Device
 * padding columns are written as zeros and host padding columns are left alone.  Both wait
 * for the copy (the stream is synchronized). */
static int h2d_rows(float *d, size_t dpitch, const float *h, size_t hpitch, size_t width, size_t rows,
		    hipStream_t stream)
{
	if (rows == 0)
		return 0;
	const size_t bytes = rows * dpitch * 4;
	float *tmp = NULL;
	const float *src = h;
	if (hpitch != width || dpitch != width) {
		tmp = (float *)calloc(rows * dpitch, 4);
		if (!tmp) {
			nmfk_set_error("out of host memory");
			return NMFK_ENOMEM;
		}
		for (size_t i = 0; i < rows; i++)
			memcpy(tmp + i * dpitch, h + i * hpitch, width * 4);
		src = tmp;
	}
	int e = hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, stream);
	if (!e)
		e = hipStreamSynchronize(stream);
	free(tmp);
	return nmfk_hip_fail(e, "host-to-device copy");
}

static int d2h_rows(float *h, size_t hpitch, const float *d, size_t dpitch, size_t width, size_t rows,
		    hipStream_t stream)
{
	if (rows == 0)
		return 0;
	const size_t bytes = rows * dpitch * 4;
	float *tmp = NULL;
	float *dst = h;
	if (hpitch != width || dpitch != width) {
		tmp = (float *)malloc(bytes);
		if (!tmp) {
			nmfk_set_error("out of host memory");
			return NMFK_ENOMEM;
		}
		dst = tmp;
	}
	int e = hipMemcpyAsync(dst, d, bytes, hipMemcpyDeviceToHost, stream);
	if (!e)
		e = hipStreamSynchronize(stream);
	if (!e && tmp)
		for (size_t i = 0; i < rows; i++)
			memcpy(h + i * hpitch, tmp + i * dpitch, width * 4);
	free(tmp);
	return nmfk_hip_fail(e, "device-to-host copy");
}

static nmfk_job *job_create(uint32_t n, uint32_t m, uint32_t k, nmfk_job *src, nmfk_arena *arena,
			    hipStream_t stream);

nmfk_job *nmfk_job_create(uint32_t n, uint32_t m, uint32_t k)
{
	return job_create(n, m, k, NULL, NULL, NULL);
}

nmfk_job *nmfk_job_create_in(nmfk_arena *arena, hipStream_t stream, uint32_t n, uint32_t m, uint32_t k,
			     nmfk_job *src)
{
	if (!arena || !stream) {
		nmfk_set_error("nmfk_job_create_in: needs an arena and a stream");
		return NULL;
	}
	if (src && (src->streamed || nmfk_job_refresh_Vt(src) || nmfk_job_sync(src)))
		return NULL;
	return job_create(n, m, k, src, arena, stream);
}

nmfk_job *nmfk_job_create_shared(nmfk_job *src, uint32_t k)
{
	if (!src) {
		nmfk_set_error("nmfk_job_create_shared: no source job");
		return NULL;
	}
	if (src->streamed) {
		nmfk_set_error("nmfk_job_create_shared: the source job's V is in host memory (streamed job)");
		return NULL;
	}
	if (nmfk_job_refresh_Vt(src) || nmfk_job_sync(src))
		return NULL;
	return job_create(src->n, src->m, k, src, NULL, NULL);
}

static nmfk_job *job_create(uint32_t n, uint32_t m, uint32_t k, nmfk_job *src, nmfk_arena *arena,
			    hipStream_t stream)
{
	if (n < 1 || m < 1 || k < 1 || k > NMFK_KMAX) {
		nmfk_set_error("nmfk_job_create: invalid shape n=%u m=%u k=%u (1 <= k <= %d)", n, m, k, NMFK_KMAX);
		return NULL;
	}
	if (k > n || k > m) {
		/* src/NMF_GPU.c:1130 */
		nmfk_set_error("nmfk_job_create: k=%u must be <= min(n=%u, m=%u)", k, n, m);
		return NULL;
	}
	nmfk_job *j = (nmfk_job *)calloc(1, sizeof(nmfk_job));
	if (!j) {
		nmfk_set_error("nmfk_job_create: out of host memory");
		return NULL;
	}
	job_layout L;
	layout_of(n, m, k, &L);
	j->arena = arena;
	j->n = n, j->m = m, j->k = k, j->kt = L.kt;
	j->ldv = L.ldv, j->ldvt = L.ldvt;
	j->planH = L.planH, j->planW = L.planW;
	j->part_cap = L.n_part;
	j->ncolW = (n + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS;
	j->ncolH = (m + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS;
	j->ndot = nmfk_dot_kl_blocks(n, m);
	int e = 0;
	j->owns_V = src == NULL;
	if (src) {
		j->V = src->V;
		j->Vt = src->Vt;
	} else {
		e |= dalloc_job(j, (void **)&j->V, 4 * L.n_V, "hipMalloc(V)");
		if (!e) e |= dalloc_job(j, (void **)&j->Vt, 4 * L.n_Vt, "hipMalloc(Vt)");
	}
	if (!e) e |= dalloc_job(j, (void **)&j->W, 4 * L.n_W, "hipMalloc(W)");
	if (!e) e |= dalloc_job(j, (void **)&j->H, 4 * L.n_H, "hipMalloc(H)");
	if (!e) e |= dalloc_job(j, (void **)&j->part, 4 * L.n_part, "hipMalloc(partials)");
	if (!e) e |= dalloc_job(j, (void **)&j->colW, 4 * L.n_colW, "hipMalloc(colW)");
	if (!e) e |= dalloc_job(j, (void **)&j->colH, 4 * L.n_colH, "hipMalloc(colH)");
	if (!e) e |= dalloc_job(j, (void **)&j->denW, 4 * L.n_den, "hipMalloc(den)");
	if (!e) j->denH = j->denW + L.kt;
	if (!e) e |= dalloc_job(j, (void **)&j->dot, 8 * L.n_dot, "hipMalloc(dot)");
	if (!e) e |= dalloc_job(j, (void **)&j->cls, 4 * (size_t)m, "hipMalloc(cls)");
	if (!e) e |= dalloc_job(j, (void **)&j->last, 4 * (size_t)m, "hipMalloc(last)");
	if (!e) e |= dalloc_job(j, (void **)&j->st, sizeof(nmfk_state), "hipMalloc(state)");
	if (!e && L.n_shadow) {
		e |= dalloc_job(j, (void **)&j->Wb, 2 * L.n_shadow, "hipMalloc(BF16 copies)");
		if (!e)
			j->Hb = j->Wb + (size_t)n * L.kt;
	}
	if (stream) {
		j->stream = stream;   /* borrowed: the creator keeps it */
		j->borrowed_stream = 1;
	} else {
		if (!e) e |= run_resources(j);
		if (!e) e |= nmfk_hip_fail(hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking), "hipStreamCreate");
	}
	if (e) {
		char msg[512];
		snprintf(msg, sizeof msg, "%s", g_err);
		nmfk_job_destroy(j);
		nmfk_set_error("%s", msg);
		return NULL;
	}
	/* Zero W and H so their padding columns start (and stay) at zero. */
	hipMemsetAsync(j->W, 0, 4 * L.n_W, j->stream);
	if (j->Wb)
		hipMemsetAsync(j->Wb, 0, 2 * L.n_shadow, j->stream);
	hipMemsetAsync(j->H, 0, 4 * L.n_H, j->stream);
	if (j->owns_V) {
		hipMemsetAsync(j->V, 0, 4 * L.n_V, j->stream);
		hipMemsetAsync(j->Vt, 0, 4 * L.n_Vt, j->stream);
	}
	if (nmfk_hip_fail(hipStreamSynchronize(j->stream), "job create")) {
		nmfk_job_destroy(j);
		return NULL;
	}
	return j;
}

void nmfk_job_destroy(nmfk_job *j)
{
	if (!j)
		return;
	if (j->stream)
		hipStreamSynchronize(j->stream);
	if (j->chunk_exec)
		hipGraphExecDestroy(j->chunk_exec);
	if (j->streamed)
		nmfk_stream_free(j);
	if (j->owns_V) {
		dfree_job(j, j->V);
		dfree_job(j, j->Vt);
	}
	void *bufs[] = {j->W, j->H, j->part, j->colW, j->colH, j->denW, j->dot, j->cls, j->last, j->Wb, j->st,
			j->fpart, j->fcolW};
	for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; i++)
		dfree_job(j, bufs[i]);
	if (j->h_st)
		hipHostFree(j->h_st);
	if (j->ev[0])
		hipEventDestroy(j->ev[0]);
	if (j->ev[1])
		hipEventDestroy(j->ev[1]);
	if (j->stream && !j->borrowed_stream)
		hipStreamDestroy(j->stream);
	free(j);
}

int nmfk_job_info_get(const nmfk_job *j, nmfk_job_info *info)
{
	info->n = j->n, info->m = j->m, info->k = j->k, info->kt = j->kt;
	info->ldv = j->ldv, info->ldvt = j->ldvt;
	info->dV = j->V, info->dVt = j->Vt, info->dW = j->W, info->dH = j->H;
	info->stream = (void *)j->stream;
	return 0;
}

int nmfk_job_set_kernel_path(nmfk_job *j, int path)
{
	if (needs_device_V(j, "nmfk_job_set_kernel_path"))
		return NMFK_EINVAL;
	if (path < NMFK_PATH_AUTO || path > NMFK_PATH_MFMA_BF16) {
		nmfk_set_error("nmfk_job_set_kernel_path: unknown path %d", path);
		return NMFK_EINVAL;
	}
	nmfk_hplan h, w;
	nmfk_plan_half_path(j->m, j->n, j->kt, path, &h);
	nmfk_plan_half_path(j->n, j->m, j->kt, path, &w);
	nmfk_plan_set_rank(&h, j->kt, j->k);
	nmfk_plan_set_rank(&w, j->kt, j->k);
	if ((size_t)h.S * j->m * j->kt > j->part_cap || (size_t)w.S * j->n * j->kt > j->part_cap) {
		nmfk_set_error("nmfk_job_set_kernel_path: plan needs more partial space than the job holds "
			       "(tunables changed after nmfk_job_create)");
		return NMFK_EINVAL;
	}
	j->planH = h;
	j->planW = w;
	if (j->chunk_exec) {
		hipGraphExecDestroy(j->chunk_exec);
		j->chunk_exec = NULL;
	}
	return 0;
}

int nmfk_job_replan(nmfk_job *j, uint32_t target_wgs)
{
	if (needs_device_V(j, "nmfk_job_replan"))
		return NMFK_EINVAL;
	nmfk_hplan h, w;
	nmfk_plan_half_target(j->m, j->n, j->kt, (int)j->planH.path, target_wgs, &h);
	nmfk_plan_half_target(j->n, j->m, j->kt, (int)j->planW.path, target_wgs, &w);
	nmfk_plan_set_rank(&h, j->kt, j->k);
	nmfk_plan_set_rank(&w, j->kt, j->k);
	if ((size_t)h.S * j->m * j->kt > j->part_cap || (size_t)w.S * j->n * j->kt > j->part_cap) {
		nmfk_set_error("nmfk_job_replan: plan needs more partial space than the job holds");
		return NMFK_EINVAL;
	}
	j->planH = h;
	j->planW = w;
	if (j->chunk_exec) {
		hipGraphExecDestroy(j->chunk_exec);
		j->chunk_exec = NULL;
	}
	return 0;
}

/* ---------------------------------------------------------------- iteration schedule */

/* The one-pass kernel runs the job: V on the device, both half-steps planned on the FP32
 * matrix cores, and a shape k_fused_mfma takes. */
static int one_pass_ok(const nmfk_job *j, uint32_t *G, uint32_t *nwv)
{
	if (j->streamed || !j->V || j->planH.path != NMFK_PATH_MFMA32 || j->planW.path != NMFK_PATH_MFMA32)
		return 0;
	return nmfk_fused_plan(j->n, j->m, j->kt, G, nwv) == 0;
}

int nmfk_job_set_schedule(nmfk_job *j, int sched)
{
	uint32_t G, nwv;
	if (sched < NMFK_SCHED_AUTO || sched > NMFK_SCHED_ONE_PASS) {
		nmfk_set_error("nmfk_job_set_schedule: unknown schedule %d", sched);
		return NMFK_EINVAL;
	}
	if (sched == NMFK_SCHED_ONE_PASS && !one_pass_ok(j, &G, &nwv)) {
		nmfk_set_error("nmfk_job_set_schedule: one pass per iteration needs V on the device, the FP32 "
			       "matrix-core path at padded rank 16 (k 9..16) and m <= 1024 (job %ux%u, k=%u)",
			       j->n, j->m, j->k);
		return NMFK_EINVAL;
	}
	j->sched = sched;
	return 0;
}

/* Automatic is two passes: on one MI355X the one-pass kernel measured from 12% faster to
 * 30% slower than two passes over 20000-80000 x 200-1024 at k = 16 (slower at the benchmarked
 * 20000 x 1000), since both are bound by the FP32 matrix-core work, which fusing does not
 * reduce.  A requested one pass falls back to two where it cannot run (kernel path changed). */
int nmfk_job_schedule(const nmfk_job *j)
{
	uint32_t G, nwv;
	if (j->sched == NMFK_SCHED_ONE_PASS && one_pass_ok(j, &G, &nwv))
		return NMFK_SCHED_ONE_PASS;
	return NMFK_SCHED_TWO_PASS;
}

/* The pass's sums and column sums, sized by the current plan (reallocated, and the captured
 * interval graph dropped, when a tunable changed the number of workgroups). */
static int one_pass_buffers(nmfk_job *j)
{
	uint32_t G = 0, nwv = 0;
	if (!one_pass_ok(j, &G, &nwv)) {
		nmfk_set_error("one pass per iteration is not available for this job");
		return NMFK_EINVAL;
	}
	if (j->fpart && j->fG == G && j->fnwv == nwv)
		return 0;
	dfree_job(j, j->fpart);
	dfree_job(j, j->fcolW);
	j->fpart = j->fcolW = NULL;
	j->fG = j->fnwv = 0;
	if (j->chunk_exec) {
		hipGraphExecDestroy(j->chunk_exec);
		j->chunk_exec = NULL;
	}
	int e = dalloc_job(j, (void **)&j->fpart, 4 * (size_t)G * j->m * j->kt, "hipMalloc(one-pass sums)");
	if (!e)
		e = dalloc_job(j, (void **)&j->fcolW, 4 * (size_t)G * j->kt, "hipMalloc(one-pass column sums)");
	if (e)
		return e;
	j->fG = G;
	j->fnwv = nwv;
	return 0;
}

/* The pass over V: new W rows (upd_w) and the next H half-step's sums; clamp on test passes. */
static int one_pass(nmfk_job *j, int upd_w, int clamp, const nmfk_state *st)
{
	return nmfk_launch_fused(j->V, j->ldv, j->W, j->H, j->n, j->m, j->k, j->kt, j->fG, j->fnwv, j->fpart,
				 j->fcolW, upd_w, clamp, st, j->stream);
}

/* The H half-step from the last pass's sums. */
static int one_pass_finish(nmfk_job *j, const nmfk_state *st)
{
	return nmfk_launch_fused_finish(j->fpart, j->fG, j->fcolW, j->H, j->m, j->k, j->kt, st, j->stream);
}

int nmfk_job_plan_get(const nmfk_job *j, uint32_t out[8])
{
	out[0] = j->planH.nab, out[1] = j->planH.S, out[2] = j->planH.bchunk, out[3] = j->planH.path;
	out[4] = j->planW.nab, out[5] = j->planW.S, out[6] = j->planW.bchunk, out[7] = j->planW.path;
	return 0;
}

int nmfk_job_kernel_path(const nmfk_job *j)
{
	return (int)j->planW.path;
}

int nmfk_job_sync(nmfk_job *j)
{
	return nmfk_hip_fail(hipStreamSynchronize(j->stream), "hipStreamSynchronize");
}

/* ---------------------------------------------------------------- inputs */

int nmfk_job_set_V(nmfk_job *j, const float *V, size_t ldv)
{
	if (!v_writable(j))
		return NMFK_EINVAL;
	if (ldv < j->m) {
		nmfk_set_error("nmfk_job_set_V: ldv=%zu < m=%u", ldv, j->m);
		return NMFK_EINVAL;
	}
	int e = h2d_rows(j->V, j->ldv, V, ldv, j->m, j->n, j->stream);
	if (e)
		return e;
	return finish_V(j);
}

int nmfk_job_set_V_device(nmfk_job *j, const float *dV, size_t ldv)
{
	if (!v_writable(j))
		return NMFK_EINVAL;
	if (ldv < j->m) {
		nmfk_set_error("nmfk_job_set_V_device: ldv=%zu < m=%u", ldv, j->m);
		return NMFK_EINVAL;
	}
	int e = hipMemcpy2DAsync(j->V, j->ldv * 4, dV, ldv * 4, (size_t)j->m * 4, j->n, hipMemcpyDeviceToDevice,
				 j->stream);
	if (e)
		return nmfk_hip_fail(e, "copy V");
	return finish_V(j);
}

int nmfk_job_gen_V_lognormal(nmfk_job *j, uint64_t seed, double mu, double sigma)
{
	if (!v_writable(j))
		return NMFK_EINVAL;
	int e = nmfk_launch_gen_lognormal(j->V, j->n, j->m, (uint32_t)j->ldv, seed, mu, sigma, j->stream);
	if (e)
		return nmfk_hip_fail(e, "gen lognormal");
	return finish_V(j);
}

int nmfk_job_gen_V_planted(nmfk_job *j, uint64_t seed, uint32_t kstar, double sigma)
{
	if (!v_writable(j))
		return NMFK_EINVAL;
	double *scratch = NULL;
	int e = dalloc_job(j, (void **)&scratch, 8 * (size_t)kstar * ((size_t)j->n + j->m), "hipMalloc(planted)");
	if (e)
		return e;
	e = nmfk_launch_gen_planted(j->V, j->n, j->m, (uint32_t)j->ldv, kstar, seed, sigma, scratch, j->stream);
	if (e) {
		hipStreamSynchronize(j->stream);
		dfree_job(j, scratch);
		return nmfk_hip_fail(e, "gen planted");
	}
	e = finish_V(j);
	dfree_job(j, scratch);
	return e;
}

int nmfk_job_get_V(nmfk_job *j, float *V, size_t ldv)
{
	if (needs_device_V(j, "nmfk_job_get_V"))
		return NMFK_EINVAL;
	if (ldv < j->m) {
		nmfk_set_error("nmfk_job_get_V: ldv < m");
		return NMFK_EINVAL;
	}
	return d2h_rows(V, ldv, j->V, j->ldv, j->m, j->n, j->stream);
}

int nmfk_job_init_random(nmfk_job *j, uint64_t seed)
{
	/* W then H from the same seed, as set_random_values() draws d_W then d_H
	 * (src/NMF_GPU.c:356-430). */
	int e = nmfk_launch_uniform(j->W, j->n, j->k, j->kt, seed, 0, j->stream);
	if (!e)
		e = nmfk_launch_uniform(j->H, j->m, j->k, j->kt, seed, 1, j->stream);
	j->colW_valid = 0;
	if (e)
		return nmfk_hip_fail(e, "init random");
	return nmfk_job_sync(j);
}

static int put_factor(nmfk_job *j, float *d, uint32_t rows, const float *h, size_t ld)
{
	if (ld < j->k) {
		nmfk_set_error("leading dimension %zu < k=%u", ld, j->k);
		return NMFK_EINVAL;
	}
	/* the padding columns k..kt-1 are written as zeros */
	return h2d_rows(d, j->kt, h, ld, j->k, rows, j->stream);
}

int nmfk_job_set_WH(nmfk_job *j, const float *W, size_t ldw, const float *H, size_t ldh)
{
	int e = put_factor(j, j->W, j->n, W, ldw);
	if (!e)
		e = put_factor(j, j->H, j->m, H, ldh);
	j->colW_valid = 0;
	if (e)
		return e;
	return nmfk_job_sync(j);
}

static int get_factor(nmfk_job *j, const float *d, uint32_t rows, float *h, size_t ld)
{
	if (ld < j->k) {
		nmfk_set_error("leading dimension %zu < k=%u", ld, j->k);
		return NMFK_EINVAL;
	}
	if (nmfk_job_sync(j))
		return NMFK_EHIP;
	return d2h_rows(h, ld, d, j->kt, j->k, rows, j->stream);
}

int nmfk_job_get_W(nmfk_job *j, float *W, size_t ldw)
{
	return get_factor(j, j->W, j->n, W, ldw);
}

int nmfk_job_get_H(nmfk_job *j, float *H, size_t ldh)
{
	return get_factor(j, j->H, j->m, H, ldh);
}

/* ---------------------------------------------------------------- half-steps */

static int partial_H(nmfk_job *j, const nmfk_state *st)
{
	if (j->planH.path == NMFK_PATH_MFMA_BF16)
		return nmfk_launch_partial_bf16(j->V, j->ldv, j->H, j->Wb, j->m, j->n, j->kt, &j->planH, j->part,
						j->colW, j->ncolW, j->denW, st, j->stream);
	return nmfk_launch_partial(j->V, j->ldv, j->H, j->W, j->m, j->n, j->kt, &j->planH, j->part, j->colW,
				   j->ncolW, j->denW, st, j->stream);
}

static int partial_W(nmfk_job *j, const nmfk_state *st)
{
	if (j->planW.path == NMFK_PATH_MFMA_BF16)
		return nmfk_launch_partial_bf16(j->Vt, j->ldvt, j->W, j->Hb, j->n, j->m, j->kt,
						&j->planW, j->part, j->colH, j->ncolH, j->denH, st, j->stream);
	return nmfk_launch_partial(j->Vt, j->ldvt, j->W, j->H, j->n, j->m, j->kt, &j->planW, j->part, j->colH,
				   j->ncolH, j->denH, st, j->stream);
}

/* H half: the partial kernel streams V against W (and reduces colsum(W) into denW), the
 * finish kernel applies the update to H and leaves colsum(H) partials for the W half (and
 * the BF16 copy of H when the W half runs in BF16).  On the BF16 path W is read from its
 * BF16 copy, which the W half's finish (or nmfk_job_shadow_for_H) keeps current. */
int nmfk_enqueue_half_H(nmfk_job *j, const nmfk_state *st)
{
	int e = partial_H(j, st);
	const int sh = j->planW.path == NMFK_PATH_MFMA_BF16;
	if (!e)
		e = nmfk_launch_finish_shadow(j->part, j->planH.S, j->H, j->m, j->k, j->kt, j->denW, j->colH, 0, st,
					      sh ? j->Hb : NULL, j->stream);
	return e;
}

int nmfk_enqueue_half_W(nmfk_job *j, const nmfk_state *st, int clamp)
{
	int e = partial_W(j, st);
	const int sh = j->planH.path == NMFK_PATH_MFMA_BF16;
	if (!e)
		e = nmfk_launch_finish_shadow(j->part, j->planW.S, j->W, j->n, j->k, j->kt, j->denH, j->colW, clamp, st,
					      sh ? j->Wb : NULL, j->stream);
	return e;
}

int nmfk_job_shadow_for_H(nmfk_job *j)
{
	if (j->planH.path != NMFK_PATH_MFMA_BF16)
		return 0;
	return nmfk_launch_shadow(j->W, j->n, j->kt, j->Wb, j->stream);
}

int nmfk_job_shadow_for_W(nmfk_job *j)
{
	if (j->planW.path != NMFK_PATH_MFMA_BF16)
		return 0;
	return nmfk_launch_shadow(j->H, j->m, j->kt, j->Hb, j->stream);
}

/* update_H(): accum_W = colsum(W) (src/NMF_routines.c:720), then the H half-step. */
int nmfk_job_update_H(nmfk_job *j)
{
	if (j->streamed)
		return nmfk_stream_update_H(j);
	int e = nmfk_launch_colsum_parts(j->W, j->n, j->kt, j->colW, j->stream);
	if (!e)
		e = nmfk_job_shadow_for_H(j);
	if (!e)
		e = nmfk_enqueue_half_H(j, NULL);
	return nmfk_hip_fail(e, "update_H");
}

/* update_W(): accum_H = colsum(H) (src/NMF_routines.c:1360), then the W half-step. */
int nmfk_job_update_W(nmfk_job *j)
{
	if (j->streamed)
		return nmfk_stream_update_W(j);
	if (nmfk_job_refresh_Vt(j))
		return NMFK_EHIP;
	int e = nmfk_launch_colsum_parts(j->H, j->m, j->kt, j->colH, j->stream);
	if (!e)
		e = nmfk_job_shadow_for_W(j);
	if (!e)
		e = nmfk_enqueue_half_W(j, NULL, 0);
	j->colW_valid = 1;
	return nmfk_hip_fail(e, "update_W");
}

int nmfk_job_adjust(nmfk_job *j)
{
	int e = nmfk_launch_adjust(j->H, j->m, j->k, j->kt, j->stream);
	if (!e)
		e = nmfk_launch_adjust(j->W, j->n, j->k, j->kt, j->stream);
	j->colW_valid = 0;
	return nmfk_hip_fail(e, "adjust");
}

int nmfk_job_get_classification(nmfk_job *j, uint32_t *cls)
{
	int e = nmfk_launch_idx_max(j->H, j->m, j->k, j->kt, j->cls, j->stream);
	if (!e)
		e = hipMemcpyAsync(cls, j->cls, 4 * (size_t)j->m, hipMemcpyDeviceToHost, j->stream);
	if (e)
		return nmfk_hip_fail(e, "classification");
	return nmfk_job_sync(j);
}

/* ---------------------------------------------------------------- the loop */

/* One test interval: T iterations, the last of which clamps (src/NMF_GPU.c:611-638), then
 * the test.  One pass per iteration: the H update from the previous pass's sums, then the
 * pass (new W rows, and the sums for the next H update). */
static int enqueue_interval(nmfk_job *j, uint32_t T, int sched)
{
	int e = 0;
	for (uint32_t i = 0; i < T && !e; i++) {
		if (sched == NMFK_SCHED_ONE_PASS) {
			e = one_pass_finish(j, j->st);
			if (!e)
				e = one_pass(j, 1, i + 1 == T, j->st);
			continue;
		}
		e = nmfk_enqueue_half_H(j, j->st);
		if (!e)
			e = nmfk_enqueue_half_W(j, j->st, i + 1 == T);
	}
	if (!e)
		e = nmfk_launch_test(j->H, j->m, j->k, j->kt, j->cls, j->last, j->st, j->stream);
	return e;
}

static int build_interval_graph(nmfk_job *j, uint32_t T, int sched)
{
	if (j->chunk_exec && j->chunk_T == T && j->chunk_sched == sched)
		return 0;
	if (j->chunk_exec) {
		hipGraphExecDestroy(j->chunk_exec);
		j->chunk_exec = NULL;
	}
	hipGraph_t g = NULL;
	int e = hipStreamBeginCapture(j->stream, hipStreamCaptureModeThreadLocal);
	if (e)
		return nmfk_hip_fail(e, "hipStreamBeginCapture");
	int e2 = enqueue_interval(j, T, sched);
	e = hipStreamEndCapture(j->stream, &g);
	if (e2)
		return nmfk_hip_fail(e2, "capture interval");
	if (e)
		return nmfk_hip_fail(e, "hipStreamEndCapture");
	e = hipGraphInstantiate(&j->chunk_exec, g, NULL, NULL, 0);
	hipGraphDestroy(g);
	if (e)
		return nmfk_hip_fail(e, "hipGraphInstantiate");
	j->chunk_T = T;
	j->chunk_sched = sched;
	return 0;
}

int nmfk_job_run(nmfk_job *j, const nmfk_options *opt, nmfk_result *res)
{
	if (!opt || opt->k != j->k || opt->niter_test_conv == 0) {
		nmfk_set_error("nmfk_job_run: options do not match the job (k=%u, niter_test_conv>0)", j->k);
		return NMFK_EINVAL;
	}
	if (j->streamed)
		return nmfk_stream_run(j, opt, res);
	if (run_resources(j))
		return NMFK_EHIP;
	const uint32_t T = opt->niter_test_conv;
	const uint32_t niter_conv = opt->nIters / T, niter_rem = opt->nIters % T;
	const int sched = nmfk_job_schedule(j);
	if (sched == NMFK_SCHED_ONE_PASS) {
		if (one_pass_buffers(j))
			return NMFK_EHIP;
	} else if (nmfk_job_refresh_Vt(j)) {
		return NMFK_EHIP;
	}
	int e = nmfk_launch_state_reset(j->st, opt->stop_threshold, UINT32_MAX, j->last, j->m, j->stream);
	if (sched == NMFK_SCHED_ONE_PASS) {
		/* the sums of H(1) against W(0), H(0); every iteration then starts with its H update */
		if (!e)
			e = one_pass(j, 0, 0, j->st);
	} else {
		if (!e)
			e = nmfk_launch_colsum_parts(j->W, j->n, j->kt, j->colW, j->stream);
		if (!e)
			e = nmfk_job_shadow_for_H(j);
	}
	if (e)
		return nmfk_hip_fail(e, "run setup");
	if (niter_conv > 0 && (e = build_interval_graph(j, T, sched)) != 0)
		return e;
	e = nmfk_hip_fail(hipStreamSynchronize(j->stream), "run setup");
	if (e)
		return e;

	const double t0 = nmfk_now();
	int converged = 0;
	for (uint32_t c = 0; c < niter_conv; c++) {
		e = hipGraphLaunch(j->chunk_exec, j->stream);
		if (!e)
			e = hipMemcpyAsync(&j->h_st[c & 1], j->st, sizeof(nmfk_state), hipMemcpyDeviceToHost, j->stream);
		if (!e)
			e = hipEventRecord(j->ev[c & 1], j->stream);
		if (e)
			return nmfk_hip_fail(e, "interval launch");
		if (c >= 1) {
			e = hipEventSynchronize(j->ev[(c - 1) & 1]);
			if (e)
				return nmfk_hip_fail(e, "interval wait");
			nmfk_trace_state(j, &j->h_st[(c - 1) & 1]);
			if (j->h_st[(c - 1) & 1].converged) {
				converged = 1;
				break;
			}
		}
	}
	if (!converged && niter_rem) {
		for (uint32_t i = 0; i < niter_rem && !e; i++) {
			if (sched == NMFK_SCHED_ONE_PASS) {
				e = one_pass_finish(j, j->st);
				if (!e)
					e = one_pass(j, 1, 0, j->st);
				continue;
			}
			e = nmfk_enqueue_half_H(j, j->st);
			if (!e)
				e = nmfk_enqueue_half_W(j, j->st, 0);
		}
		if (e)
			return nmfk_hip_fail(e, "remaining iterations");
	}
	e = hipMemcpyAsync(&j->h_st[0], j->st, sizeof(nmfk_state), hipMemcpyDeviceToHost, j->stream);
	if (!e)
		e = hipStreamSynchronize(j->stream);
	if (e)
		return nmfk_hip_fail(e, "run");
	const double t1 = nmfk_now();
	const nmfk_state s = j->h_st[0];
	nmfk_trace_state(j, &s);
	j->colW_valid = 1;
	if (res) {
		res->converged = s.converged == 1;
		res->ntests = s.ntests;
		res->last_diff = s.last_diff;
		res->iterations = s.converged ? s.stop_test * T : opt->nIters;
		res->seconds = t1 - t0;
		res->pad_ = 0;
	}
	return 0;
}

int nmfk_job_run_traced(nmfk_job *j, const nmfk_options *opt, nmfk_result *res, uint64_t *diffs, uint32_t cap,
			uint32_t *ntrace)
{
	if (!diffs && cap) {
		nmfk_set_error("nmfk_job_run_traced: no trace buffer for %u entries", cap);
		return NMFK_EINVAL;
	}
	j->tracing = 1;
	j->diff_trace = diffs;
	j->trace_cap = diffs ? cap : 0;
	j->trace_len = 0;
	const int e = nmfk_job_run(j, opt, res);
	if (ntrace)
		*ntrace = j->trace_len;
	j->tracing = 0;
	j->diff_trace = NULL;
	j->trace_cap = j->trace_len = 0;
	return e;
}

/* ---------------------------------------------------------------- checks */

int nmfk_job_dot_kl(nmfk_job *j, double *dot_V, double *dot_VWH, double *kl)
{
	if (j->streamed)
		return nmfk_stream_dot_kl(j, dot_V, dot_VWH, kl);
	double *h = (double *)malloc(sizeof(double) * 3 * j->ndot);
	if (!h) {
		nmfk_set_error("out of host memory");
		return NMFK_ENOMEM;
	}
	int e = nmfk_launch_dot_kl(j->V, j->ldv, j->W, j->H, j->n, j->m, j->k, j->kt, j->dot, j->stream);
	if (!e)
		e = hipMemcpyAsync(h, j->dot, sizeof(double) * 3 * j->ndot, hipMemcpyDeviceToHost, j->stream);
	if (!e)
		e = hipStreamSynchronize(j->stream);
	if (e) {
		free(h);
		return nmfk_hip_fail(e, "dot/KL");
	}
	double s[3] = {0.0, 0.0, 0.0};
	for (uint32_t b = 0; b < j->ndot; b++)
		for (int q = 0; q < 3; q++)
			s[q] += h[3 * (size_t)b + q];
	free(h);
	if (dot_V)
		*dot_V = s[0];
	if (dot_VWH)
		*dot_VWH = s[1];
	if (kl)
		*kl = s[2];
	return 0;
}

/* ---------------------------------------------------------------- timing */

static int time_loop(nmfk_job *j, uint32_t iters, int kind, double *ms)
{
	hipEvent_t a, b;
	int e = hipEventCreate(&a);
	if (!e)
		e = hipEventCreate(&b);
	if (e)
		return nmfk_hip_fail(e, "hipEventCreate");
	e = hipEventRecord(a, j->stream);
	for (uint32_t i = 0; i < iters && !e; i++) {
		switch (kind) {
		case 0:
			e = nmfk_enqueue_half_H(j, NULL);
			if (!e)
				e = nmfk_enqueue_half_W(j, NULL, 0);
			break;
		case 1:
			e = partial_H(j, NULL);
			break;
		case 2:
			e = nmfk_launch_finish(j->part, j->planH.S, j->H, j->m, j->k, j->kt, j->denW, j->colH, 0,
					       NULL, j->stream);
			break;
		case 3:
			e = partial_W(j, NULL);
			break;
		case 4:
			e = nmfk_launch_finish(j->part, j->planW.S, j->W, j->n, j->k, j->kt, j->denH, j->colW, 0,
					       NULL, j->stream);
			break;
		}
	}
	if (!e)
		e = hipEventRecord(b, j->stream);
	if (!e)
		e = hipEventSynchronize(b);
	float f = 0.0f;
	if (!e)
		e = hipEventElapsedTime(&f, a, b);
	hipEventDestroy(a);
	hipEventDestroy(b);
	*ms = (double)f / (iters ? iters : 1);
	return nmfk_hip_fail(e, "timing");
}

int nmfk_measure_stream(size_t bytes, uint32_t iters, uint32_t blocks, double *read_GBps, double *copy_GBps)
{
	float *src = NULL, *dst = NULL;
	hipEvent_t a = NULL, b = NULL;
	hipStream_t s = NULL;
	const size_t nf = bytes / 16 * 4;
	int e = hipMalloc((void **)&src, nf * 4);
	if (!e)
		e = hipMalloc((void **)&dst, nf * 4);
	if (!e)
		e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
	if (!e)
		e = hipEventCreate(&a);
	if (!e)
		e = hipEventCreate(&b);
	if (!e)
		e = hipMemsetAsync(src, 0, nf * 4, s);
	if (!e)
		e = hipMemsetAsync(dst, 0, nf * 4, s);
	double ms[2] = {0.0, 0.0};
	for (int kind = 0; kind < 2 && !e; kind++) {
		e = nmfk_launch_stream(kind, src, dst, nf, blocks, s);   /* warm-up */
		if (!e)
			e = hipEventRecord(a, s);
		for (uint32_t i = 0; i < iters && !e; i++)
			e = nmfk_launch_stream(kind, src, dst, nf, blocks, s);
		if (!e)
			e = hipEventRecord(b, s);
		if (!e)
			e = hipEventSynchronize(b);
		float f = 0.0f;
		if (!e)
			e = hipEventElapsedTime(&f, a, b);
		ms[kind] = (double)f / (iters ? iters : 1);
	}
	if (a)
		hipEventDestroy(a);
	if (b)
		hipEventDestroy(b);
	if (s)
		hipStreamDestroy(s);
	hipFree(src);
	hipFree(dst);
	if (e)
		return nmfk_hip_fail(e, "nmfk_measure_stream");
	if (read_GBps)
		*read_GBps = (double)(nf * 4) / (ms[0] * 1e-3) / 1e9;
	if (copy_GBps)
		*copy_GBps = (double)(2 * nf * 4) / (ms[1] * 1e-3) / 1e9;
	return 0;
}

static int time_one_pass_loop(nmfk_job *j, uint32_t iters, int kind, double *ms)
{
	hipEvent_t a, b;
	int e = hipEventCreate(&a);
	if (!e)
		e = hipEventCreate(&b);
	if (e)
		return nmfk_hip_fail(e, "hipEventCreate");
	e = hipEventRecord(a, j->stream);
	for (uint32_t i = 0; i < iters && !e; i++) {
		if (kind != 1)
			e = one_pass_finish(j, NULL);
		if (!e && kind != 2)
			e = one_pass(j, 1, 0, NULL);
	}
	if (!e)
		e = hipEventRecord(b, j->stream);
	if (!e)
		e = hipEventSynchronize(b);
	float f = 0.0f;
	if (!e)
		e = hipEventElapsedTime(&f, a, b);
	hipEventDestroy(a);
	hipEventDestroy(b);
	*ms = (double)f / (iters ? iters : 1);
	return nmfk_hip_fail(e, "timing");
}

int nmfk_job_time_one_pass(nmfk_job *j, uint32_t iters, nmfk_timing_one_pass *t)
{
	if (needs_device_V(j, "nmfk_job_time_one_pass"))
		return NMFK_EINVAL;
	if (one_pass_buffers(j))
		return NMFK_EINVAL;
	int e = one_pass(j, 0, 0, NULL);
	if (e)
		return nmfk_hip_fail(e, "timing setup");
	double ms[3];
	for (int kind = 0; kind < 3; kind++) {
		e = time_one_pass_loop(j, iters, kind, &ms[kind]);
		if (e)
			return e;
	}
	t->ms_per_iter = ms[0];
	t->ms_pass = ms[1];
	t->ms_finish = ms[2];
	t->workgroups = j->fG;
	t->waves_per_workgroup = j->fnwv;
	return 0;
}

int nmfk_job_time_iterations(nmfk_job *j, uint32_t iters, nmfk_timing *t)
{
	if (needs_device_V(j, "nmfk_job_time_iterations"))
		return NMFK_EINVAL;
	int e = nmfk_launch_colsum_parts(j->W, j->n, j->kt, j->colW, j->stream);
	if (!e)
		e = nmfk_job_shadow_for_H(j);
	if (!e)
		e = nmfk_job_shadow_for_W(j);
	if (e)
		return nmfk_hip_fail(e, "timing setup");
	double ms[5];
	for (int kind = 0; kind < 5; kind++) {
		e = time_loop(j, iters, kind, &ms[kind]);
		if (e)
			return e;
	}
	t->ms_per_iter = ms[0];
	t->ms_partial_H = ms[1];
	t->ms_finish_H = ms[2];
	t->ms_partial_W = ms[3];
	t->ms_finish_W = ms[4];
	j->colW_valid = 0;
	return 0;
}
