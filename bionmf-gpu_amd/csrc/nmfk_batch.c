/*
 * nmfk_batch.c -- continuous batching of many KL-NMF factorizations on one MI355X.
 *
 * The reference runs one factorization per process (src/NMF_GPU.c:312-924).  Here a batch
 * holds up to <cap> jobs of uneven shape (n, m, k) in slots and advances all of them with
 * one launch per phase: partial and finish of the H half, partial and finish of the W
 * half, and every niter_test_conv iterations the test (two launches) and a gather of the
 * 32-byte stop-rule states, which are copied to pinned memory after the chunk.  A launch
 * walks a table of tiles; a tile is exactly one workgroup of the job's own single-job grid,
 * run by the same device code, so each job's results are bitwise those of nmfk_job_run on
 * the same job.
 *
 * Jobs join (nmfk_batch_add) and leave (nmfk_batch_poll) between chunks.  A finished job's
 * tiles stay in the tables as no-ops (its slot is marked empty on the device) until the
 * dead share of the tables passes a threshold, then the tables are rebuilt and the chunk
 * graph re-captured; joining jobs also trigger a rebuild.  One chunk = T iterations +
 * test + state gather, captured once as a hipGraph and replayed.
 */
#include "nmfk_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Partial tables per half and body class: VALU half-steps without and with lane groups
 * (a factor with few rows), matrix-core half-steps of kt 16 and of kt 32 in 1 x 4 waves, and
 * the same in 2 x 2 waves (a lone job's plan over a short contraction range), each launched
 * on its own so that it runs at its own occupancy.  Empty classes launch nothing. */
#define NCLS 6
enum { TAB_PART_H0, TAB_PART_W0 = TAB_PART_H0 + NCLS, TAB_FIN_H = TAB_PART_W0 + NCLS, TAB_FIN_W, TAB_TEST, NTAB };

static int half_cls(const nmfk_job *j, int half)
{
	const nmfk_hplan *p = half ? &j->planW : &j->planH;
	if (j->kt <= 8)
		return p->lb > 1 ? 1 : 0;
	return (j->kt == 16 ? 2 : 3) + (p->ap == 2 ? 2 : 0);
}

struct nmfk_batch {
	uint32_t T, cap, nslots;      /* nslots: highest slot in use + 1 */
	hipStream_t stream;
	hipEvent_t ev;
	nmfk_job **job;               /* per slot, NULL when free */
	nmfk_options *opt;
	double *t_add;
	uint32_t *slot_tiles;         /* tiles a slot holds in the current tables */
	uint8_t *fresh;               /* added after the last launched chunk: its gathered state is not yet valid */
	/* Job descriptors: h_desc is the host's working copy, rewritten by add and poll at any
	 * time; uploads go through two pinned staging copies, each rewritten only after its
	 * previous upload has completed (up_ev), so no buffer is ever changed under a copy. */
	nmfk_bjob *h_desc, *d_desc;
	nmfk_bjob *h_up[2];
	hipEvent_t up_ev[2];
	int up_pending[2], up_next;
	nmfk_state *d_states, *h_states;
	nmfk_tile *h_tiles, *d_tiles;
	size_t tiles_cap, dtiles_cap;
	uint32_t off[NTAB], cnt[NTAB];
	uint32_t total_tiles, dead_tiles;
	uint32_t active;
	int tables_dirty, desc_dirty;
	hipGraphExec_t exec;
	nmfk_batch_tuning tune;
};

void nmfk_batch_default_tuning(nmfk_batch_tuning *t)
{
	t->grid_partial = 0;
	t->grid_finish = 0;
	t->order = 0;
	t->chunks_per_poll = 1;
	t->compact_pct = 50;
}

static int hipck(int e, const char *what)
{
	return nmfk_hip_fail(e, what);
}

nmfk_batch *nmfk_batch_create(uint32_t niter_test_conv, uint32_t cap)
{
	if (niter_test_conv == 0 || cap == 0) {
		nmfk_set_error("nmfk_batch_create: niter_test_conv and capacity must be positive");
		return NULL;
	}
	nmfk_batch *b = (nmfk_batch *)calloc(1, sizeof(nmfk_batch));
	if (!b) {
		nmfk_set_error("nmfk_batch_create: out of host memory");
		return NULL;
	}
	b->T = niter_test_conv;
	b->cap = cap;
	nmfk_batch_default_tuning(&b->tune);
	b->job = (nmfk_job **)calloc(cap, sizeof(nmfk_job *));
	b->opt = (nmfk_options *)calloc(cap, sizeof(nmfk_options));
	b->t_add = (double *)calloc(cap, sizeof(double));
	b->slot_tiles = (uint32_t *)calloc(cap, sizeof(uint32_t));
	b->fresh = (uint8_t *)calloc(cap, 1);
	b->h_desc = (nmfk_bjob *)calloc(cap, sizeof(nmfk_bjob));
	int e = (!b->job || !b->opt || !b->t_add || !b->slot_tiles || !b->fresh || !b->h_desc) ? NMFK_ENOMEM : 0;
	if (e)
		nmfk_set_error("nmfk_batch_create: out of host memory");
	if (!e) e = hipck(hipMalloc((void **)&b->d_desc, cap * sizeof(nmfk_bjob)), "hipMalloc(batch jobs)");
	if (!e) e = hipck(hipMalloc((void **)&b->d_states, cap * sizeof(nmfk_state)), "hipMalloc(batch states)");
	if (!e) e = hipck(hipHostMalloc((void **)&b->h_states, cap * sizeof(nmfk_state), hipHostMallocDefault),
			  "hipHostMalloc(batch states)");
	if (!e) e = hipck(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
	if (!e) e = hipck(hipEventCreateWithFlags(&b->ev, hipEventDisableTiming), "hipEventCreate");
	for (int i = 0; i < 2 && !e; i++) {
		e = hipck(hipHostMalloc((void **)&b->h_up[i], cap * sizeof(nmfk_bjob), hipHostMallocDefault),
			  "hipHostMalloc(batch job staging)");
		if (!e)
			e = hipck(hipEventCreateWithFlags(&b->up_ev[i], hipEventDisableTiming), "hipEventCreate");
	}
	if (!e) e = hipck(hipMemsetAsync(b->d_desc, 0, cap * sizeof(nmfk_bjob), b->stream), "clear batch");
	if (!e) e = hipck(hipStreamSynchronize(b->stream), "clear batch");
	if (e) {
		char msg[512];
		snprintf(msg, sizeof msg, "%s", nmfk_last_error());
		nmfk_batch_destroy(b);
		nmfk_set_error("%s", msg);
		return NULL;
	}
	return b;
}

void nmfk_batch_destroy(nmfk_batch *b)
{
	if (!b)
		return;
	if (b->stream)
		hipStreamSynchronize(b->stream);
	if (b->exec)
		hipGraphExecDestroy(b->exec);
	hipFree(b->d_desc);
	hipFree(b->d_states);
	hipFree(b->d_tiles);
	if (b->h_states)
		hipHostFree(b->h_states);
	for (int i = 0; i < 2; i++) {
		if (b->h_up[i])
			hipHostFree(b->h_up[i]);
		if (b->up_ev[i])
			hipEventDestroy(b->up_ev[i]);
	}
	if (b->ev)
		hipEventDestroy(b->ev);
	if (b->stream)
		hipStreamDestroy(b->stream);
	free(b->job), free(b->opt), free(b->t_add), free(b->slot_tiles), free(b->fresh), free(b->h_desc), free(b->h_tiles);
	free(b);
}

int nmfk_batch_set_tuning(nmfk_batch *b, const nmfk_batch_tuning *t)
{
	if (t->chunks_per_poll == 0 || t->compact_pct > 100 || t->order > 1) {
		nmfk_set_error("nmfk_batch_set_tuning: chunks_per_poll >= 1, compact_pct <= 100, order 0 or 1");
		return NMFK_EINVAL;
	}
	b->tune = *t;
	b->tables_dirty = 1;
	return 0;
}

int nmfk_batch_get_tuning(const nmfk_batch *b, nmfk_batch_tuning *t)
{
	*t = b->tune;
	return 0;
}

uint32_t nmfk_batch_active(const nmfk_batch *b)
{
	return b->active;
}

/* Plans the batched bodies run: VALU <KT, 4, 4> for KT 4/8, MFMA <KT, 1, 4> and <KT, 2, 2>
 * for KT 16/32. */
static int plan_batchable(const nmfk_hplan *p, uint32_t kt)
{
	if (kt <= 8)
		return p->path == NMFK_PATH_VALU && p->ap == 4 && p->nw == 4;
	return (kt == 16 || kt == 32) && p->path == NMFK_PATH_MFMA32 && (p->ap == 1 || p->ap == 2) && p->nw == 4;
}

static void fill_desc(nmfk_bjob *d, const nmfk_job *j)
{
	memset(d, 0, sizeof *d);
	d->V = j->V, d->Vt = j->Vt, d->W = j->W, d->H = j->H, d->part = j->part;
	d->colW = j->colW, d->colH = j->colH, d->denW = j->denW, d->denH = j->denH;
	d->st = j->st, d->cls = j->cls, d->last = j->last;
	d->ldv = j->ldv, d->ldvt = j->ldvt;
	d->n = j->n, d->m = j->m, d->k = j->k, d->kt = j->kt;
	d->ncolW = j->ncolW, d->ncolH = j->ncolH;
	d->planH = j->planH, d->planW = j->planW;
	d->live = 1;
}

int nmfk_batch_add(nmfk_batch *b, nmfk_job *j, const nmfk_options *opt)
{
	if (!opt || opt->k != j->k || opt->niter_test_conv != b->T || opt->nIters < b->T) {
		nmfk_set_error("nmfk_batch_add: options must match the job's k and the batch's niter_test_conv, "
			       "with nIters >= niter_test_conv");
		return -1;
	}
	if (j->streamed) {
		nmfk_set_error("nmfk_batch_add: a streamed job's V stays in host memory; batched jobs need it on the device");
		return -1;
	}
	if (j->kt > NMFK_BATCH_KTMAX || !plan_batchable(&j->planH, j->kt) || !plan_batchable(&j->planW, j->kt)) {
		nmfk_set_error("nmfk_batch_add: job (k=%u) has a kernel plan the batched launches do not run "
			       "(batched: k <= %d on the default kernel paths)", j->k, NMFK_BATCH_KTMAX);
		return -1;
	}
	uint32_t s = 0;
	while (s < b->cap && b->job[s])
		s++;
	if (s == b->cap) {
		nmfk_set_error("nmfk_batch_add: batch is full (%u jobs)", b->cap);
		return -1;
	}
	/* The job's own stream may still hold work (V upload, initialisation). */
	if (nmfk_job_refresh_Vt(j) || nmfk_job_sync(j))
		return -1;
	int e = nmfk_launch_state_reset(j->st, opt->stop_threshold, opt->nIters / b->T, j->last, j->m, b->stream);
	if (!e)
		e = nmfk_launch_colsum_parts(j->W, j->n, j->kt, j->colW, b->stream);
	if (e) {
		hipck(e, "nmfk_batch_add");
		return -1;
	}
	b->job[s] = j;
	b->fresh[s] = 1;
	b->opt[s] = *opt;
	b->t_add[s] = nmfk_now();
	fill_desc(&b->h_desc[s], j);
	if (s + 1 > b->nslots)
		b->nslots = s + 1;
	b->active++;
	b->tables_dirty = 1;
	b->desc_dirty = 1;
	return (int)s;
}

/* Uploads descriptors [0, count) to the device, ordered on the batch stream, through the
 * staging copy whose last upload has finished. */
static int upload_desc(nmfk_batch *b, uint32_t count)
{
	if (count == 0)
		return 0;
	const int i = b->up_next;
	if (b->up_pending[i]) {
		int e = hipck(hipEventSynchronize(b->up_ev[i]), "batch job upload wait");
		if (e)
			return e;
		b->up_pending[i] = 0;
	}
	memcpy(b->h_up[i], b->h_desc, count * sizeof(nmfk_bjob));
	int e = hipMemcpyAsync(b->d_desc, b->h_up[i], count * sizeof(nmfk_bjob), hipMemcpyHostToDevice, b->stream);
	if (!e)
		e = hipEventRecord(b->up_ev[i], b->stream);
	if (e)
		return hipck(e, "upload jobs");
	b->up_pending[i] = 1;
	b->up_next = i ^ 1;
	return 0;
}

/* ---------------------------------------------------------------- tile tables */

static uint32_t cdiv(uint32_t a, uint32_t b)
{
	return (a + b - 1) / b;
}

static size_t tiles_of_job(const nmfk_job *j)
{
	return (size_t)j->planH.nab * j->planH.S + 1 + (size_t)j->planW.nab * j->planW.S + 1 +
	       cdiv(j->m, NMFK_FIN_ROWS) + cdiv(j->n, NMFK_FIN_ROWS) + cdiv(j->m, NMFK_TEST_ROWS);
}

static void push(nmfk_tile *t, uint32_t *n, uint32_t job, uint32_t bx, uint32_t by)
{
	t[*n].job = job, t[*n].bx = bx, t[*n].by = by, t[*n].pad_ = 0;
	(*n)++;
}

/* Partial tiles of one half for every live job of body class cls.  order 0: job by job;
 * order 1: split by split across jobs (round-robin), so concurrently running workgroups
 * belong to many jobs. */
static void push_partial(nmfk_batch *b, nmfk_tile *t, uint32_t *n, int half, int cls)
{
	uint32_t maxS = 0;
	for (uint32_t s = 0; s < b->nslots; s++)
		if (b->job[s] && half_cls(b->job[s], half) == cls) {
			const nmfk_hplan *p = half ? &b->job[s]->planW : &b->job[s]->planH;
			push(t, n, s, p->nab, 0);   /* denominator tile first: the finish needs it */
			if (p->S > maxS)
				maxS = p->S;
		}
	if (b->tune.order == 0) {
		for (uint32_t s = 0; s < b->nslots; s++)
			if (b->job[s] && half_cls(b->job[s], half) == cls) {
				const nmfk_hplan *p = half ? &b->job[s]->planW : &b->job[s]->planH;
				for (uint32_t by = 0; by < p->S; by++)
					for (uint32_t bx = 0; bx < p->nab; bx++)
						push(t, n, s, bx, by);
			}
	} else {
		for (uint32_t by = 0; by < maxS; by++)
			for (uint32_t s = 0; s < b->nslots; s++)
				if (b->job[s] && half_cls(b->job[s], half) == cls) {
					const nmfk_hplan *p = half ? &b->job[s]->planW : &b->job[s]->planH;
					if (by < p->S)
						for (uint32_t bx = 0; bx < p->nab; bx++)
							push(t, n, s, bx, by);
				}
	}
}

static int capture_chunk(nmfk_batch *b);

static int build_tables(nmfk_batch *b)
{
	size_t need = 0;
	for (uint32_t s = 0; s < b->nslots; s++)
		if (b->job[s])
			need += tiles_of_job(b->job[s]);
	if (need > UINT32_MAX) {
		nmfk_set_error("batch tile table too large");
		return NMFK_EINVAL;
	}
	if (need > b->tiles_cap) {
		nmfk_tile *t = (nmfk_tile *)realloc(b->h_tiles, need * sizeof(nmfk_tile));
		if (!t) {
			nmfk_set_error("out of host memory (tile table)");
			return NMFK_ENOMEM;
		}
		b->h_tiles = t;
		b->tiles_cap = need;
	}
	uint32_t n = 0;
	for (int half = 0; half < 2; half++)
		for (int cls = 0; cls < NCLS; cls++) {
			const int tab = TAB_PART_H0 + NCLS * half + cls;
			b->off[tab] = n;
			push_partial(b, b->h_tiles, &n, half, cls);
			b->cnt[tab] = n - b->off[tab];
		}
	b->off[TAB_FIN_H] = n;
	for (uint32_t s = 0; s < b->nslots; s++)
		if (b->job[s])
			for (uint32_t bx = 0; bx < cdiv(b->job[s]->m, NMFK_FIN_ROWS); bx++)
				push(b->h_tiles, &n, s, bx, 0);
	b->cnt[TAB_FIN_H] = n - b->off[TAB_FIN_H];
	b->off[TAB_FIN_W] = n;
	for (uint32_t s = 0; s < b->nslots; s++)
		if (b->job[s])
			for (uint32_t bx = 0; bx < cdiv(b->job[s]->n, NMFK_FIN_ROWS); bx++)
				push(b->h_tiles, &n, s, bx, 0);
	b->cnt[TAB_FIN_W] = n - b->off[TAB_FIN_W];
	b->off[TAB_TEST] = n;
	for (uint32_t s = 0; s < b->nslots; s++)
		if (b->job[s])
			for (uint32_t bx = 0; bx < cdiv(b->job[s]->m, NMFK_TEST_ROWS); bx++)
				push(b->h_tiles, &n, s, bx, 0);
	b->cnt[TAB_TEST] = n - b->off[TAB_TEST];
	for (uint32_t s = 0; s < b->cap; s++)
		b->slot_tiles[s] = b->job[s] ? (uint32_t)tiles_of_job(b->job[s]) : 0;
	b->total_tiles = n;
	b->dead_tiles = 0;

	/* Nothing of the batch is running (callers synchronise first), so the device tables can
	 * be replaced in place. */
	int e = 0;
	if (n > b->dtiles_cap) {
		/* grow by half again, so that a growing batch reallocates rarely */
		const size_t want = (size_t)n + n / 2;
		hipFree(b->d_tiles);
		b->d_tiles = NULL;
		b->dtiles_cap = 0;
		e = hipck(hipMalloc((void **)&b->d_tiles, want * sizeof(nmfk_tile)), "hipMalloc(tiles)");
		if (e)
			return e;
		b->dtiles_cap = want;
	}
	if (n)
		e = hipck(hipMemcpyAsync(b->d_tiles, b->h_tiles, (size_t)n * sizeof(nmfk_tile), hipMemcpyHostToDevice,
					 b->stream), "upload tiles");
	if (!e)
		e = upload_desc(b, b->cap);
	/* the pageable tile table is read by the copy above: wait for it before returning */
	if (!e)
		e = hipck(hipStreamSynchronize(b->stream), "upload tables");
	if (e)
		return e;
	b->desc_dirty = 0;
	b->tables_dirty = 0;
	return capture_chunk(b);
}

/* One iteration of every live job: four launches. */
static int enqueue_iteration(nmfk_batch *b, int clamp)
{
	const nmfk_tile *t = b->d_tiles;
	int e = 0;
	for (int cls = 0; cls < NCLS && !e; cls++)
		e = nmfk_launch_batch_partial(b->d_desc, t + b->off[TAB_PART_H0 + cls], b->cnt[TAB_PART_H0 + cls], 0,
					      cls, b->tune.grid_partial, b->stream);
	if (!e)
		e = nmfk_launch_batch_finish(b->d_desc, t + b->off[TAB_FIN_H], b->cnt[TAB_FIN_H], 0, 0,
					     b->tune.grid_finish, b->stream);
	for (int cls = 0; cls < NCLS && !e; cls++)
		e = nmfk_launch_batch_partial(b->d_desc, t + b->off[TAB_PART_W0 + cls], b->cnt[TAB_PART_W0 + cls], 1,
					      cls, b->tune.grid_partial, b->stream);
	if (!e)
		e = nmfk_launch_batch_finish(b->d_desc, t + b->off[TAB_FIN_W], b->cnt[TAB_FIN_W], 1, clamp,
					     b->tune.grid_finish, b->stream);
	return e;
}

static int enqueue_chunk(nmfk_batch *b)
{
	int e = 0;
	for (uint32_t i = 0; i < b->T && !e; i++)
		e = enqueue_iteration(b, i + 1 == b->T);
	if (!e)
		e = nmfk_launch_batch_test(b->d_desc, b->d_tiles + b->off[TAB_TEST], b->cnt[TAB_TEST], b->nslots,
					   b->tune.grid_finish, b->stream);
	if (!e)
		e = nmfk_launch_batch_gather_states(b->d_desc, b->nslots, b->d_states, b->stream);
	return e;
}

static int capture_chunk(nmfk_batch *b)
{
	if (b->exec) {
		hipGraphExecDestroy(b->exec);
		b->exec = NULL;
	}
	hipGraph_t g = NULL;
	int e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
	if (e)
		return hipck(e, "hipStreamBeginCapture");
	int e2 = enqueue_chunk(b);
	e = hipStreamEndCapture(b->stream, &g);
	if (e2)
		return hipck(e2, "capture batch chunk");
	if (e)
		return hipck(e, "hipStreamEndCapture");
	e = hipGraphInstantiate(&b->exec, g, NULL, NULL, 0);
	hipGraphDestroy(g);
	return hipck(e, "hipGraphInstantiate");
}

int nmfk_batch_step(nmfk_batch *b, uint32_t chunks)
{
	if (b->active == 0)
		return 0;
	int e = 0;
	if (b->tables_dirty) {
		e = hipck(hipStreamSynchronize(b->stream), "batch step");
		if (!e)
			e = build_tables(b);
	} else if (b->desc_dirty) {
		e = upload_desc(b, b->nslots);
		b->desc_dirty = 0;
	}
	for (uint32_t c = 0; c < chunks && !e; c++)
		e = hipck(hipGraphLaunch(b->exec, b->stream), "batch chunk");
	/* The last chunk's gathered states to pinned memory, after the graph rather than as a
	 * node in it, as the single-job loop copies its state: the chunk graph holds kernels
	 * only.  A batch whose slots grew was re-captured above, so the gather covers nslots. */
	if (!e && chunks)
		e = hipck(hipMemcpyAsync(b->h_states, b->d_states, b->nslots * sizeof(nmfk_state),
					 hipMemcpyDeviceToHost, b->stream), "batch states");
	if (!e)
		e = hipck(hipEventRecord(b->ev, b->stream), "batch event");
	if (!e && chunks)
		memset(b->fresh, 0, b->cap);   /* every added job is in the tables these chunks ran */
	return e;
}

/* Result fields as nmfk_job_run reports them. */
static void fill_result(const nmfk_state *s, const nmfk_options *o, uint32_t T, double seconds, nmfk_result *r)
{
	r->converged = s->converged == 1;
	r->ntests = s->ntests;
	r->last_diff = s->last_diff;
	r->iterations = s->converged == 1 ? s->stop_test * T : o->nIters;
	r->seconds = seconds;
	r->pad_ = 0;
}

int nmfk_batch_poll(nmfk_batch *b, nmfk_job **done, nmfk_result *res, uint32_t cap, uint32_t *ndone)
{
	*ndone = 0;
	int e = hipck(hipEventSynchronize(b->ev), "batch poll");
	if (e)
		return e;
	const double now = nmfk_now();
	for (uint32_t s = 0; s < b->nslots && *ndone < cap; s++) {
		nmfk_job *j = b->job[s];
		if (!j || b->fresh[s])
			continue;
		const nmfk_state st = b->h_states[s];
		if (!st.converged)
			continue;
		/* Budget spent without converging: the reference's leftover iterations
		 * (src/NMF_GPU.c:760-795), no test, on the job's own stream. */
		const uint32_t rem = b->opt[s].nIters % b->T;
		if (st.converged == 2 && rem) {
			for (uint32_t i = 0; i < rem && !e; i++) {
				e = nmfk_enqueue_half_H(j, NULL);
				if (!e)
					e = nmfk_enqueue_half_W(j, NULL, 0);
			}
			if (!e)
				e = hipStreamSynchronize(j->stream);
			if (e)
				return hipck(e, "leftover iterations");
		}
		j->colW_valid = 1;
		if (done)
			done[*ndone] = j;
		if (res)
			fill_result(&st, &b->opt[s], b->T, now - b->t_add[s], &res[*ndone]);
		(*ndone)++;
		b->job[s] = NULL;
		b->h_desc[s].live = 0;
		b->desc_dirty = 1;
		b->dead_tiles += b->slot_tiles[s];
		b->active--;
	}
	while (b->nslots > 0 && !b->job[b->nslots - 1] && !b->slot_tiles[b->nslots - 1])
		b->nslots--;
	if (b->active && (uint64_t)b->dead_tiles * 100 > (uint64_t)b->tune.compact_pct * b->total_tiles)
		b->tables_dirty = 1;
	return 0;
}

/* ---------------------------------------------------------------- autotuner */

/* Times <iters> batched iterations (no tests) of the current jobs under tuning <t>. */
static int time_config(nmfk_batch *b, const nmfk_batch_tuning *t, uint32_t iters, double *ms)
{
	b->tune = *t;
	int e = build_tables(b);
	if (e)
		return e;
	hipEvent_t a, z;
	e = hipck(hipEventCreate(&a), "hipEventCreate");
	if (!e)
		e = hipck(hipEventCreate(&z), "hipEventCreate");
	if (e)
		return e;
	e = enqueue_iteration(b, 0);   /* warm-up */
	if (!e)
		e = hipEventRecord(a, b->stream);
	for (uint32_t i = 0; i < iters && !e; i++)
		e = enqueue_iteration(b, 0);
	if (!e)
		e = hipEventRecord(z, b->stream);
	if (!e)
		e = hipEventSynchronize(z);
	float f = 0.0f;
	if (!e)
		e = hipEventElapsedTime(&f, a, z);
	hipEventDestroy(a);
	hipEventDestroy(z);
	*ms = (double)f / (iters ? iters : 1);
	return hipck(e, "autotune timing");
}

int nmfk_batch_autotune(nmfk_batch *b, uint32_t iters, nmfk_batch_tuning *best_out, double *best_ms)
{
	if (b->active == 0)
		return 0;
	/* The trial iterations change W and H: keep copies and restore them afterwards. */
	size_t bytes = 0;
	for (uint32_t s = 0; s < b->nslots; s++)
		if (b->job[s])
			bytes += 4 * ((size_t)b->job[s]->n + b->job[s]->m) * b->job[s]->kt;
	float *save = NULL;
	int e = hipck(hipMalloc((void **)&save, bytes ? bytes : 4), "hipMalloc(autotune)");
	if (e)
		return e;
	size_t o = 0;
	for (uint32_t s = 0; s < b->nslots && !e; s++) {
		nmfk_job *j = b->job[s];
		if (!j)
			continue;
		e = hipMemcpyAsync(save + o, j->W, 4 * (size_t)j->n * j->kt, hipMemcpyDeviceToDevice, b->stream);
		o += (size_t)j->n * j->kt;
		if (!e)
			e = hipMemcpyAsync(save + o, j->H, 4 * (size_t)j->m * j->kt, hipMemcpyDeviceToDevice, b->stream);
		o += (size_t)j->m * j->kt;
	}
	if (e) {
		hipFree(save);
		return hipck(e, "autotune save");
	}
	uint32_t tiles_max = 0;
	for (uint32_t s = 0; s < b->nslots; s++)
		if (b->job[s]) {
			const uint32_t a = b->job[s]->planH.nab * b->job[s]->planH.S;
			const uint32_t c = b->job[s]->planW.nab * b->job[s]->planW.S;
			tiles_max += a > c ? a : c;
		}
	/* Candidates: persistent grid of the partial launches (0 = one workgroup per tile, else
	 * several tiles per workgroup) x tile order (job-major / split-major). */
	const uint32_t grids[] = {0, 4096, 2048, 1024};
	nmfk_batch_tuning cur = b->tune, best = b->tune;
	double bms = 1e30;
	for (int gi = 0; gi < 4 && !e; gi++) {
		if (grids[gi] && grids[gi] * 2 > tiles_max)
			continue;
		for (uint32_t order = 0; order < 2 && !e; order++) {
			nmfk_batch_tuning t = cur;
			t.grid_partial = grids[gi];
			t.order = order;
			double ms = 0.0;
			e = time_config(b, &t, iters, &ms);
			if (!e && ms < bms) {
				bms = ms;
				best = t;
			}
		}
	}
	o = 0;
	for (uint32_t s = 0; s < b->nslots; s++) {
		nmfk_job *j = b->job[s];
		if (!j)
			continue;
		int e2 = hipMemcpyAsync(j->W, save + o, 4 * (size_t)j->n * j->kt, hipMemcpyDeviceToDevice, b->stream);
		o += (size_t)j->n * j->kt;
		if (!e2)
			e2 = hipMemcpyAsync(j->H, save + o, 4 * (size_t)j->m * j->kt, hipMemcpyDeviceToDevice, b->stream);
		o += (size_t)j->m * j->kt;
		if (!e2)
			e2 = nmfk_launch_colsum_parts(j->W, j->n, j->kt, j->colW, b->stream);
		if (!e && e2)
			e = e2;
	}
	int e3 = hipStreamSynchronize(b->stream);
	hipFree(save);
	if (!e)
		e = e3;
	if (e)
		return hipck(e, "autotune");
	b->tune = best;
	b->tables_dirty = 1;
	if (best_out)
		*best_out = best;
	if (best_ms)
		*best_ms = bms;
	return 0;
}
