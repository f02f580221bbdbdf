/*
 * nmfk_serve.c -- serving loop: a stream of factorization requests on one MI355X.
 *
 * A request is one input matrix (generated on the device from its seed, as the synthetic
 * benchmark inputs are) factorized at rank k from <restarts> seeded starting points.  The
 * reference serves one request per process run (src/NMF_GPU.c:1034-1269); here requests
 * arrive on a wall-clock schedule and are admitted in arrival order as soon as their
 * arrival time has passed, their device memory (nmfk_job_bytes; restarts share V) fits in
 * the HBM budget and the batch has free slots.  Every restart is a job of one shared
 * nmfk_batch, so work of requests that arrived at different times overlaps on the device;
 * a request retires when all its restarts have met the stop rule (or nIters), its memory
 * is released and its latency (finish - arrival) recorded.
 *
 * The host loop enqueues the next chunk before doing admission work (matrix generation,
 * job set-up), so the device keeps iterating while new requests are prepared.
 *
 * Device memory comes from an arena reserved within the HBM budget (nmfk_arena.c), and all
 * set-up work runs on one stream created before the loop: admitting and retiring requests
 * calls no hipMalloc, hipFree or stream creation, so the host thread never waits for the
 * device to drain (hipFree would).
 */
#include "nmfk_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

void nmfk_serve_default_config(nmfk_serve_config *c)
{
	memset(c, 0, sizeof *c);
	c->hbm_budget = (size_t)200 << 30;
	c->capacity = 1024;
	c->nIters = 2000;
	c->niter_test_conv = 10;
	c->stop_threshold = 40;
	c->mu = 5.0;
	c->sigma = 1.5;
	c->chunks_per_poll = 1;
	c->plan_target = 64;   /* 100 requests at 5/s: 4.73 req/s, p99 10.0 s (32: 4.47, 11.3; 16: 4.01, 13.9) */
}

/* Device bytes of one request: its own job (V, Vt and buffers) plus restarts - 1 jobs that
 * share V, each buffer counted as the arena places it. */
size_t nmfk_request_bytes(uint32_t n, uint32_t m, uint32_t k, uint32_t restarts)
{
	return nmfk_job_bytes(n, m, k) + (size_t)(restarts ? restarts - 1 : 0) * nmfk_job_bytes_shared(n, m, k);
}

typedef struct req_state {
	nmfk_job **jobs;      /* restarts jobs; jobs[0] owns V */
	uint32_t left;        /* restarts not finished */
	size_t bytes;
	int admitted;
} req_state;

static void free_request(req_state *r, uint32_t restarts)
{
	if (!r->jobs)
		return;
	for (uint32_t i = restarts; i-- > 0;)   /* shared jobs before the owner of V */
		nmfk_job_destroy(r->jobs[i]);
	free(r->jobs);
	r->jobs = NULL;
}

static int cmp_double(const void *a, const void *b)
{
	const double x = *(const double *)a, y = *(const double *)b;
	return x < y ? -1 : x > y;
}

static double quantile(double *sorted, uint32_t n, double q)
{
	if (n == 0)
		return 0.0;
	/* linear interpolation between closest ranks (numpy's default) */
	const double pos = q * (double)(n - 1);
	const uint32_t i = (uint32_t)pos;
	const double f = pos - (double)i;
	return i + 1 < n ? sorted[i] * (1.0 - f) + sorted[i + 1] * f : sorted[n - 1];
}

static void sleep_until(double t0, double t)
{
	double now = nmfk_now() - t0;
	if (t <= now)
		return;
	const double d = t - now;
	struct timespec ts;
	ts.tv_sec = (time_t)d;
	ts.tv_nsec = (long)((d - (double)ts.tv_sec) * 1e9);
	nanosleep(&ts, NULL);
}

/* Creates the request's jobs: V from its seed, restart r from nmfk_sweep_seed(seed, k, r).
 * Their buffers come from the arena and their set-up runs on the serving loop's stream. */
static int admit(const nmfk_serve_config *cfg, const nmfk_request *q, req_state *r, nmfk_batch *b,
		 nmfk_arena *arena, hipStream_t stream)
{
	r->jobs = (nmfk_job **)calloc(q->restarts, sizeof(nmfk_job *));
	if (!r->jobs) {
		nmfk_set_error("out of host memory");
		return NMFK_ENOMEM;
	}
	nmfk_options opt;
	opt.k = q->k;
	opt.nIters = cfg->nIters;
	opt.niter_test_conv = cfg->niter_test_conv;
	opt.stop_threshold = cfg->stop_threshold;
	int e = 0;
	r->jobs[0] = nmfk_job_create_in(arena, stream, q->n, q->m, q->k, NULL);
	if (!r->jobs[0])
		return NMFK_EBUDGET;
	e = nmfk_job_gen_V_lognormal(r->jobs[0], q->seed, cfg->mu, cfg->sigma);
	for (uint32_t i = 1; i < q->restarts && !e; i++) {
		r->jobs[i] = nmfk_job_create_in(arena, stream, q->n, q->m, q->k, r->jobs[0]);
		if (!r->jobs[i])
			e = NMFK_EBUDGET;
	}
	if (e)
		return e;   /* nothing was added to the batch yet */
	for (uint32_t i = 0; i < q->restarts && !e; i++) {
		if (cfg->plan_target)
			e = nmfk_job_replan(r->jobs[i], cfg->plan_target);
		if (!e)
			e = nmfk_job_init_random(r->jobs[i], nmfk_sweep_seed(q->seed, q->k, i));
		if (!e && nmfk_batch_add(b, r->jobs[i], &opt) < 0)
			e = NMFK_EINVAL;
	}
	return e;
}

/* Copies a retired restart's outcome into the caller's record (buffers optional). */
static int record_restart(nmfk_job *j, const nmfk_result *r, double kl, nmfk_restart_result *o)
{
	o->iterations = r->iterations;
	o->converged = r->converged;
	o->ntests = r->ntests;
	o->last_diff = r->last_diff;
	o->kl = kl;
	int e = 0;
	/* the class vector of the final H, as nmfk_job_get_classification gives it */
	if (o->cls && (e = nmfk_job_get_classification(j, o->cls)) != 0)
		return e;
	if (o->W && (e = nmfk_job_get_W(j, o->W, j->k)) != 0)
		return e;
	if (o->H && (e = nmfk_job_get_H(j, o->H, j->k)) != 0)
		return e;
	return nmfk_job_sync(j);
}

int nmfk_serve_run(const nmfk_serve_config *cfg, const nmfk_request *req, uint32_t nreq, nmfk_request_result *out,
		   nmfk_serve_stats *stats)
{
	return nmfk_serve_run_detail(cfg, req, nreq, out, stats, NULL);
}

int nmfk_serve_run_detail(const nmfk_serve_config *cfg, const nmfk_request *req, uint32_t nreq,
			  nmfk_request_result *out, nmfk_serve_stats *stats, nmfk_restart_result *restarts)
{
	for (uint32_t i = 0; i < nreq; i++) {
		const nmfk_request *q = &req[i];
		if (q->restarts == 0 || q->k < 1 || q->k > NMFK_BATCH_KTMAX || (i && q->arrival < req[i - 1].arrival)) {
			nmfk_set_error("nmfk_serve_run: request %u: need restarts >= 1, 1 <= k <= %d, arrivals in order", i,
				       NMFK_BATCH_KTMAX);
			return NMFK_EINVAL;
		}
		if (q->restarts > cfg->capacity || nmfk_request_bytes(q->n, q->m, q->k, q->restarts) > cfg->hbm_budget) {
			nmfk_set_error("nmfk_serve_run: request %u can never fit (restarts > capacity or bytes > budget)", i);
			return NMFK_EBUDGET;
		}
	}
	nmfk_batch *b = nmfk_batch_create(cfg->niter_test_conv, cfg->capacity);
	if (!b)
		return NMFK_EINVAL;
	/* The arena's first segment is reserved here, before the clock starts: the largest
	 * request, at least 16 GiB, within the budget. */
	size_t first_seg = (size_t)16 << 30;
	for (uint32_t i = 0; i < nreq; i++) {
		const size_t need = nmfk_request_bytes(req[i].n, req[i].m, req[i].k, req[i].restarts);
		if (need > first_seg)
			first_seg = need;
	}
	if (first_seg > cfg->hbm_budget)
		first_seg = cfg->hbm_budget;
	nmfk_arena *arena = nmfk_arena_create(cfg->hbm_budget, first_seg);
	hipStream_t stream = NULL;
	req_state *rs = (req_state *)calloc(nreq ? nreq : 1, sizeof(req_state));
	nmfk_job **done = (nmfk_job **)malloc(cfg->capacity * sizeof(nmfk_job *));
	nmfk_result *dres = (nmfk_result *)malloc(cfg->capacity * sizeof(nmfk_result));
	double *lat = (double *)malloc((nreq ? nreq : 1) * sizeof(double));
	size_t *roff = (size_t *)malloc((nreq ? nreq : 1) * sizeof(size_t));   /* first restart record of request i */
	/* job -> request map, by slot-independent linear search over active requests' jobs */
	if (!arena || !rs || !done || !dres || !lat || !roff) {
		free(rs), free(done), free(dres), free(lat), free(roff);
		nmfk_arena_destroy(arena);
		nmfk_batch_destroy(b);
		nmfk_set_error("out of host memory");
		return NMFK_ENOMEM;
	}
	for (uint32_t i = 0, o = 0; i < nreq; o += req[i].restarts, i++)
		roff[i] = o;
	int e = nmfk_hip_fail(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate(serve)");
	if (!e && nreq) {   /* reserve the first segment now */
		void *p = nmfk_arena_alloc(arena, first_seg);
		if (!p) {
			nmfk_set_error("nmfk_serve_run: cannot reserve %zu bytes of device memory", first_seg);
			e = NMFK_EBUDGET;
		}
		nmfk_arena_free(arena, p);
	}
	for (uint32_t i = 0; i < nreq; i++) {
		memset(&out[i], 0, sizeof out[i]);
		out[i].kl_best = INFINITY;
	}
	size_t used = 0, peak = 0;
	uint32_t active_jobs = 0, peak_jobs = 0, next = 0, finished = 0, first_active = 0;
	const double t0 = nmfk_now();
	double t_step = 0.0, t_admit = 0.0, t_wait = 0.0, t_retire = 0.0, tm;
	uint32_t nchunks = 0;
	while (!e && finished < nreq) {
		/* 1. keep the device busy */
		tm = nmfk_now();
		if (nmfk_batch_active(b) > 0) {
			e = nmfk_batch_step(b, cfg->chunks_per_poll);
			nchunks++;
		}
		t_step += nmfk_now() - tm;
		tm = nmfk_now();
		/* 2. admit what has arrived and fits, in arrival order */
		while (!e && next < nreq) {
			const nmfk_request *q = &req[next];
			const double now = nmfk_now() - t0;
			if (q->arrival > now)
				break;
			const size_t need = nmfk_request_bytes(q->n, q->m, q->k, q->restarts);
			if (used + need > cfg->hbm_budget || active_jobs + q->restarts > cfg->capacity)
				break;
			e = admit(cfg, q, &rs[next], b, arena, stream);
			if (e == NMFK_EBUDGET && rs[next].jobs && rs[next].jobs[q->restarts - 1] == NULL) {
				/* The budget has room but the arena has no block that large (its free space is
				 * split): undo, and wait for retirements; with nothing running, release the
				 * arena's empty segments and try once more. */
				free_request(&rs[next], q->restarts);
				e = 0;
				if (active_jobs > 0)
					break;
				nmfk_arena_trim(arena);
				e = admit(cfg, q, &rs[next], b, arena, stream);
			}
			if (e)
				break;
			rs[next].bytes = need;
			rs[next].left = q->restarts;
			rs[next].admitted = 1;
			out[next].admitted = nmfk_now() - t0;
			used += need;
			active_jobs += q->restarts;
			if (used > peak)
				peak = used;
			if (active_jobs > peak_jobs)
				peak_jobs = active_jobs;
			next++;
		}
		t_admit += nmfk_now() - tm;
		if (e)
			break;
		if (nmfk_batch_active(b) == 0) {
			if (next < nreq)
				sleep_until(t0, req[next].arrival);
			continue;
		}
		/* 3. retire finished jobs and requests */
		uint32_t nd = 0;
		tm = nmfk_now();
		e = nmfk_batch_poll(b, done, dres, cfg->capacity, &nd);
		t_wait += nmfk_now() - tm;
		tm = nmfk_now();
		for (uint32_t d = 0; d < nd && !e; d++) {
			uint32_t qi = first_active, ri = 0;
			int found = 0;
			for (; qi < next && !found; qi++) {
				if (!rs[qi].jobs)
					continue;
				for (ri = 0; ri < req[qi].restarts; ri++)
					if (rs[qi].jobs[ri] == done[d]) {
						found = 1;
						break;
					}
				if (found)
					break;
			}
			if (!found) {
				nmfk_set_error("nmfk_serve_run: finished job not found");
				e = NMFK_EINVAL;
				break;
			}
			nmfk_request_result *o = &out[qi];
			double kl = 0.0;
			e = nmfk_job_dot_kl(done[d], NULL, NULL, &kl);
			if (!e && restarts)
				e = record_restart(done[d], &dres[d], kl, &restarts[roff[qi] + ri]);
			if (e)
				break;
			if (dres[d].iterations > o->iterations_max)
				o->iterations_max = dres[d].iterations;
			o->iterations_sum += dres[d].iterations;
			o->converged += dres[d].converged;
			if (kl < o->kl_best) {
				o->kl_best = kl;
				o->best_restart = ri;
			}
			if (--rs[qi].left == 0) {
				o->finished = nmfk_now() - t0;
				o->latency = o->finished - req[qi].arrival;
				lat[finished++] = o->latency;
				if (cfg->progress_every && finished % cfg->progress_every == 0) {
					fprintf(stderr, "nmfk_serve: %u/%u requests done at %.1f s (%u jobs active)\n", finished,
						nreq, o->finished, active_jobs - req[qi].restarts);
					fflush(stderr);
				}
				used -= rs[qi].bytes;
				active_jobs -= req[qi].restarts;
				free_request(&rs[qi], req[qi].restarts);
				while (first_active < next && !rs[first_active].jobs)
					first_active++;
			}
		}
		t_retire += nmfk_now() - tm;
	}
	const double t1 = nmfk_now() - t0;
	/* The batch first: destroying it waits for its stream, which may still read jobs. */
	nmfk_batch_destroy(b);
	for (uint32_t i = 0; i < nreq; i++)
		free_request(&rs[i], req[i].restarts);
	if (stream)
		hipStreamDestroy(stream);
	nmfk_arena_destroy(arena);
	if (!e && stats) {
		qsort(lat, finished, sizeof(double), cmp_double);
		uint64_t jobs = 0;
		for (uint32_t i = 0; i < nreq; i++)
			jobs += req[i].restarts;
		memset(stats, 0, sizeof *stats);
		stats->seconds = t1;
		stats->requests = finished;
		stats->requests_per_s = t1 > 0 ? finished / t1 : 0.0;
		stats->factorizations_per_s = t1 > 0 ? (double)jobs / t1 : 0.0;
		stats->p50 = quantile(lat, finished, 0.50);
		stats->p99 = quantile(lat, finished, 0.99);
		double s = 0.0;
		for (uint32_t i = 0; i < finished; i++)
			s += lat[i];
		stats->mean_latency = finished ? s / finished : 0.0;
		stats->peak_bytes = peak;
		stats->peak_jobs = peak_jobs;
		stats->step_s = t_step;
		stats->admit_s = t_admit;
		stats->wait_s = t_wait;
		stats->retire_s = t_retire;
		stats->chunks = nchunks;
		stats->pad_ = 0;
	}
	free(rs), free(done), free(dres), free(lat), free(roff);
	return e;
}
