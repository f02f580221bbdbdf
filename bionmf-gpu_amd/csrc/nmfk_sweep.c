/*
 * nmfk_sweep.c -- rank sweep x random restarts over one input matrix, with consensus
 * matrices.
 *
 * The reference factorizes one (V, k, seed) per run of bin/NMF_GPU (src/NMF_GPU.c:1034-1269);
 * choosing k by the stability of the class vectors over restarts means running it many
 * times.  Here every (k, restart) pair is a job reading the same device-resident V
 * (nmfk_job_create_shared), all jobs advance together through the batched launches
 * (nmfk_batch.c), and for every k the m x m consensus matrix
 *   C_k[i][j] = #{restarts r : class_r(i) == class_r(j)}
 * is accumulated exactly in int32 on the device from each job's final class vector.
 * The one-at-a-time mode runs the identical jobs (same seeds, same kernel plans) with
 * nmfk_job_run, one after the other: the plain path the batched form is measured against.
 */
#include "nmfk_internal.h"

#include <stdlib.h>
#include <string.h>

uint64_t nmfk_sweep_seed(uint64_t base, uint32_t k, uint32_t restart)
{
	/* splitmix64 finaliser (Steele, Lea, Flood 2014) of base ^ (k << 32 | restart) */
	uint64_t z = base ^ (((uint64_t)k << 32) | restart);
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

void nmfk_sweep_default_config(nmfk_sweep_config *c)
{
	memset(c, 0, sizeof *c);
	c->restarts = 1;
	c->nIters = 2000;
	c->niter_test_conv = 10;
	c->stop_threshold = 40;
	c->mode = NMFK_SWEEP_BATCHED;
	c->autotune_iters = 20;
}

static int sweep_fail(nmfk_job **jobs, uint32_t nj, nmfk_batch *b, int32_t *dC, int e)
{
	if (b)
		nmfk_batch_destroy(b);
	for (uint32_t i = 0; i < nj; i++)
		nmfk_job_destroy(jobs[i]);
	free(jobs);
	hipFree(dC);
	return e ? e : NMFK_EINVAL;
}

int nmfk_sweep_run(nmfk_job *src, const nmfk_sweep_config *cfg, nmfk_sweep_out *out)
{
	if (!src || !cfg || !cfg->ks || cfg->nk == 0 || cfg->restarts == 0 || cfg->niter_test_conv == 0) {
		nmfk_set_error("nmfk_sweep_run: need a source job, ranks, restarts >= 1 and niter_test_conv >= 1");
		return NMFK_EINVAL;
	}
	if (cfg->mode == NMFK_SWEEP_BATCHED && cfg->nIters < cfg->niter_test_conv) {
		nmfk_set_error("nmfk_sweep_run: batched mode needs nIters >= niter_test_conv");
		return NMFK_EINVAL;
	}
	const uint32_t R = cfg->restarts, nj = cfg->nk * R, m = src->m;
	nmfk_job **jobs = (nmfk_job **)calloc(nj, sizeof(nmfk_job *));
	if (!jobs) {
		nmfk_set_error("out of host memory");
		return NMFK_ENOMEM;
	}
	int32_t *dC = NULL;
	nmfk_batch *b = NULL;
	int e = 0;
	for (uint32_t q = 0; q < cfg->nk && !e; q++)
		for (uint32_t r = 0; r < R && !e; r++) {
			const uint32_t i = q * R + r;
			jobs[i] = nmfk_job_create_shared(src, cfg->ks[q]);
			if (!jobs[i]) {
				e = NMFK_EINVAL;
				break;
			}
			if (cfg->plan_target)
				e = nmfk_job_replan(jobs[i], cfg->plan_target);
			if (!e)
				e = nmfk_job_init_random(jobs[i], nmfk_sweep_seed(cfg->base_seed, cfg->ks[q], r));
		}
	if (e)
		return sweep_fail(jobs, nj, b, dC, e);

	nmfk_options opt;
	opt.nIters = cfg->nIters;
	opt.niter_test_conv = cfg->niter_test_conv;
	opt.stop_threshold = cfg->stop_threshold;
	double tune_s = 0.0;
	const double t0 = nmfk_now();
	if (out)
		out->plan_target = cfg->plan_target;
	if (cfg->mode == NMFK_SWEEP_ONE_AT_A_TIME) {
		for (uint32_t i = 0; i < nj && !e; i++) {
			opt.k = jobs[i]->k;
			nmfk_result r;
			e = nmfk_job_run(jobs[i], &opt, &r);
			if (!e && out && out->res)
				out->res[i] = r;
		}
	} else {
		nmfk_batch_tuning best_tn;
		nmfk_batch_default_tuning(&best_tn);
		uint32_t best_target = cfg->plan_target;
		if (cfg->autotune && cfg->plan_target == 0) {
			/* Plan search: every candidate per-job plan (workgroups each job's partial
			 * launch aims for; 0 = the process tunables) is tried on a trial batch of all
			 * jobs, whose launch shape the batch autotuner picks; the fastest pair wins. */
			const double ta = nmfk_now();
			static const uint32_t cands[] = {8, 16, 32, 64, 0};
			double best_ms = 1e30;
			for (size_t c = 0; c < sizeof cands / sizeof cands[0] && !e; c++) {
				for (uint32_t i = 0; i < nj && !e; i++)
					e = nmfk_job_replan(jobs[i], cands[c]);
				nmfk_batch *tb = e ? NULL : nmfk_batch_create(cfg->niter_test_conv, nj);
				if (!e && !tb)
					e = NMFK_EINVAL;
				for (uint32_t i = 0; i < nj && !e; i++) {
					opt.k = jobs[i]->k;
					if (nmfk_batch_add(tb, jobs[i], &opt) < 0)
						e = NMFK_EINVAL;
				}
				nmfk_batch_tuning tn;
				double ms = 0.0;
				if (!e)
					e = nmfk_batch_autotune(tb, cfg->autotune_iters, &tn, &ms);
				if (!e && ms < best_ms) {
					best_ms = ms;
					best_tn = tn;
					best_target = cands[c];
				}
				nmfk_batch_destroy(tb);
			}
			for (uint32_t i = 0; i < nj && !e; i++)
				e = nmfk_job_replan(jobs[i], best_target);
			tune_s = nmfk_now() - ta;
		}
		b = e ? NULL : nmfk_batch_create(cfg->niter_test_conv, nj);
		if (!e && !b)
			e = NMFK_EINVAL;
		for (uint32_t i = 0; i < nj && !e; i++) {
			opt.k = jobs[i]->k;
			if (nmfk_batch_add(b, jobs[i], &opt) < 0)
				e = NMFK_EINVAL;
		}
		if (!e && cfg->autotune) {
			if (cfg->plan_target == 0) {
				e = nmfk_batch_set_tuning(b, &best_tn);
			} else {
				const double ta = nmfk_now();
				e = nmfk_batch_autotune(b, cfg->autotune_iters, &best_tn, NULL);
				tune_s = nmfk_now() - ta;
			}
		}
		if (out)
			out->plan_target = best_target;
		nmfk_batch_tuning tn;
		nmfk_batch_get_tuning(b, &tn);
		nmfk_job **done = (nmfk_job **)malloc(nj * sizeof(nmfk_job *));
		nmfk_result *rs = (nmfk_result *)malloc(nj * sizeof(nmfk_result));
		if (!done || !rs) {
			nmfk_set_error("out of host memory");
			e = NMFK_ENOMEM;
		}
		while (!e && nmfk_batch_active(b) > 0) {
			e = nmfk_batch_step(b, tn.chunks_per_poll);
			uint32_t nd = 0;
			if (!e)
				e = nmfk_batch_poll(b, done, rs, nj, &nd);
			for (uint32_t d = 0; d < nd; d++)
				for (uint32_t i = 0; i < nj; i++)
					if (jobs[i] == done[d]) {
						if (out && out->res)
							out->res[i] = rs[d];
						break;
					}
		}
		free(done);
		free(rs);
		if (!e && out)
			nmfk_batch_get_tuning(b, &out->tuning);
	}
	const double t1 = nmfk_now();
	if (e)
		return sweep_fail(jobs, nj, b, dC, e);
	if (out) {
		out->seconds = t1 - t0 - tune_s;
		out->tune_seconds = tune_s;
	}

	/* Final class vectors (argmax of the final H) and the consensus matrices. */
	const size_t mm = (size_t)m * m;
	hipStream_t s = src->stream;
	e = hipMalloc((void **)&dC, cfg->nk * mm * sizeof(int32_t));
	if (!e)
		e = hipMemsetAsync(dC, 0, cfg->nk * mm * sizeof(int32_t), s);
	for (uint32_t q = 0; q < cfg->nk && !e; q++)
		for (uint32_t r = 0; r < R && !e; r++) {
			nmfk_job *j = jobs[q * R + r];
			e = nmfk_job_sync(j);
			if (!e)
				e = nmfk_launch_idx_max(j->H, j->m, j->k, j->kt, j->cls, s);
			if (!e)
				e = nmfk_launch_consensus_add(j->cls, m, dC + q * mm, s);
			if (!e && out && out->cls)
				e = hipMemcpyAsync(out->cls + (size_t)(q * R + r) * m, j->cls, 4 * (size_t)m,
						   hipMemcpyDeviceToHost, s);
		}
	if (!e && out && out->consensus)
		e = hipMemcpyAsync(out->consensus, dC, cfg->nk * mm * sizeof(int32_t), hipMemcpyDeviceToHost, s);
	if (!e)
		e = hipStreamSynchronize(s);
	if (e) {
		nmfk_hip_fail(e, "sweep consensus");
		return sweep_fail(jobs, nj, b, dC, NMFK_EHIP);
	}
	if (out && out->W_H_out) {
		/* optional per-job factor export, k-major (row-major n x k then m x k) */
		size_t o = 0;
		for (uint32_t i = 0; i < nj && !e; i++) {
			e = nmfk_job_get_W(jobs[i], out->W_H_out + o, jobs[i]->k);
			o += (size_t)jobs[i]->n * jobs[i]->k;
			if (!e)
				e = nmfk_job_get_H(jobs[i], out->W_H_out + o, jobs[i]->k);
			o += (size_t)jobs[i]->m * jobs[i]->k;
		}
	}
	if (b)
		nmfk_batch_destroy(b);
	for (uint32_t i = 0; i < nj; i++)
		nmfk_job_destroy(jobs[i]);
	free(jobs);
	hipFree(dC);
	return e;
}
