/*
 * serve_sweep.c -- host-memory check of the batch engine: a served request stream (60
 * requests arriving at 20/s, so admissions, retirements, table rebuilds and graph
 * re-captures overlap) and an autotuned batched rank sweep, through the C API.  Built by the
 * Makefile beside it with the library's host C code under AddressSanitizer (the HIP
 * kernels object is the in-tree build, not instrumented).  Needs an MI355X.
 */
#include "nmfk.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static uint64_t st = 12345;
static double urand(void) { st = st * 6364136223846793005ULL + 1442695040888963407ULL; return (double)(st >> 11) / 9007199254740992.0; }

int main(void)
{
	enum { NR = 60 };
	nmfk_request req[NR];
	double t = 0;
	for (int i = 0; i < NR; i++) {
		t += -log(1.0 - urand()) / 20.0;
		req[i].n = (uint32_t)exp(log(1000) + urand() * (log(20000) - log(1000)));
		req[i].m = (uint32_t)exp(log(20) + urand() * (log(3000) - log(20)));
		uint32_t kmax = req[i].m < 32 ? req[i].m : 32;
		req[i].k = 2 + (uint32_t)(urand() * (kmax - 1));
		if (req[i].k > kmax) req[i].k = kmax;
		req[i].restarts = 1 + (uint32_t)(urand() * 10);
		req[i].seed = (uint64_t)(urand() * 1e15);
		req[i].arrival = t;
	}
	nmfk_serve_config cfg;
	nmfk_serve_default_config(&cfg);
	cfg.progress_every = 10;
	nmfk_request_result out[NR];
	nmfk_serve_stats stats;
	int e = nmfk_serve_run(&cfg, req, NR, out, &stats);
	printf("serve rc=%d %s requests=%u req/s=%.2f p50=%.2f p99=%.2f\n", e, e ? nmfk_last_error() : "", stats.requests,
	       stats.requests_per_s, stats.p50, stats.p99);
	fflush(stdout);
	if (e)
		return 1;
	nmfk_job *src = nmfk_job_create(2000, 200, 1);
	if (!src || nmfk_job_gen_V_planted(src, 77, 6, 0.3))
		return 2;
	uint32_t ks[] = {2, 3, 5, 9, 16, 24, 32};
	nmfk_sweep_config sc;
	nmfk_sweep_default_config(&sc);
	sc.ks = ks, sc.nk = 7, sc.restarts = 6, sc.base_seed = 5, sc.autotune = 1, sc.autotune_iters = 3;
	nmfk_result res[7 * 6];
	int32_t *cons = malloc(sizeof(int32_t) * 7 * 200 * 200);
	nmfk_sweep_out so = {0};
	so.res = res, so.consensus = cons;
	e = nmfk_sweep_run(src, &sc, &so);
	printf("sweep rc=%d %s seconds=%.3f\n", e, e ? nmfk_last_error() : "", so.seconds);
	free(cons);
	nmfk_job_destroy(src);
	return e ? 3 : 0;
}
