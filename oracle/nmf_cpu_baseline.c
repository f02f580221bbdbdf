/*
 * nmf_cpu_baseline.c -- times the oracle's scalar C rendering of the reference update
 * sequence (src/NMF_GPU.c:522-795 driving src/NMF_routines.c update_H/update_W) on the
 * host cores.  TEST INFRASTRUCTURE: reported beside the GPU numbers, never the product.
 *
 * Usage:
 *   nmf_cpu_baseline --n N --m M --k K [--iters I] [--T t] [--S s]
 *                    [--gen lognormal|planted] [--kstar r] [--sigma x] [--mu x]
 *                    [--seed-v s] [--seed-init s] [--threads p] [--fp64]
 * Prints one JSON line.  Only the iteration loop is timed (CLOCK_MONOTONIC); V generation
 * and W0/H0 initialisation are excluded.
 */
#include "nmf_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char **argv)
{
	uint32_t n = 5000, m = 38, k = 3, iters = 2000, T = 10, S = 40, kstar = 6, threads = 1;
	double sigma = 1.5, mu = 5.0;
	uint64_t seed_v = 1, seed_init = 2;
	int planted = 0, fp64 = 0;
	for (int a = 1; a < argc; a++) {
		const char *s = argv[a];
		const char *v = (a + 1 < argc) ? argv[a + 1] : "0";
		if (!strcmp(s, "--n")) n = (uint32_t)strtoul(v, 0, 10), a++;
		else if (!strcmp(s, "--m")) m = (uint32_t)strtoul(v, 0, 10), a++;
		else if (!strcmp(s, "--k")) k = (uint32_t)strtoul(v, 0, 10), a++;
		else if (!strcmp(s, "--iters")) iters = (uint32_t)strtoul(v, 0, 10), a++;
		else if (!strcmp(s, "--T")) T = (uint32_t)strtoul(v, 0, 10), a++;
		else if (!strcmp(s, "--S")) S = (uint32_t)strtoul(v, 0, 10), a++;
		else if (!strcmp(s, "--kstar")) kstar = (uint32_t)strtoul(v, 0, 10), a++;
		else if (!strcmp(s, "--threads")) threads = (uint32_t)strtoul(v, 0, 10), a++;
		else if (!strcmp(s, "--sigma")) sigma = strtod(v, 0), a++;
		else if (!strcmp(s, "--mu")) mu = strtod(v, 0), a++;
		else if (!strcmp(s, "--seed-v")) seed_v = strtoull(v, 0, 10), a++;
		else if (!strcmp(s, "--seed-init")) seed_init = strtoull(v, 0, 10), a++;
		else if (!strcmp(s, "--gen")) planted = !strcmp(v, "planted"), a++;
		else if (!strcmp(s, "--fp64")) fp64 = 1;
		else {
			fprintf(stderr, "unknown option %s\n", s);
			return 2;
		}
	}
#ifdef _OPENMP
	omp_set_num_threads((int)threads);
#else
	threads = 1;
#endif
	float *V = (float *)malloc(sizeof(float) * (size_t)n * m);
	float *W = (float *)malloc(sizeof(float) * (size_t)n * k);
	float *H = (float *)malloc(sizeof(float) * (size_t)m * k);
	uint32_t *cls = (uint32_t *)malloc(sizeof(uint32_t) * m);
	if (!V || !W || !H || !cls)
		return 3;
	if (planted)
		orc_gen_planted(V, n, m, m, kstar, seed_v, sigma);
	else
		orc_gen_lognormal(V, n, m, m, seed_v, mu, sigma);
	orc_uniform_fill(W, n, k, k, seed_init, 0);
	orc_uniform_fill(H, m, k, k, seed_init, 1);

	uint32_t it = 0, ntests = 0;
	int conv = 0;
	double t0, t1, kl;
	if (fp64) {
		size_t nv = (size_t)n * m, nw = (size_t)n * k, nh = (size_t)m * k;
		double *Vd = (double *)malloc(sizeof(double) * nv);
		double *Wd = (double *)malloc(sizeof(double) * nw);
		double *Hd = (double *)malloc(sizeof(double) * nh);
		for (size_t q = 0; q < nv; q++) Vd[q] = V[q];
		for (size_t q = 0; q < nw; q++) Wd[q] = W[q];
		for (size_t q = 0; q < nh; q++) Hd[q] = H[q];
		t0 = now_s();
		orc_nmf_f64(Vd, n, m, m, Wd, Hd, k, k, iters, T, S, cls, &it, &conv, 0, 0, 0, &ntests);
		t1 = now_s();
		kl = orc_kl_div_f64(Vd, n, m, m, Wd, Hd, k, k);
		free(Vd), free(Wd), free(Hd);
	} else {
		t0 = now_s();
		orc_nmf_f32(V, n, m, m, W, H, k, k, iters, T, S, cls, &it, &conv, 0, 0, 0, &ntests);
		t1 = now_s();
		kl = orc_kl_div_f32(V, n, m, m, W, H, k, k);
	}
	uint32_t done = conv ? it : iters;
	printf("{\"n\": %u, \"m\": %u, \"k\": %u, \"iters_requested\": %u, \"iters_done\": %u, "
	       "\"converged\": %d, \"tests\": %u, \"seconds\": %.6f, \"s_per_iter\": %.9f, "
	       "\"threads\": %u, \"precision\": \"%s\", \"kl\": %.10e}\n",
	       n, m, k, iters, done, conv, ntests, t1 - t0, (t1 - t0) / (done ? done : 1), threads,
	       fp64 ? "f64" : "f32", kl);
	free(V), free(W), free(H), free(cls);
	return 0;
}
