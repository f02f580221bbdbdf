/*
 * nmf_oracle.c -- CPU oracle for the KL multiplicative-update loop of NMF-mGPU.
 * TEST INFRASTRUCTURE ONLY (see nmf_oracle.h for scope and citations).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no -ffast-math).
 */
#include "nmf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon et al., SC'11, section 3.2).                        */
/* ------------------------------------------------------------------------ */

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
	uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
	uint32_t k0 = key[0], k1 = key[1];
	for (int r = 0; r < 10; r++) {
		uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
		uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
		uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
		uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
		uint32_t n0 = hi1 ^ c1 ^ k0;
		uint32_t n2 = hi0 ^ c3 ^ k1;
		c0 = n0;
		c1 = lo1;
		c2 = n2;
		c3 = lo0;
		k0 += PHILOX_W0;
		k1 += PHILOX_W1;
	}
	out[0] = c0;
	out[1] = c1;
	out[2] = c2;
	out[3] = c3;
}

/* Domain word placed in the last counter lane so these streams never collide with
 * other uses of the same key. */
#define ORC_DOMAIN 0x4E4D4621u

static void orc_block(uint64_t seed, uint32_t tag, uint64_t q, uint32_t out[4])
{
	uint32_t ctr[4] = {(uint32_t)q, (uint32_t)(q >> 32), tag, ORC_DOMAIN};
	uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
	orc_philox4x32_10(ctr, key, out);
}

/* 24 random bits -> (0, 1]: ((x >> 8) + 1) * 2^-24, exact in FP32. */
float orc_uniform01(uint64_t seed, uint32_t tag, uint64_t idx)
{
	uint32_t out[4];
	orc_block(seed, tag, idx >> 2, out);
	uint32_t x = out[idx & 3u];
	return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
}

void orc_uniform_fill(float *A, uint32_t height, uint32_t width, uint32_t pitch, uint64_t seed, uint32_t tag)
{
	for (uint32_t i = 0; i < height; i++)
		for (uint32_t j = 0; j < pitch; j++)
			A[(size_t)i * pitch + j] =
				(j < width) ? orc_uniform01(seed, tag, (uint64_t)i * width + j) : 0.0f;
}

/* 53-bit uniform in (0,1] from two words. */
static double orc_u53(uint32_t hi, uint32_t lo)
{
	uint64_t b = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6); /* 27 + 26 bits */
	return ((double)b + 1.0) * (1.0 / 9007199254740992.0);
}

double orc_normal(uint64_t seed, uint32_t tag, uint64_t idx)
{
	uint32_t out[4];
	orc_block(seed, tag, idx, out);
	double u1 = orc_u53(out[0], out[1]);
	double u2 = orc_u53(out[2], out[3]);
	return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

/* Stream tags of the synthetic-data generators. */
#define TAG_NOISE 0x100u
#define TAG_WSTAR 0x101u
#define TAG_HSTAR 0x102u

void orc_gen_lognormal(float *V, uint32_t n, uint32_t m, uint32_t pitch, uint64_t seed, double mu, double sigma)
{
	for (uint32_t i = 0; i < n; i++)
		for (uint32_t j = 0; j < pitch; j++)
			V[(size_t)i * pitch + j] =
				(j < m) ? (float)exp(mu + sigma * orc_normal(seed, TAG_NOISE, (uint64_t)i * m + j)) : 0.0f;
}

static double orc_exp1(uint64_t seed, uint32_t tag, uint64_t idx)
{
	uint32_t out[4];
	orc_block(seed, tag, idx, out);
	return -log(orc_u53(out[0], out[1]));
}

void orc_gen_planted(float *V, uint32_t n, uint32_t m, uint32_t pitch, uint32_t kstar, uint64_t seed, double sigma)
{
	double *Hs = (double *)malloc(sizeof(double) * (size_t)kstar * m);
	double *Wi = (double *)malloc(sizeof(double) * kstar);
	for (uint32_t c = 0; c < kstar; c++)
		for (uint32_t j = 0; j < m; j++)
			Hs[(size_t)c * m + j] = orc_exp1(seed, TAG_HSTAR, (uint64_t)c * m + j);
	for (uint32_t i = 0; i < n; i++) {
		for (uint32_t c = 0; c < kstar; c++)
			Wi[c] = orc_exp1(seed, TAG_WSTAR, (uint64_t)i * kstar + c);
		for (uint32_t j = 0; j < pitch; j++) {
			if (j >= m) {
				V[(size_t)i * pitch + j] = 0.0f;
				continue;
			}
			double s = 0.0;
			for (uint32_t c = 0; c < kstar; c++)
				s += Wi[c] * Hs[(size_t)c * m + j];
			double z = orc_normal(seed, TAG_NOISE, (uint64_t)i * m + j);
			V[(size_t)i * pitch + j] = (float)(s * exp(sigma * z));
		}
	}
	free(Wi);
	free(Hs);
}

/* ------------------------------------------------------------------------ */
/* Integer parts.                                                            */
/* ------------------------------------------------------------------------ */

/* src/common.c:1195-1209: pairs (i<j) whose "same class" relation changed. */
size_t orc_get_difference(const uint32_t *cls, const uint32_t *last, uint32_t m)
{
	size_t diff = 0;
	for (uint32_t i = 0; i + 1 < m; i++)
		for (uint32_t j = i + 1; j < m; j++) {
			int conn = (cls[j] == cls[i]);
			int conn_last = (last[j] == last[i]);
			diff += (size_t)(conn != conn_last);
		}
	return diff;
}

/* src/NMF_GPU.c:695-708. */
uint32_t orc_stop_rule(const size_t *diffs, uint32_t ntests, uint32_t stop_threshold)
{
	uint32_t inc = 0;
	for (uint32_t t = 0; t < ntests; t++) {
		if (diffs[t])
			inc = 0;
		else if (inc < stop_threshold)
			inc++;
		else
			return t + 1;
	}
	return 0;
}

/* ------------------------------------------------------------------------ */
/* Per-precision routines.                                                   */
/* ------------------------------------------------------------------------ */

/* Both forms clamp at FLT_MIN, the R_MIN of the default FP32 build
 * (include/real_type.h:120-126), so that they run the same algorithm. */
#define ORC_KMAX 256

#define REAL double
#define SFX f64
#define ORC_RMIN FLT_MIN
#include "nmf_oracle_impl.h"
#undef ORC_RMIN
#undef SFX
#undef REAL

#define REAL float
#define SFX f32
#define ORC_RMIN FLT_MIN
#include "nmf_oracle_impl.h"
#undef ORC_RMIN
#undef SFX
#undef REAL
