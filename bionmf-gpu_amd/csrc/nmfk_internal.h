/*
 * nmfk_internal.h -- private definitions shared by the C host files of libnmfgpu_amd.
 */
#ifndef NMFK_INTERNAL_H
#define NMFK_INTERNAL_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "nmfk.h"
#include "nmfk_kernels.h"

#define NMFK_EINVAL 2
#define NMFK_EHIP 3
#define NMFK_ENOMEM 4
#define NMFK_EBUDGET 5

/* Device memory arena (nmfk_arena.c): segments reserved once, sub-allocated per job. */
typedef struct nmfk_arena nmfk_arena;
nmfk_arena *nmfk_arena_create(size_t limit, size_t seg_bytes);
void nmfk_arena_destroy(nmfk_arena *a);
void *nmfk_arena_alloc(nmfk_arena *a, size_t bytes);   /* NULL when it does not fit */
int nmfk_arena_free(nmfk_arena *a, void *p);
size_t nmfk_arena_round(size_t bytes);                  /* bytes a block of this size takes */
size_t nmfk_arena_reserved(const nmfk_arena *a);
size_t nmfk_arena_in_use(const nmfk_arena *a);
void nmfk_arena_trim(nmfk_arena *a);                   /* releases segments with nothing in use */

struct nmfk_job {
	uint32_t n, m, k, kt;
	size_t ldv, ldvt;
	float *V, *Vt, *W, *H;         /* device */
	float *part;                   /* partial contractions, S x p x kt */
	size_t part_cap;               /* floats allocated for part */
	float *colW, *colH;            /* column-sum partials, one row per NMFK_FIN_ROWS rows */
	uint32_t ncolW, ncolH;
	float *denW, *denH;            /* colsum(W), colsum(H): the half-steps' denominators */
	double *dot;                   /* FP64 partial sums of the Frobenius / KL kernel */
	uint32_t ndot;
	uint32_t *cls, *last;          /* device class vectors */
	nmfk_state *st;                /* device stop-rule state */
	nmfk_state *h_st;              /* pinned, 2 slots */
	nmfk_hplan planH, planW;
	hipStream_t stream;
	hipEvent_t ev[2];
	hipGraphExec_t chunk_exec;
	uint32_t chunk_T;
	int chunk_sched;               /* schedule the captured interval graph runs */
	/* One pass over V per iteration (nmfk_job_set_schedule): requested schedule, and the
	 * pass's per-workgroup sums (fG x m x kt) and column sums of W (fG x kt), allocated on
	 * the first run that uses it. */
	int sched;
	float *fpart, *fcolW;
	uint32_t fG, fnwv;
	int colW_valid;
	int vt_stale;                  /* V changed since Vt was built */
	int owns_V;                    /* 0: V and Vt belong to another job (nmfk_job_create_shared) */
	/* BF16 copies of W and H for the BF16 kernel path (kt % 32 == 0 jobs only), row-major
	 * (rows x kt) like the factors. */
	uint16_t *Wb, *Hb;
	/* Streamed jobs (nmfk_job_create_streamed): V stays in host memory and passes through
	 * two device slots one row block at a time; V and Vt above stay NULL. */
	int streamed;
	const float *hV;               /* host V, row-major, leading dimension hldv */
	size_t hldv;
	int hV_pinned;                 /* hV is pinned (nmfk_host_alloc): copied without staging */
	uint32_t brows, nblocks;       /* rows per block (multiple of NMFK_FIN_ROWS), blocks */
	float *Vb[2], *Vtb[2];         /* device slots: a row block (ld ldv) and its transpose */
	size_t ldvtb;
	float *stage[2];               /* pinned staging buffers for a pageable hV */
	nmfk_hplan bplanH[2], bplanW[2]; /* plans of a full block [0] and of the last block [1] */
	float *partH;                  /* the H half's partial sums of all blocks, in block order */
	uint32_t SH;                   /* slices in partH */
	float *Hc;                     /* clamped copy of H that test passes contract against */
	float *den_scratch;            /* where the block partials' denominator tile writes */
	hipStream_t cstream;           /* host-to-device copies */
	hipEvent_t up[2], done[2];     /* slot s: block uploaded / block's kernels finished */
	double stream_bytes;           /* bytes of V copied to the device by the last run */
	int slot0_written;             /* a caller wrote slot 0 on the job's stream (drop-in upload) */
	/* Per-test connectivity differences of the current run (nmfk_job_run_traced): the host
	 * copies the device state after every test interval and keeps its last_diff. */
	nmfk_arena *arena;             /* buffers come from (and return to) this arena; NULL: hipMalloc */
	int borrowed_stream;           /* the stream belongs to the job's creator */
	int tracing;
	uint64_t *diff_trace;
	uint32_t trace_cap, trace_len;
};

/* Records the difference of the test a state copy reports, once per test. */
static inline void nmfk_trace_state(nmfk_job *j, const nmfk_state *s)
{
	if (j->tracing && s->ntests == j->trace_len + 1) {
		if (j->trace_len < j->trace_cap)
			j->diff_trace[j->trace_len] = s->last_diff;
		j->trace_len++;
	}
}

/* A job whose device buffers come from <arena> and whose work runs on the caller's <stream>
 * (nmfk_job_create, or nmfk_job_create_shared when src is given, without any hipMalloc,
 * hipFree or stream creation; nmfk_job_destroy returns the buffers to the arena). */
nmfk_job *nmfk_job_create_in(nmfk_arena *arena, hipStream_t stream, uint32_t n, uint32_t m, uint32_t k,
			     nmfk_job *src);
/* nmfk_job_bytes without V and Vt: a job sharing another job's V. */
size_t nmfk_job_bytes_shared(uint32_t n, uint32_t m, uint32_t k);

/* Streamed jobs: the loop and the Frobenius/KL pass (nmfk_stream.c). */
int nmfk_stream_run(nmfk_job *j, const nmfk_options *opt, nmfk_result *res);
int nmfk_stream_dot_kl(nmfk_job *j, double *dot_V, double *dot_VWH, double *kl);
int nmfk_stream_update_H(nmfk_job *j);
int nmfk_stream_update_W(nmfk_job *j);
void nmfk_stream_free(nmfk_job *j);

/* Refreshes the BF16 copy of the factor the next half-step reads as B, if its plan is BF16. */
int nmfk_job_shadow_for_H(nmfk_job *j);
int nmfk_job_shadow_for_W(nmfk_job *j);

/* Rebuilds Vt from V if V was written since (asynchronous on the job's stream). */
int nmfk_job_refresh_Vt(nmfk_job *j);

/* The factorization the drop-in globals (GPU_setup.h, NMF_routines.h) refer to. */
extern nmfk_job *nmfk_dropin_job;
extern uint64_t nmfk_dropin_seed;
extern uint32_t nmfk_dropin_rng_calls;
extern int nmfk_dropin_rng_ready;

void nmfk_set_error(const char *fmt, ...);
int nmfk_hip_fail(int err, const char *what);
double nmfk_now(void);

int nmfk_enqueue_half_H(nmfk_job *j, const nmfk_state *st);
int nmfk_enqueue_half_W(nmfk_job *j, const nmfk_state *st, int clamp);

#endif /* NMFK_INTERNAL_H */
