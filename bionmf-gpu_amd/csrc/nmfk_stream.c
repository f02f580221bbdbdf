/*
 * nmfk_stream.c -- streamed (out-of-core) jobs: V stays in host memory.
 *
 * The reference handles an input larger than device memory by processing V in blocks of
 * rows and of columns that it copies to the device as each half-step needs them
 * (src/GPU_setup.c:902-1270, src/NMF_routines.c:543-1000), i.e. V crosses the bus twice
 * per iteration.  Here it crosses once, in row blocks through two device slots, and the
 * updates keep the reference's order: with H(i+1) known, one pass over the blocks
 *   - updates the block's rows of W:  W(i+1) = W(i) . ((V / (W(i) H(i+1))) H(i+1)) / colsum(H(i+1)),
 *     which is row-local and needs only the block (transposed on the device) and H(i+1);
 *   - adds the block's share of the next H half-step, the sum over its rows of
 *     W(i+1)^T (V / (W(i+1) H(i+1))), which needs the same block and its new rows of W.
 * After the pass, colsum(W(i+1)) and the blocks' partial sums (added in block order) finish
 * H(i+2).  The first pass only adds the partial sums of H(1) against the initial W.
 *
 * At the end of every test interval the reference clamps W and H to R_MIN before testing
 * (src/NMF_GPU.c:611,631): update_W of that iteration used H before the clamp, and the next
 * update_H uses the clamped H.  Such passes therefore contract the W rows against H and the
 * H partial sums against a clamped copy of it; the test kernel then clamps H in place and
 * the W finish clamps the new rows, as in the resident loop.
 *
 * Block starts are multiples of NMFK_FIN_ROWS, so the column-sum partials of W that the
 * per-block finish kernels write are the resident loop's, in the same places.
 */
#include "nmfk_internal.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { PASS_H0, PASS_WH, PASS_W, PASS_DOT };

static size_t rup(size_t x, size_t a)
{
	return (x + a - 1) / a * a;
}

static int salloc(void **p, size_t bytes, const char *what)
{
	*p = NULL;
	return nmfk_hip_fail(hipMalloc(p, bytes ? bytes : 4), what);
}

void *nmfk_host_alloc(size_t bytes)
{
	void *p = NULL;
	if (nmfk_hip_fail(hipHostMalloc(&p, bytes ? bytes : 4, hipHostMallocDefault), "hipHostMalloc"))
		return NULL;
	return p;
}

void nmfk_host_free(void *p)
{
	if (p)
		hipHostFree(p);
}

void nmfk_stream_free(nmfk_job *j)
{
	if (j->cstream)
		hipStreamSynchronize(j->cstream);
	for (int s = 0; s < 2; s++) {
		hipFree(j->Vb[s]);
		hipFree(j->Vtb[s]);
		if (j->stage[s])
			hipHostFree(j->stage[s]);
		if (j->up[s])
			hipEventDestroy(j->up[s]);
		if (j->done[s])
			hipEventDestroy(j->done[s]);
	}
	hipFree(j->partH);
	hipFree(j->Hc);
	hipFree(j->den_scratch);
	if (j->cstream)
		hipStreamDestroy(j->cstream);
}

/* Kernel paths of a streamed job: FP32 throughout. */
static int stream_path(uint32_t kt)
{
	return kt % 16 == 0 ? NMFK_PATH_MFMA32 : NMFK_PATH_VALU;
}

/* Sizes of a streamed job's buffers, shared by create() and nmfk_streamed_bytes(). */
typedef struct slayout {
	uint32_t kt, br, nblocks, rlast, SH, ncolW, ncolH, ndot;
	size_t ldv, ldvtb, part_cap, bytes;
	nmfk_hplan bH[2], bW[2];
} slayout;

static void stream_layout(uint32_t n, uint32_t m, uint32_t k, uint32_t block_rows, slayout *L)
{
	L->kt = nmfk_job_kt(n, m, k);
	L->ldv = rup(m, 4);
	uint32_t br = block_rows;
	if (br == 0) {
		/* two slots of a block and its transpose in 1 GiB */
		const size_t per_row = 2 * 4 * (L->ldv + m);
		br = (uint32_t)(((size_t)1 << 30) / per_row / NMFK_FIN_ROWS * NMFK_FIN_ROWS);
		if (br < NMFK_FIN_ROWS)
			br = NMFK_FIN_ROWS;
	}
	if (br > rup(n, NMFK_FIN_ROWS))
		br = (uint32_t)rup(n, NMFK_FIN_ROWS);
	L->br = br;
	L->nblocks = (n + br - 1) / br;
	L->rlast = n - (L->nblocks - 1) * br;
	L->ldvtb = rup(br, 4);
	const int path = stream_path(L->kt);
	nmfk_plan_half_path(m, br, L->kt, path, &L->bH[0]);
	nmfk_plan_half_path(m, L->rlast, L->kt, path, &L->bH[1]);
	nmfk_plan_half_path(br, m, L->kt, path, &L->bW[0]);
	nmfk_plan_half_path(L->rlast, m, L->kt, path, &L->bW[1]);
	for (int i = 0; i < 2; i++) {
		nmfk_plan_set_rank(&L->bH[i], L->kt, k);
		nmfk_plan_set_rank(&L->bW[i], L->kt, k);
	}
	L->SH = (L->nblocks - 1) * L->bH[0].S + L->bH[1].S;
	const size_t pw0 = (size_t)L->bW[0].S * br, pw1 = (size_t)L->bW[1].S * L->rlast;
	L->part_cap = (pw0 > pw1 ? pw0 : pw1) * L->kt;
	L->ncolW = (n + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS;
	L->ncolH = (m + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS;
	L->ndot = nmfk_dot_kl_blocks(br, m);
	const size_t kt = L->kt;
	L->bytes = 4 * ((size_t)n * kt + 2 * (size_t)m * kt + L->part_cap + (size_t)L->SH * m * kt +
			((size_t)L->ncolW + L->ncolH + 3) * kt + 2 * (size_t)m) +
		   8 * 3 * (size_t)L->ndot + sizeof(nmfk_state) +
		   2 * 4 * ((size_t)br * L->ldv + (size_t)m * L->ldvtb) + 20 * 256;
}

size_t nmfk_streamed_bytes(uint32_t n, uint32_t m, uint32_t k, uint32_t block_rows)
{
	if (n < 1 || m < 1 || k < 1 || k > NMFK_KMAX || block_rows % NMFK_FIN_ROWS != 0)
		return 0;
	slayout L;
	stream_layout(n, m, k, block_rows, &L);
	return L.bytes;
}

nmfk_job *nmfk_job_create_streamed(uint32_t n, uint32_t m, uint32_t k, const float *V, size_t ldv,
				   uint32_t block_rows, int pinned)
{
	if (n < 1 || m < 1 || k < 1 || k > NMFK_KMAX || k > n || k > m) {
		nmfk_set_error("nmfk_job_create_streamed: invalid shape n=%u m=%u k=%u (1 <= k <= min(n, m, %d))", n, m,
			       k, NMFK_KMAX);
		return NULL;
	}
	if (!V || ldv < m) {
		nmfk_set_error("nmfk_job_create_streamed: need V and ldv >= m=%u", m);
		return NULL;
	}
	if (block_rows % NMFK_FIN_ROWS != 0) {
		nmfk_set_error("nmfk_job_create_streamed: block_rows=%u is not a multiple of %d", block_rows,
			       NMFK_FIN_ROWS);
		return NULL;
	}
	nmfk_job *j = (nmfk_job *)calloc(1, sizeof(nmfk_job));
	if (!j) {
		nmfk_set_error("nmfk_job_create_streamed: out of host memory");
		return NULL;
	}
	slayout L;
	stream_layout(n, m, k, block_rows, &L);
	j->n = n, j->m = m, j->k = k, j->kt = L.kt;
	j->ldv = L.ldv;
	j->streamed = 1;
	j->hV = V, j->hldv = ldv, j->hV_pinned = pinned != 0;
	j->brows = L.br, j->nblocks = L.nblocks, j->ldvtb = L.ldvtb;
	const uint32_t br = L.br;
	for (int i = 0; i < 2; i++)
		j->bplanH[i] = L.bH[i], j->bplanW[i] = L.bW[i];
	j->planH = j->bplanH[0];
	j->planW = j->bplanW[0];
	j->SH = L.SH;
	j->part_cap = L.part_cap;
	j->ncolW = L.ncolW, j->ncolH = L.ncolH, j->ndot = L.ndot;
	const size_t kt = j->kt;
	int e = 0;
	e |= salloc((void **)&j->W, 4 * (size_t)n * kt, "hipMalloc(W)");
	if (!e) e |= salloc((void **)&j->H, 4 * (size_t)m * kt, "hipMalloc(H)");
	if (!e) e |= salloc((void **)&j->Hc, 4 * (size_t)m * kt, "hipMalloc(H copy)");
	if (!e) e |= salloc((void **)&j->part, 4 * j->part_cap, "hipMalloc(partials)");
	if (!e) e |= salloc((void **)&j->partH, 4 * (size_t)j->SH * m * kt, "hipMalloc(H partials)");
	if (!e) e |= salloc((void **)&j->colW, 4 * (size_t)j->ncolW * kt, "hipMalloc(colW)");
	if (!e) e |= salloc((void **)&j->colH, 4 * (size_t)j->ncolH * kt, "hipMalloc(colH)");
	if (!e) e |= salloc((void **)&j->denW, 4 * 2 * kt, "hipMalloc(den)");
	if (!e) j->denH = j->denW + kt;
	if (!e) e |= salloc((void **)&j->den_scratch, 4 * kt, "hipMalloc(den scratch)");
	if (!e) e |= salloc((void **)&j->dot, 8 * 3 * (size_t)j->ndot, "hipMalloc(dot)");
	if (!e) e |= salloc((void **)&j->cls, 4 * (size_t)m, "hipMalloc(cls)");
	if (!e) e |= salloc((void **)&j->last, 4 * (size_t)m, "hipMalloc(last)");
	if (!e) e |= salloc((void **)&j->st, sizeof(nmfk_state), "hipMalloc(state)");
	for (int s = 0; s < 2 && !e; s++) {
		e |= salloc((void **)&j->Vb[s], 4 * (size_t)br * j->ldv, "hipMalloc(V block)");
		if (!e) e |= salloc((void **)&j->Vtb[s], 4 * (size_t)m * j->ldvtb, "hipMalloc(V block transposed)");
		if (!e && !j->hV_pinned)
			e |= nmfk_hip_fail(hipHostMalloc((void **)&j->stage[s], 4 * (size_t)br * j->ldv, hipHostMallocDefault),
					   "hipHostMalloc(staging)");
		if (!e) e |= nmfk_hip_fail(hipEventCreateWithFlags(&j->up[s], hipEventDisableTiming), "hipEventCreate");
		if (!e) e |= nmfk_hip_fail(hipEventCreateWithFlags(&j->done[s], hipEventDisableTiming), "hipEventCreate");
	}
	if (!e) e |= nmfk_hip_fail(hipHostMalloc((void **)&j->h_st, 2 * sizeof(nmfk_state), hipHostMallocDefault),
				   "hipHostMalloc(state)");
	if (!e) e |= nmfk_hip_fail(hipStreamCreateWithFlags(&j->stream, hipStreamNonBlocking), "hipStreamCreate");
	if (!e) e |= nmfk_hip_fail(hipStreamCreateWithFlags(&j->cstream, hipStreamNonBlocking), "hipStreamCreate");
	if (!e) e |= nmfk_hip_fail(hipEventCreateWithFlags(&j->ev[0], hipEventDisableTiming), "hipEventCreate");
	if (!e) e |= nmfk_hip_fail(hipEventCreateWithFlags(&j->ev[1], hipEventDisableTiming), "hipEventCreate");
	if (!e) {
		/* padding columns of W and H start (and stay) at zero, and so do the slots' */
		e |= hipMemsetAsync(j->W, 0, 4 * (size_t)n * kt, j->stream);
		e |= hipMemsetAsync(j->H, 0, 4 * (size_t)m * kt, j->stream);
		e |= hipMemsetAsync(j->Hc, 0, 4 * (size_t)m * kt, j->stream);
		e |= hipMemsetAsync(j->st, 0, sizeof(nmfk_state), j->stream);
		for (int s = 0; s < 2; s++) {
			e |= hipMemsetAsync(j->Vb[s], 0, 4 * (size_t)br * j->ldv, j->stream);
			e |= hipMemsetAsync(j->Vtb[s], 0, 4 * (size_t)m * j->ldvtb, j->stream);
		}
		if (!e)
			e = (int)hipStreamSynchronize(j->stream);
		e = nmfk_hip_fail(e, "streamed job set-up");
	}
	if (e) {
		char msg[512];
		snprintf(msg, sizeof msg, "%s", nmfk_last_error());
		nmfk_job_destroy(j);
		nmfk_set_error("%s", msg);
		return NULL;
	}
	return j;
}

int nmfk_job_streamed_info(const nmfk_job *j, uint32_t *block_rows, uint32_t *nblocks, double *bytes_copied)
{
	if (!j->streamed) {
		nmfk_set_error("nmfk_job_streamed_info: not a streamed job");
		return NMFK_EINVAL;
	}
	if (block_rows)
		*block_rows = j->brows;
	if (nblocks)
		*nblocks = j->nblocks;
	if (bytes_copied)
		*bytes_copied = j->stream_bytes;
	return 0;
}

/* Rows [r0, r1) of a block copied into the staging layout (pitch ldv). */
typedef struct stage_part {
	float *dst;
	const float *src;
	size_t ldv, hldv, row, r0, r1;
} stage_part;

static void *stage_rows(void *arg)
{
	const stage_part *t = (const stage_part *)arg;
	for (size_t i = t->r0; i < t->r1; i++)
		memcpy(t->dst + i * t->ldv, t->src + i * t->hldv, t->row);
	return NULL;
}

/* A block's staging copy, split over up to 8 host threads: one thread's memcpy runs at
 * about half the rate the bus takes. */
static void stage_block(float *dst, const float *src, size_t ldv, size_t hldv, size_t row, size_t rows)
{
	enum { MAXT = 8 };
	const size_t bytes = rows * row;
	size_t nt = bytes / ((size_t)4 << 20);   /* at least 4 MB per thread */
	if (nt > MAXT)
		nt = MAXT;
	if (nt > rows)
		nt = rows;
	stage_part part[MAXT];
	pthread_t tid[MAXT];
	int started[MAXT] = {0};
	for (size_t t = 0; t < (nt ? nt : 1); t++) {
		part[t].dst = dst, part[t].src = src, part[t].ldv = ldv, part[t].hldv = hldv, part[t].row = row;
		part[t].r0 = rows * t / (nt ? nt : 1);
		part[t].r1 = rows * (t + 1) / (nt ? nt : 1);
	}
	for (size_t t = 1; t < nt; t++)
		started[t] = pthread_create(&tid[t], NULL, stage_rows, &part[t]) == 0;
	stage_rows(&part[0]);
	for (size_t t = 1; t < nt; t++) {
		if (started[t])
			pthread_join(tid[t], NULL);
		else
			stage_rows(&part[t]);   /* no thread: copy its rows here */
	}
}

/* Copies block b of V into slot s on the copy stream and makes the job's stream wait for
 * it.  The slot is reused once block b - 2's kernels are done (done[s]); a staging buffer
 * once block b - 2's copy is done (up[s]). */
static int upload(nmfk_job *j, uint32_t b, int s, uint32_t *i0, uint32_t *r)
{
	*i0 = b * j->brows;
	*r = j->n - *i0 < j->brows ? j->n - *i0 : j->brows;
	const float *src = j->hV + (size_t)*i0 * j->hldv;
	const size_t row = 4 * (size_t)j->m;
	int e = hipStreamWaitEvent(j->cstream, j->done[s], 0);
	if (!e && j->hV_pinned) {
		if (j->hldv == j->m && j->ldv == j->m)   /* unpadded rows: one contiguous copy */
			e = hipMemcpyAsync(j->Vb[s], src, (size_t)*r * row, hipMemcpyHostToDevice, j->cstream);
		else
			e = hipMemcpy2DAsync(j->Vb[s], j->ldv * 4, src, j->hldv * 4, row, *r, hipMemcpyHostToDevice,
					     j->cstream);
	} else if (!e) {
		e = hipEventSynchronize(j->up[s]);
		if (!e) {
			float *dst = j->stage[s];
			stage_block(dst, src, j->ldv, j->hldv, row, *r);
			/* rows in the slot's layout; the last row without its padding */
			const size_t bytes = ((size_t)(*r - 1) * j->ldv + j->m) * 4;
			e = hipMemcpyAsync(j->Vb[s], dst, bytes, hipMemcpyHostToDevice, j->cstream);
		}
	}
	if (!e)
		e = hipEventRecord(j->up[s], j->cstream);
	if (!e)
		e = hipStreamWaitEvent(j->stream, j->up[s], 0);
	j->stream_bytes += (double)*r * (double)row;
	return e;
}

/* One pass over the blocks of V.  PASS_H0: the H partial sums against the current W.
 * PASS_WH: update the block's rows of W (clamped if <clamp>), then the H partial sums
 * against the new rows and <Hp>.  PASS_W: the W update only.  PASS_DOT: the Frobenius / KL
 * partials of every block, summed into acc[3] in block order.  st: the stop state the
 * kernels skip their work on (null: none). */
static int pass(nmfk_job *j, int kind, const float *Hp, int clamp, double acc[3], const nmfk_state *st)
{
	uint32_t hoff = 0;
	double *h = NULL;
	if (kind == PASS_DOT) {
		h = (double *)malloc(sizeof(double) * 3 * j->ndot);
		if (!h) {
			nmfk_set_error("out of host memory");
			return NMFK_ENOMEM;
		}
	}
	/* after a caller's own upload into slot 0 (the drop-in driver's first block, on the
	 * job's stream), copies into the slots wait for it */
	int e = 0;
	if (j->slot0_written) {
		e = hipEventRecord(j->ev[0], j->stream);
		if (!e)
			e = hipStreamWaitEvent(j->cstream, j->ev[0], 0);
		j->slot0_written = 0;
	}
	for (uint32_t b = 0; b < j->nblocks && !e; b++) {
		const int s = (int)(b & 1u), lastb = b + 1 == j->nblocks;
		uint32_t i0, r;
		e = upload(j, b, s, &i0, &r);
		const nmfk_hplan *ph = &j->bplanH[lastb], *pw = &j->bplanW[lastb];
		float *Wb = j->W + (size_t)i0 * j->kt;
		if (!e && (kind == PASS_WH || kind == PASS_W)) {
			e = nmfk_launch_transpose(j->Vb[s], r, j->m, j->ldv, j->Vtb[s], j->ldvtb, j->stream);
			if (!e)
				e = nmfk_launch_partial(j->Vtb[s], j->ldvtb, Wb, j->H, r, j->m, j->kt, pw, j->part, j->colH,
							j->ncolH, j->denH, st, j->stream);
			if (!e)
				e = nmfk_launch_finish(j->part, pw->S, Wb, r, j->k, j->kt, j->denH,
						       j->colW + (size_t)(i0 / NMFK_FIN_ROWS) * j->kt, clamp, st, j->stream);
		}
		if (!e && (kind == PASS_H0 || kind == PASS_WH)) {
			/* the partial kernel's denominator tile reduces no parts into a scratch buffer:
			 * colsum(W) is complete only after the pass */
			e = nmfk_launch_partial(j->Vb[s], j->ldv, Hp, Wb, j->m, r, j->kt, ph,
						j->partH + (size_t)hoff * j->m * j->kt, j->colW, 0, j->den_scratch, st,
						j->stream);
			hoff += ph->S;
		}
		if (!e && kind == PASS_DOT) {
			e = nmfk_launch_dot_kl(j->Vb[s], j->ldv, Wb, j->H, r, j->m, j->k, j->kt, j->dot, j->stream);
			if (!e)
				e = hipMemcpyAsync(h, j->dot, sizeof(double) * 3 * j->ndot, hipMemcpyDeviceToHost, j->stream);
			if (!e)
				e = hipStreamSynchronize(j->stream);
			for (uint32_t q = 0; q < j->ndot && !e; q++)
				for (int c = 0; c < 3; c++)
					acc[c] += h[3 * (size_t)q + c];
		}
		if (!e)
			e = hipEventRecord(j->done[s], j->stream);
	}
	free(h);
	return e ? nmfk_hip_fail(e, "streamed pass") : 0;
}

int nmfk_stream_run(nmfk_job *j, const nmfk_options *opt, nmfk_result *res)
{
	const uint32_t T = opt->niter_test_conv;
	j->stream_bytes = 0.0;
	int e = nmfk_launch_state_reset(j->st, opt->stop_threshold, UINT32_MAX, j->last, j->m, j->stream);
	if (!e)
		e = nmfk_launch_colsum_parts(j->W, j->n, j->kt, j->colW, j->stream);
	if (!e)
		e = nmfk_launch_den(j->colW, j->ncolW, j->kt, j->denW, j->stream);
	if (!e)
		e = hipStreamSynchronize(j->stream);
	if (e)
		return nmfk_hip_fail(e, "streamed run set-up");
	const double t0 = nmfk_now();
	if (opt->nIters > 0) {
		/* iteration 0's H half-step */
		e = pass(j, PASS_H0, j->H, 0, NULL, j->st);
		if (!e)
			e = nmfk_hip_fail(nmfk_launch_finish(j->partH, j->SH, j->H, j->m, j->k, j->kt, j->denW, j->colH, 0,
							     j->st, j->stream),
					  "H finish");
	}
	/* Pass it completes iteration it - 1 (its W half-step) and adds iteration it's H partial sums. */
	for (uint32_t it = 1; it <= opt->nIters && !e; it++) {
		const int test = it % T == 0, more = it < opt->nIters;
		const float *Hp = j->H;
		if (test && more) {
			e = hipMemcpyAsync(j->Hc, j->H, 4 * (size_t)j->m * j->kt, hipMemcpyDeviceToDevice, j->stream);
			if (!e)
				e = nmfk_launch_adjust(j->Hc, j->m, j->k, j->kt, j->stream);
			if (e) {
				e = nmfk_hip_fail(e, "clamped H");
				break;
			}
			Hp = j->Hc;
		}
		e = pass(j, more ? PASS_WH : PASS_W, Hp, test, NULL, j->st);
		if (!e && test)
			e = nmfk_hip_fail(nmfk_launch_test(j->H, j->m, j->k, j->kt, j->cls, j->last, j->st, j->stream),
					  "test");
		if (!e && more) {
			e = nmfk_launch_den(j->colW, j->ncolW, j->kt, j->denW, j->stream);
			if (!e)
				e = nmfk_launch_finish(j->partH, j->SH, j->H, j->m, j->k, j->kt, j->denW, j->colH, 0, j->st,
						       j->stream);
			e = nmfk_hip_fail(e, "H finish");
		}
		if (!e && test) {
			e = hipMemcpyAsync(&j->h_st[0], j->st, sizeof(nmfk_state), hipMemcpyDeviceToHost, j->stream);
			if (!e)
				e = hipStreamSynchronize(j->stream);
			if (e)
				e = nmfk_hip_fail(e, "test state");
			else {
				nmfk_trace_state(j, &j->h_st[0]);
				if (j->h_st[0].converged)
					break;
			}
		}
	}
	if (e)
		return e;
	e = hipMemcpyAsync(&j->h_st[0], j->st, sizeof(nmfk_state), hipMemcpyDeviceToHost, j->stream);
	if (!e)
		e = hipStreamSynchronize(j->stream);
	if (e)
		return nmfk_hip_fail(e, "streamed run");
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

int nmfk_stream_dot_kl(nmfk_job *j, double *dot_V, double *dot_VWH, double *kl)
{
	double acc[3] = {0.0, 0.0, 0.0};
	const double run_bytes = j->stream_bytes;   /* reported for runs only */
	int e = pass(j, PASS_DOT, NULL, 0, acc, NULL);
	j->stream_bytes = run_bytes;
	if (e)
		return e;
	if (dot_V)
		*dot_V = acc[0];
	if (dot_VWH)
		*dot_VWH = acc[1];
	if (kl)
		*kl = acc[2];
	return 0;
}

/* update_H() of a streamed job (the reference's blockwise half-step): colsum(W), one pass
 * over V adding the H partial sums, then the H finish.  Not gated by the stop state. */
int nmfk_stream_update_H(nmfk_job *j)
{
	int e = nmfk_launch_colsum_parts(j->W, j->n, j->kt, j->colW, j->stream);
	if (!e)
		e = nmfk_launch_den(j->colW, j->ncolW, j->kt, j->denW, j->stream);
	if (e)
		return nmfk_hip_fail(e, "update_H");
	e = pass(j, PASS_H0, j->H, 0, NULL, NULL);
	if (!e)
		e = nmfk_hip_fail(nmfk_launch_finish(j->partH, j->SH, j->H, j->m, j->k, j->kt, j->denW, j->colH, 0, NULL,
						     j->stream),
				  "update_H");
	return e;
}

/* update_W() of a streamed job: colsum(H), then one pass over V updating the rows of W
 * block by block. */
int nmfk_stream_update_W(nmfk_job *j)
{
	int e = nmfk_launch_colsum_parts(j->H, j->m, j->kt, j->colH, j->stream);
	if (e)
		return nmfk_hip_fail(e, "update_W");
	return pass(j, PASS_W, NULL, 0, NULL, NULL);
}
