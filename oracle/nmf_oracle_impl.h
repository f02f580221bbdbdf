/*
 * nmf_oracle_impl.h -- per-precision body of the CPU oracle (TEST
 * INFRASTRUCTURE ONLY; see nmf_oracle.h).  Included twice by nmf_oracle.c with
 * REAL/SFX set to (double, f64) and (float, f32).
 *
 * All sums run in the natural sequential order over their index, in REAL,
 * with no fused multiply-add (the oracle is compiled with -ffp-contract=off).
 */

#define ORC_CAT_(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT_(a, b)
#define ORC_FN(name) ORC_CAT(name, SFX)

/* acc[c] = sum_i A[i][c] -- matrix_to_row / kernel_reduce_to_row
 * (src/matrix_operations.c:744, src/GPU_kernels.cu:870-1066). */
void ORC_FN(orc_colsum)(const REAL *A, uint32_t height, uint32_t width, uint32_t pitch, REAL *acc)
{
	for (uint32_t c = 0; c < width; c++)
		acc[c] = (REAL)0;
	for (uint32_t i = 0; i < height; i++)
		for (uint32_t c = 0; c < width; c++)
			acc[c] += A[(size_t)i * pitch + c];
}

/* WH[i][j] = sum_c W[i][c] * H[j][c]  (the first GEMM, src/NMF_routines.c:400,1038). */
static inline REAL ORC_FN(orc_wh)(const REAL *Wi, const REAL *Hj, uint32_t k)
{
	REAL s = (REAL)0;
	for (uint32_t c = 0; c < k; c++)
		s += Wi[c] * Hj[c];
	return s;
}

/*
 * H half-step, update_H() (src/NMF_routines.c:704-1000):
 *   accum_W = colsum(W)                         :720
 *   WH      = W * H^T                           :400
 *   Aux     = V ./ WH                           :436  (kernel: A = B / A, GPU_kernels.cu:1690)
 *   Haux    = W^T * Aux                         :464
 *   H       = (H .* Haux) ./ accum_W            :498  (kernel: mul first, then div, GPU_kernels.cu:2029-2033)
 * Rows of V are streamed in order, so every Haux[j][c] is summed over i ascending.
 */
void ORC_FN(orc_update_H)(const REAL *V, uint32_t n, uint32_t m, uint32_t ldv,
			  const REAL *W, REAL *H, uint32_t k, uint32_t ldk)
{
	REAL *accW = (REAL *)calloc(k, sizeof(REAL));
	REAL *Haux = (REAL *)calloc((size_t)m * k, sizeof(REAL));
	ORC_FN(orc_colsum)(W, n, k, ldk, accW);
	/* Column blocks are independent and each keeps the ascending-i order, so the result
	 * does not depend on the number of OpenMP threads. */
	const long nblk = ((long)m + 7) / 8;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
	for (long b = 0; b < nblk; b++) {
		uint32_t j0 = (uint32_t)b * 8u, j1 = j0 + 8u < m ? j0 + 8u : m;
		for (uint32_t i = 0; i < n; i++) {
			const REAL *Wi = W + (size_t)i * ldk;
			const REAL *Vi = V + (size_t)i * ldv;
			for (uint32_t j = j0; j < j1; j++) {
				REAL wh = ORC_FN(orc_wh)(Wi, H + (size_t)j * ldk, k);
				REAL a = Vi[j] / wh;
				REAL *hx = Haux + (size_t)j * k;
				for (uint32_t c = 0; c < k; c++)
					hx[c] += Wi[c] * a;
			}
		}
	}
	for (uint32_t j = 0; j < m; j++)
		for (uint32_t c = 0; c < k; c++) {
			REAL t = H[(size_t)j * ldk + c] * Haux[(size_t)j * k + c];
			H[(size_t)j * ldk + c] = t / accW[c];
		}
	free(Haux);
	free(accW);
}

/*
 * W half-step, update_W() (src/NMF_routines.c:1344-1636), using the NEW H:
 *   accum_H = colsum(H)                         :1360
 *   WH      = W * H^T                           :1038
 *   Aux     = V ./ WH                           :1074
 *   Waux    = Aux * H                           :1102
 *   W       = (W .* Waux) ./ accum_H            :1137
 * Waux[i][c] is summed over j ascending.  k <= ORC_KMAX.
 */
void ORC_FN(orc_update_W)(const REAL *V, uint32_t n, uint32_t m, uint32_t ldv,
			  REAL *W, const REAL *H, uint32_t k, uint32_t ldk)
{
	REAL *accH = (REAL *)calloc(k, sizeof(REAL));
	ORC_FN(orc_colsum)(H, m, k, ldk, accH);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
	for (long il = 0; il < (long)n; il++) {
		uint32_t i = (uint32_t)il;
		REAL Waux[ORC_KMAX];
		REAL *Wi = W + (size_t)i * ldk;
		const REAL *Vi = V + (size_t)i * ldv;
		for (uint32_t c = 0; c < k; c++)
			Waux[c] = (REAL)0;
		for (uint32_t j = 0; j < m; j++) {
			const REAL *Hj = H + (size_t)j * ldk;
			REAL wh = ORC_FN(orc_wh)(Wi, Hj, k);
			REAL a = Vi[j] / wh;
			for (uint32_t c = 0; c < k; c++)
				Waux[c] += a * Hj[c];
		}
		for (uint32_t c = 0; c < k; c++) {
			REAL t = Wi[c] * Waux[c];
			Wi[c] = t / accH[c];
		}
	}
	free(accH);
}

/* x = (isnan(x) || x < R_MIN) ? R_MIN : x  -- kernel_adjust (src/GPU_kernels.cu:2338-2345).
 * R_MIN = FLT_MIN in both forms (the default FP32 build's value, include/real_type.h:120-126). */
void ORC_FN(orc_adjust)(REAL *A, uint32_t height, uint32_t width, uint32_t pitch)
{
	const REAL rmin = ORC_RMIN;
	for (uint32_t i = 0; i < height; i++)
		for (uint32_t c = 0; c < width; c++) {
			REAL x = A[(size_t)i * pitch + c];
			if (x != x || x < rmin)
				A[(size_t)i * pitch + c] = rmin;
		}
}

/*
 * idx[i] = index of the row maximum -- kernel_idx_max (src/GPU_kernels.cu:2508-3267).
 * Emulates the reference thread layout exactly, NaN cases included:
 *   block width bw = 32 when the reference pitch (K rounded up to 32) is <= 96,
 *   else next_pow2(ceil(pitch/3))            (src/matrix_operations.c:190-203);
 *   thread t (t < bw) scans columns t, t+bw, t+2bw with a strict '>' and starts from its
 *   first column, or from (-1.0, 0) when it has none (src/GPU_kernels.cu:2603-2640);
 *   a tree over halves keeps the lower thread unless 'lower < upper'
 *   (src/GPU_kernels.cu:2761-2830).
 * Consequence: among tied maxima the column with the smallest (col mod bw) wins, then the
 * smallest column; for k <= 32 that is the lowest index.
 */
void ORC_FN(orc_idx_max)(const REAL *A, uint32_t height, uint32_t width, uint32_t pitch, uint32_t *idx)
{
	uint32_t kp_ref = (width + 31u) & ~31u;
	uint32_t bw = 32;
	if (kp_ref > 96u) {
		uint32_t need = (kp_ref + 2u) / 3u;
		bw = 1;
		while (bw < need)
			bw <<= 1;
	}
	REAL *tv = (REAL *)malloc(sizeof(REAL) * bw);
	uint32_t *ti = (uint32_t *)malloc(sizeof(uint32_t) * bw);
	for (uint32_t r = 0; r < height; r++) {
		const REAL *row = A + (size_t)r * pitch;
		for (uint32_t t = 0; t < bw; t++) {
			if (t < width) {
				REAL mv = row[t];
				uint32_t mi = t;
				for (uint32_t c = t + bw; c < width; c += bw)
					if (row[c] > mv) {
						mv = row[c];
						mi = c;
					}
				tv[t] = mv;
				ti[t] = mi;
			} else {
				tv[t] = (REAL)-1.0;
				ti[t] = 0;
			}
		}
		for (uint32_t half = bw >> 1; half >= 1; half >>= 1)
			for (uint32_t t = 0; t < half; t++)
				if (tv[t] < tv[t + half]) {
					tv[t] = tv[t + half];
					ti[t] = ti[t + half];
				}
		idx[r] = ti[0];
	}
	free(ti);
	free(tv);
}

/* A[i][c] = (A[i][c] * Aux[i][c]) / acc[c] -- kernel_mul_div (src/GPU_kernels.cu:2015-2040). */
void ORC_FN(orc_mul_div)(REAL *A, const REAL *Aux, const REAL *acc, uint32_t height,
			 uint32_t width, uint32_t pitch)
{
	for (uint32_t i = 0; i < height; i++)
		for (uint32_t c = 0; c < width; c++) {
			size_t p = (size_t)i * pitch + c;
			REAL t = A[p] * Aux[p];
			A[p] = t / acc[c];
		}
}

/* A = B ./ A  or  A = B - A -- kernel_div_sub (src/GPU_kernels.cu:1676-1700). */
void ORC_FN(orc_div_sub)(REAL *A, const REAL *B, uint32_t height, uint32_t width,
			 uint32_t pitch, int div_operator)
{
	for (uint32_t i = 0; i < height; i++)
		for (uint32_t c = 0; c < width; c++) {
			size_t p = (size_t)i * pitch + c;
			A[p] = div_operator ? (B[p] / A[p]) : (B[p] - A[p]);
		}
}

/* dot_V = <V,V>, dot_VWH = <V-WH, V-WH> -- dot_product_VWH (src/NMF_routines.c:1758-2794).
 * WH and V-WH are formed in REAL as in the reference; the squares are accumulated in FP64
 * (the reference accumulates per-row cublasSdot results in FP32). */
void ORC_FN(orc_dot_VWH)(const REAL *V, uint32_t n, uint32_t m, uint32_t ldv,
			 const REAL *W, const REAL *H, uint32_t k, uint32_t ldk,
			 double *dot_V, double *dot_VWH)
{
	double sv = 0.0, sd = 0.0;
	for (uint32_t i = 0; i < n; i++)
		for (uint32_t j = 0; j < m; j++) {
			REAL v = V[(size_t)i * ldv + j];
			REAL wh = ORC_FN(orc_wh)(W + (size_t)i * ldk, H + (size_t)j * ldk, k);
			REAL d = v - wh;
			sv += (double)v * (double)v;
			sd += (double)d * (double)d;
		}
	*dot_V = sv;
	*dot_VWH = sd;
}

/* Generalized KL divergence D(V||WH) = sum V*log(V/WH) - V + WH, accumulated in FP64.
 * Not computed by the reference (it reports only the Frobenius ratio); it is the objective
 * the multiplicative updates decrease. 0*log(0/x) = 0. */
double ORC_FN(orc_kl_div)(const REAL *V, uint32_t n, uint32_t m, uint32_t ldv,
			  const REAL *W, const REAL *H, uint32_t k, uint32_t ldk)
{
	double s = 0.0;
	for (uint32_t i = 0; i < n; i++)
		for (uint32_t j = 0; j < m; j++) {
			double v = (double)V[(size_t)i * ldv + j];
			double wh = 0.0;
			for (uint32_t c = 0; c < k; c++)
				wh += (double)W[(size_t)i * ldk + c] * (double)H[(size_t)j * ldk + c];
			s += (v > 0.0 ? v * log(v / wh) : 0.0) - v + wh;
		}
	return s;
}

/*
 * The nmf() driver loop (src/NMF_GPU.c:522-795):
 *   nIters/T blocks of T iterations { update_H; update_W }, then adjust(H), adjust(W),
 *   classification, connectivity difference against the previous vector (initially all
 *   zeros, src/NMF_GPU.c:1188), and the stop rule (stop after stop_threshold+1 unchanged
 *   tests).  If not converged, nIters % T more iterations with no adjust and no test.
 * Outputs: W, H (in place), the class vector = argmax of the final H, the iteration count
 * the reference prints (tests*T when converged, else nIters), and optional per-test traces.
 */
int ORC_FN(orc_nmf)(const REAL *V, uint32_t n, uint32_t m, uint32_t ldv,
		    REAL *W, REAL *H, uint32_t k, uint32_t ldk,
		    uint32_t nIters, uint32_t T, uint32_t S,
		    uint32_t *cls_out, uint32_t *iters_out, int *converged_out,
		    uint32_t *cls_trace, size_t *diff_trace, uint32_t max_trace,
		    uint32_t *ntests_out)
{
	if (T == 0 || k == 0 || n == 0 || m == 0)
		return 1;
	uint32_t *cls = (uint32_t *)calloc(m, sizeof(uint32_t));
	uint32_t *last = (uint32_t *)calloc(m, sizeof(uint32_t));
	uint32_t niter_conv = nIters / T, niter_rem = nIters % T;
	uint32_t inc = 0, iter = 0, ntests = 0;
	int converged = 0;

	for (; iter < niter_conv; iter++) {
		for (uint32_t i = 0; i < T; i++) {
			ORC_FN(orc_update_H)(V, n, m, ldv, W, H, k, ldk);
			ORC_FN(orc_update_W)(V, n, m, ldv, W, H, k, ldk);
		}
		ORC_FN(orc_adjust)(H, m, k, ldk);
		ORC_FN(orc_adjust)(W, n, k, ldk);
		ORC_FN(orc_idx_max)(H, m, k, ldk, cls);
		size_t diff = orc_get_difference(cls, last, m);
		if (ntests < max_trace) {
			if (cls_trace)
				memcpy(cls_trace + (size_t)ntests * m, cls, sizeof(uint32_t) * m);
			if (diff_trace)
				diff_trace[ntests] = diff;
		}
		ntests++;
		uint32_t *tmp = cls;
		cls = last;
		last = tmp;
		if (diff)
			inc = 0;
		else if (inc < S)
			inc++;
		else {
			iter++;
			converged = 1;
			break;
		}
	}
	if (!converged)
		for (uint32_t i = 0; i < niter_rem; i++) {
			ORC_FN(orc_update_H)(V, n, m, ldv, W, H, k, ldk);
			ORC_FN(orc_update_W)(V, n, m, ldv, W, H, k, ldk);
		}
	if (cls_out)
		ORC_FN(orc_idx_max)(H, m, k, ldk, cls_out);
	if (iters_out)
		*iters_out = converged ? iter * T : nIters;
	if (converged_out)
		*converged_out = converged;
	if (ntests_out)
		*ntests_out = ntests;
	free(cls);
	free(last);
	return 0;
}

#undef ORC_FN
#undef ORC_CAT
#undef ORC_CAT_
