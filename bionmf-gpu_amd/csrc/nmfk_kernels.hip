// nmfk_kernels.hip -- hand-written gfx950 kernels for the KL multiplicative-update loop
// of NMF-mGPU (update_H / update_W, src/NMF_routines.c:704-1000 and :1344-1636).
//
// Data layout in HBM (all FP32, row-major):
//   V  n x ldv, and its transpose Vt  m x ldvt, both resident for the whole run, so that
//   each half-step streams its operand with fully coalesced rows;
//   W  n x kt, H  m x kt (H transposed, as in the reference), kt = padded rank.
//
// A half-step "updates factor A (p x kt) against factor B (q x kt) through U (q x p)":
//   H half: U = V,  A = H (p = m), B = W (q = n)     (src/NMF_routines.c:376-525)
//   W half: U = Vt, A = W (p = n), B = H (q = m)     (src/NMF_routines.c:1014-1164)
//   A[a][c] <- A[a][c] * (sum_b B[b][c] U[b][a] / (A[a] . B[b])) / (sum_b B[b][c])
// The reference materialises WH and Aux (n x m) in HBM between cuBLAS GEMMs and two
// elementwise kernels; here the product, the ratio and the contraction are one streaming
// pass in registers: a lane owns AP consecutive rows a of the factor being updated, the
// other factor's row B[b] is wave-uniform (scalar loads), and U[b][a..] is read once with
// coalesced vector loads.  The b range is split over workgroups; partial sums are reduced
// in a fixed order (no float atomics), so every result is bit-reproducible.

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "nmfk_kernels.h"

#define NMFK_RMIN 1.17549435e-38f   // FLT_MIN, R_MIN of the FP32 build (include/real_type.h:120-126)

// --------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Identical to the host generator of the oracle.
// --------------------------------------------------------------------------------------

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u
#define NMFK_DOMAIN 0x4E4D4621u
#define TAG_NOISE 0x100u
#define TAG_WSTAR 0x101u
#define TAG_HSTAR 0x102u

__device__ __forceinline__ void philox_block(uint64_t seed, uint32_t tag, uint64_t q, uint32_t out[4])
{
	uint32_t c0 = (uint32_t)q, c1 = (uint32_t)(q >> 32), c2 = tag, c3 = NMFK_DOMAIN;
	uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
	for (int r = 0; r < 10; r++) {
		const uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
		const uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
		const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
		const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
		c0 = hi1 ^ c1 ^ k0;
		c1 = lo1;
		c2 = hi0 ^ c3 ^ k1;
		c3 = lo0;
		k0 += PHILOX_W0;
		k1 += PHILOX_W1;
	}
	out[0] = c0;
	out[1] = c1;
	out[2] = c2;
	out[3] = c3;
}

__device__ __forceinline__ float uniform01(uint64_t seed, uint32_t tag, uint64_t idx)
{
	uint32_t o[4];
	philox_block(seed, tag, idx >> 2, o);
	const uint32_t x = o[idx & 3u];
	return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ double u53(uint32_t hi, uint32_t lo)
{
	const uint64_t b = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
	return ((double)b + 1.0) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ double normal01(uint64_t seed, uint32_t tag, uint64_t idx)
{
	uint32_t o[4];
	philox_block(seed, tag, idx, o);
	const double u1 = u53(o[0], o[1]);
	const double u2 = u53(o[2], o[3]);
	return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925286766559 * u2);
}

__device__ __forceinline__ float clamp_rmin(float x)
{
	// (isnan(x) || x < R_MIN) ? R_MIN : x   (src/GPU_kernels.cu:2338-2345)
	return (x >= NMFK_RMIN) ? x : NMFK_RMIN;
}

// Sum over a wave64 with a fixed xor tree; every lane ends with the same bits.
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1)
		v += __shfl_xor(v, off, 64);
	return v;
}

__device__ __forceinline__ double wave_sum_d(double v)
{
#pragma unroll
	for (int off = 32; off >= 1; off >>= 1)
		v += __shfl_xor(v, off, 64);
	return v;
}

// Denominator of a half-step: den[c] = sum over the nparts column-sum partials of the
// other factor.  Four interleaved ascending partial sums per column, combined in a fixed
// order, so every kernel that forms it gets the same bits.  Needs 4*KT <= blockDim.x.
// Workgroup size of every half-step kernel (single-job and batched launches alike).
#define NMFK_WG 256

// den[c] = sum_g parts[g][c]: four interleaved running sums per column, added in a fixed
// order at the end.  s4 is 4*KT floats of LDS; needs 4*KT <= NMFK_WG threads.
template <int KT>
__device__ __forceinline__ void den_reduce(const float *__restrict__ parts, int nparts, float *__restrict__ den,
					   float *__restrict__ s4)
{
	static_assert(4 * KT <= NMFK_WG, "den_reduce: one thread per (column, lane of 4)");
	const int t = threadIdx.x;
	if (t < 4 * KT) {
		const int c = t >> 2, j = t & 3;
		float s = 0.0f;
		int g = j;
		// 8 loads in flight per step; the additions keep ascending g order.
		for (; g + 28 < nparts; g += 32) {
			float v[8];
#pragma unroll
			for (int u = 0; u < 8; u++)
				v[u] = parts[(size_t)(g + 4 * u) * KT + c];
#pragma unroll
			for (int u = 0; u < 8; u++)
				s += v[u];
		}
		for (; g < nparts; g += 4)
			s += parts[(size_t)g * KT + c];
		s4[t] = s;
	}
	__syncthreads();
	if (t < KT)
		den[t] = ((s4[4 * t] + s4[4 * t + 1]) + s4[4 * t + 2]) + s4[4 * t + 3];
}

// One finish block = NMFK_FIN_ROWS (64) rows of A with 4 threads per row, each thread
// KT/4 consecutive columns (NMFK_WG threads).  Column sums of the block's rows (ascending
// rows) are written as one partial row; shared by the finish and column-sum kernels so
// both give the same bits.
__host__ __device__ constexpr int fin_tpr(int)
{
	return 4;
}
__host__ __device__ constexpr int fin_q(int kt)
{
	return kt / fin_tpr(kt);
}
static_assert(NMFK_FIN_ROWS * 4 == NMFK_WG, "finish block shape");

// tile: NMFK_FIN_ROWS x (KT + 1) floats of LDS.
template <int KT>
__device__ __forceinline__ void block_colsum_store(const float (&x)[fin_q(KT)], float *__restrict__ out,
						   float *__restrict__ tile)
{
	constexpr int TPR = fin_tpr(KT), Q = fin_q(KT);
	const int row = threadIdx.x / TPR, q = threadIdx.x % TPR;
#pragma unroll
	for (int i = 0; i < Q; i++)
		tile[row * (KT + 1) + q * Q + i] = x[i];
	__syncthreads();
	if (threadIdx.x < KT) {
		float s = 0.0f;
		for (int r = 0; r < NMFK_FIN_ROWS; r++)
			s += tile[r * (KT + 1) + threadIdx.x];
		out[threadIdx.x] = s;
	}
}

// --------------------------------------------------------------------------------------
// Half-step kernels.
// --------------------------------------------------------------------------------------

// Vector load of AP consecutive U values.  Every lane reads in bounds: rows are padded
// to a multiple of 4 (so a0 + AP <= ldu whenever a0 < p) and lanes past p read column 0;
// values past p are then zeroed with selects (no branches in the streaming loop).
template <int AP>
__device__ __forceinline__ void load_u(const float *__restrict__ src, int avail, float (&u)[AP])
{
	if constexpr (AP == 4) {
		const float4 t = *reinterpret_cast<const float4 *>(src);
		u[0] = t.x, u[1] = t.y, u[2] = t.z, u[3] = t.w;
	} else if constexpr (AP == 2) {
		const float2 t = *reinterpret_cast<const float2 *>(src);
		u[0] = t.x, u[1] = t.y;
	} else {
		u[0] = src[0];
	}
#pragma unroll
	for (int r = 0; r < AP; r++)
		u[r] = (r < avail) ? u[r] : 0.0f;
}

// Arguments of one half-step's partial kernel: update A (p x KT) against B (q x KT)
// through U (q x p, leading dimension ldu).
struct half_args {
	const float *U;
	size_t ldu;
	const float *A, *B;
	int p, q, bchunk, nab;
	float *part;
	const float *colB;
	int ncolB;
	float *den;
	int lb;   // VALU lane groups (see partial_valu_body)
};

// LDS floats the partial bodies need (VALU: NW/2 wave slots of AP*KT*64; MFMA: NB-1 wave
// groups of NA*4*KT/16*4*64).
__host__ __device__ constexpr int smem_valu(int kt, int ap, int nw)
{
	return (nw / 2 > 0 ? nw / 2 : 1) * ap * kt * 64;
}
__host__ __device__ constexpr int smem_mfma(int kt, int na, int nb)
{
	return (nb > 1 ? nb - 1 : 1) * na * 4 * (kt / 16) * 4 * 64;
}

// Workgroup (bx, by) of a half-step's partial grid, bx in [0, nab], by in [0, S):
// part[by][a][c] = sum_{b in split by} B[b][c] * U[b][a] / (A[a] . B[b]).
// Column bx == nab is the denominator column: its first workgroup reduces the other
// factor's column-sum partials into the denominators the finish kernel divides by, off
// the critical path of the streaming workgroups; the rest return at once.
// A wave's 64 lanes are LB groups of 64 / LB: lane (la, lb) owns rows a of A
// (AP per lane) and takes every LB-th of the wave's rows of U, so that a factor with few
// rows (p <= 64 AP / LB) still keeps every lane busy.  The groups' sums are added by a
// fixed xor tree at the end, then the waves' by a fixed LDS tree.
template <int KT, int AP, int NW, int LB = 1>
__device__ __forceinline__ void partial_valu_body(int bx, int by, const half_args &h, float *__restrict__ smem)
{
	static_assert(LB == 1 || LB == 2 || LB == 4 || LB == 8, "lane groups: 1, 2, 4 or 8");
	constexpr int LA = 64 / LB;
	const float *__restrict__ U = h.U;
	const size_t ldu = h.ldu;
	const float *__restrict__ A = h.A;
	const float *__restrict__ B = h.B;
	const int p = h.p, q = h.q, bchunk = h.bchunk;
	if (bx == h.nab) {
		if (by == 0)
			den_reduce<KT>(h.colB, h.ncolB, h.den, smem);
		return;
	}
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int la = lane % LA, lg = lane / LA;
	const int a0 = (bx * LA + la) * AP;
	const int avail = p - a0;

	float av[AP][KT], acc[AP][KT];
#pragma unroll
	for (int r = 0; r < AP; r++) {
		const bool ok = r < avail;
		const float *Ar = A + (size_t)(ok ? a0 + r : 0) * KT;
#pragma unroll
		for (int c = 0; c < KT; c++) {
			av[r][c] = ok ? Ar[c] : 0.0f;
			acc[r][c] = 0.0f;
		}
	}

	const int b_begin = by * bchunk;
	const int b_end = min(q, b_begin + bchunk);
	const float *Ucol = U + (avail > 0 ? a0 : 0);
	// The next DR rows of this wave (U values and the B row) are in flight while a row is
	// computed; a consumed slot is refilled with the row DR steps ahead.  Rows are still
	// accumulated in ascending order.  (DR = 2 measured as fast as 4, with fewer registers,
	// and 13% faster than one row ahead at 20000 x 1000, k = 3.)
	constexpr int DR = 2;
	float un[DR][AP], bn[DR][KT];
	auto fetch = [&](int rb, float (&u)[AP], float (&bv)[KT]) {
		if (rb < b_end) {
			load_u<AP>(Ucol + (size_t)rb * ldu, avail, u);
			const float *Bb = B + (size_t)rb * KT;
#pragma unroll
			for (int c = 0; c < KT; c++)
				bv[c] = Bb[c];
		}
	};
	// rows of this lane: b_begin + wave * LB + lg, then every NW * LB-th
	constexpr int STEP = NW * LB;
#pragma unroll
	for (int d = 0; d < DR; d++)
		fetch(b_begin + wave * LB + lg + d * STEP, un[d], bn[d]);
	for (int b = b_begin + wave * LB + lg; b < b_end; b += DR * STEP) {
#pragma unroll
		for (int d = 0; d < DR; d++) {
			const int rb = b + d * STEP;
			if (rb >= b_end)
				break;
			float u[AP], bv[KT];
#pragma unroll
			for (int r = 0; r < AP; r++)
				u[r] = un[d][r];
#pragma unroll
			for (int c = 0; c < KT; c++)
				bv[c] = bn[d][c];
			fetch(rb + DR * STEP, un[d], bn[d]);
#pragma unroll
			for (int r = 0; r < AP; r++) {
				float wh = 0.0f;
#pragma unroll
				for (int c = 0; c < KT; c++)
					wh = fmaf(av[r][c], bv[c], wh);
				// ratio = U / WH with the hardware reciprocal (<= 1 ulp) and one rounding of
				// the product, the fast division the reference's default build uses
				// (include/real_type.h:283-288, --use_fast_math).
				const float ratio = u[r] * __builtin_amdgcn_rcpf(wh);
#pragma unroll
				for (int c = 0; c < KT; c++)
					acc[r][c] = fmaf(ratio, bv[c], acc[r][c]);
			}
		}
	}

	// Fixed-order xor tree over the lane groups: every lane of a group ends with the sum
	// (the two partners of a step add the same two values).
	if constexpr (LB > 1) {
#pragma unroll
		for (int o = 32; o >= LA; o >>= 1)
#pragma unroll
			for (int r = 0; r < AP; r++)
#pragma unroll
				for (int c = 0; c < KT; c++)
					acc[r][c] += __shfl_xor(acc[r][c], o, 64);
	}
	// Fixed-order tree over the NW waves of the workgroup.
	constexpr int E = AP * KT;
	float *__restrict__ red = smem;
#pragma unroll
	for (int half = NW / 2; half >= 1; half >>= 1) {
		if (wave >= half && wave < 2 * half) {
			float *dst = red + (size_t)(wave - half) * E * 64;
#pragma unroll
			for (int r = 0; r < AP; r++)
#pragma unroll
				for (int c = 0; c < KT; c++)
					dst[(r * KT + c) * 64 + lane] = acc[r][c];
		}
		__syncthreads();
		if (wave < half) {
			const float *src = red + (size_t)wave * E * 64;
#pragma unroll
			for (int r = 0; r < AP; r++)
#pragma unroll
				for (int c = 0; c < KT; c++)
					acc[r][c] += src[(r * KT + c) * 64 + lane];
		}
		__syncthreads();
	}
	if (wave == 0 && lg == 0) {
		float *dst = h.part + ((size_t)by * p + a0) * KT;
#pragma unroll
		for (int r = 0; r < AP; r++)
			if (r < avail) {
#pragma unroll
				for (int c = 0; c < KT; c += 4)
					if constexpr (KT % 4 == 0)
						*reinterpret_cast<float4 *>(dst + r * KT + c) =
							make_float4(acc[r][c], acc[r][c + 1], acc[r][c + 2], acc[r][c + 3]);
			}
	}
}

// Single-job partial grids are 1-D: the (nab + 1) x S tiles, rounded up to a multiple of
// the 8 XCDs.  The dispatcher places workgroup L on XCD L % 8; tile T = (L % 8) * per + L / 8
// gives every XCD a contiguous run of tiles in by-major order, so the workgroups that share
// a split of B (same by) mostly share one L2 instead of fetching it on every XCD.  Returns
// false for the padding workgroups.
__device__ __forceinline__ bool xcd_tile(int gx, int total, int &bx, int &by)
{
	const int per = (total + 7) >> 3;
	const int L = blockIdx.x;
	const int T = (L & 7) * per + (L >> 3);
	if (T >= total)
		return false;
	by = T / gx;
	bx = T - by * gx;
	return true;
}

template <int KT, int AP, int NW, int LB>
__global__ __launch_bounds__(64 * NW) void k_partial(half_args h, const nmfk_state *__restrict__ st, int S)
{
	__shared__ float smem[smem_valu(KT, AP, NW)];
	if (st != nullptr && st->converged)
		return;
	int bx, by;
	if (!xcd_tile(h.nab + 1, (h.nab + 1) * S, bx, by))
		return;
	partial_valu_body<KT, AP, NW, LB>(bx, by, h, smem);
}

// --------------------------------------------------------------------------------------
// The same half-step on the matrix cores (KT a multiple of 16), FP32 MFMA 16x16x4.
//
// Wave w of a workgroup owns 64 rows a of the factor being updated, as 4 tiles t of 16
// with row a = a0 + 4*j + t for lane column j = lane % 16 (so a lane's 4 values of one U
// row are one float4); the NW waves own adjacent 64-row groups and walk the same 16-row
// blocks of U in lockstep, so each U row is read as one contiguous 64*NW*4-byte segment
// per workgroup.  Lane group g = lane / 16.  Per block:
//   GEMM1  S[b][a] = sum_c B[b][c] A[a][c]   M = 16 rows b, N = 16 a's of tile t, K = c;
//          the output lands as lane (g, j) holding S[bb + 4g + r][a0 + 4j + t], r = 0..3;
//   ratio  R = U / S in registers (hardware reciprocal, then one product);
//   GEMM2  acc[a][c] += sum_b R[b][a] B[b][c]  M = 16 a's, N = 16 c's, K = b: GEMM1's
//          output register r is used directly as the A operand of the K-chunk made of rows
//          {bb + 4g + r : g = 0..3}; the B operand is B at those same rows.
// In both products the summed index is assigned to lane groups identically on the two
// operands, which is all a contraction needs, so no data moves between them.
// Streaming loads are buffer loads: each lane's byte offset within a block is fixed for the
// whole kernel and the block's row offset is a wave-uniform scalar, so the loop does no
// per-lane address arithmetic.
// --------------------------------------------------------------------------------------

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef int intx4 __attribute__((ext_vector_type(4)));
typedef int intx2 __attribute__((ext_vector_type(2)));

#define NMFK_RSRC_FLAGS 0x00020000   // buffer resource word 3 of the gfx9 family

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const float *p, size_t bytes)
{
	const int n = bytes > 0x7fffffffull ? 0x7fffffff : (int)bytes;
	return __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0, n, NMFK_RSRC_FLAGS);
}

__device__ __forceinline__ float4 bload4(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
	const intx4 v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0);
	return make_float4(__int_as_float(v.x), __int_as_float(v.y), __int_as_float(v.z), __int_as_float(v.w));
}

__device__ __forceinline__ float2 bload2(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
	const intx2 v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
	return make_float2(__int_as_float(v.x), __int_as_float(v.y));
}

__device__ __forceinline__ float bload1(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
	return __int_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// QN consecutive floats (1, 2 or 4) into the first QN entries of a float4.
template <int QN>
__device__ __forceinline__ float4 bloadq(__amdgpu_buffer_rsrc_t r, int voff)
{
	if constexpr (QN == 4)
		return bload4(r, voff, 0);
	else if constexpr (QN == 2) {
		const float2 v = bload2(r, voff, 0);
		return make_float4(v.x, v.y, 0.f, 0.f);
	} else
		return make_float4(bload1(r, voff, 0), 0.f, 0.f, 0.f);
}
template <int QN>
__device__ __forceinline__ float4 loadq(const float *p)
{
	if constexpr (QN == 4)
		return *reinterpret_cast<const float4 *>(p);
	else if constexpr (QN == 2) {
		const float2 v = *reinterpret_cast<const float2 *>(p);
		return make_float4(v.x, v.y, 0.f, 0.f);
	} else
		return make_float4(p[0], 0.f, 0.f, 0.f);
}

// Workgroup = NA x NB waves: NA waves side by side in a (lockstep over the same rows of U),
// NB groups of them taking the split's 16-row blocks round-robin; the NB partial sums of
// each row a are added through LDS in a fixed order.
// KE: columns of the rank GEMM1 sums (KT, or 4 / 8 at KT 16 for a rank up to 4 / 8, whose
// other columns are zero).  Lane group g supplies QN = KE / 4 consecutive columns per 16-column
// group, column 16h + QN g + qq for K-chunk qq < QN, so a rank up to 8 takes two GEMM1 chunks
// instead of four and a rank up to 4 one.
// GEMM2 at KE up to NMFK_G2N_MAX_KE (G2N) runs on the 16-block 4x4x1 FP32 MFMA instead, which
// does the same multiply-adds per cycle with 4 output columns instead of 16: block b = 4g + jj of lane
// (g, j = 4jj + n) takes row 4g + r of the 16-row block, rows a = a0 + 16jj + 4i + t (i < 4) and
// columns c = QN n + cg, so no padding column is computed.  Each lane group g then holds the
// sums over its own rows of every 16-row block; they are added across the four groups once per
// split.  Columns KE..KT-1 of the partials are left unwritten: the finish never reads a column
// at or past the rank.  Only KE 4 takes this: at KE 8 the accumulators double.  The single-job
// kernel would then run at 140 VGPRs (3 waves per SIMD) and 3.5-5% faster per iteration, but
// the batched kt-16 classes would spill 160-200 bytes at their 128-register bound (the batched
// rank sweep slowed by 8%), and a job must give the same bits in a batch as alone, so both
// keep the 16-wide GEMM2 at KE 8.
#define NMFK_G2N_MAX_KE 4
template <int KT, int NA, int NB, int KE = KT>
__device__ __forceinline__ void partial_mfma_body(int bx, int by, const half_args &ha, float *__restrict__ smem)
{
	static_assert(KT % 16 == 0, "MFMA path needs KT % 16 == 0");
	static_assert(KE == KT || (KT == 16 && (KE == 4 || KE == 8)), "GEMM1 width: KT, or 4 / 8 at KT 16");
	constexpr int H = KT / 16;     // 16-column groups of the rank
	constexpr int QN = KE >= 16 ? 4 : KE / 4;
	constexpr bool G2N = KT == 16 && KE <= NMFK_G2N_MAX_KE;
	constexpr int CW = G2N ? QN : H; // GEMM2 B operands per row and accumulators per t
	const float *__restrict__ U = ha.U;
	const size_t ldu = ha.ldu;
	const float *__restrict__ A = ha.A;
	const float *__restrict__ B = ha.B;
	const int p = ha.p, q = ha.q, bchunk = ha.bchunk;
	// Denominator column (see partial_valu_body).
	if (bx == ha.nab) {
		if (by == 0)
			den_reduce<KT>(ha.colB, ha.ncolB, ha.den, smem);
		return;
	}

	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int wa = wave % NA, wb = wave / NA;
	const int j = lane & 15, g = lane >> 4;
	const int a0 = (bx * NA + wa) * 64;
	const bool idle = a0 >= p;               // wave-uniform; idle waves still join the LDS step
	const int aj = a0 + 4 * j;               // first of this lane's 4 rows a (one per tile)
	const int avail = p - aj;                // tiles t < avail are real rows

	// GEMM1 B operand: lane (g, j) supplies A[aj + t][16h + QN g + qq] for K-chunk (h, qq).
	floatx4 aop[4][H];
#pragma unroll
	for (int t = 0; t < 4; t++) {
		const bool ok = t < avail;
		const float *Ar = A + (size_t)(ok ? aj + t : 0) * KT + QN * g;
#pragma unroll
		for (int h = 0; h < H; h++) {
			const float4 v = loadq<QN>(Ar + 16 * h);
			aop[t][h] = ok ? floatx4{v.x, v.y, v.z, v.w} : floatx4{0.f, 0.f, 0.f, 0.f};
		}
	}
	floatx4 acc[4][CW];
#pragma unroll
	for (int t = 0; t < 4; t++)
#pragma unroll
		for (int h = 0; h < CW; h++)
			acc[t][h] = floatx4{0.f, 0.f, 0.f, 0.f};

	const int b_begin = by * bchunk;
	const int b_end = min(q, b_begin + bchunk);

	// Lanes past p read column 0 (in bounds); their rows a only reach MFMA output rows of
	// the same invalid a, which are never stored, so no column mask is needed.
	// The buffer descriptors are based at each 16-row block (wave-uniform scalar work), so
	// byte offsets stay within one block and U and B may exceed the 32-bit range of an offset.
	const int ucol = avail > 0 ? aj : 0;
	const int c2 = G2N ? QN * (j & 3) : j; // first GEMM2 output column of this lane
	int uoff[4], b2off[4][H], b1off[H];
#pragma unroll
	for (int r = 0; r < 4; r++) {
		uoff[r] = ((4 * g + r) * (int)ldu + ucol) * 4;
#pragma unroll
		for (int h = 0; h < H; h++)
			b2off[r][h] = ((4 * g + r) * KT + 16 * h + c2) * 4;
	}
#pragma unroll
	for (int h = 0; h < H; h++)
		b1off[h] = (j * KT + 16 * h + QN * g) * 4;

	// Everything one 16-row block needs:
	//   u[r]      U[bb + 4g + r][aj .. aj+3]
	//   b1[h]     GEMM1 A operand, B[bb + j][16h + QN g + qq], qq < QN
	//   b2[r][h]  GEMM2 B operand, B[bb + 4g + r][16h + j]; with G2N b2[r][cg] = B[bb + 4g + r][c2 + cg]
	struct blk {
		float4 u[4];
		float4 b1[H];
		float b2[4][CW];
	};
	auto load_b2 = [&](blk &x, int r, float4 v) {
		x.b2[r][0] = v.x;
		if constexpr (CW > 1)
			x.b2[r][1] = v.y;
	};
	auto load_full = [&](int bb, blk &x) {
		const __amdgpu_buffer_rsrc_t rU = make_rsrc(U + (size_t)bb * ldu, (size_t)16 * ldu * 4);
		const __amdgpu_buffer_rsrc_t rB = make_rsrc(B + (size_t)bb * KT, (size_t)16 * KT * 4);
#pragma unroll
		for (int r = 0; r < 4; r++) {
			x.u[r] = bload4(rU, uoff[r], 0);
			if constexpr (G2N)
				load_b2(x, r, bloadq<QN>(rB, b2off[r][0]));
			else
#pragma unroll
				for (int h = 0; h < H; h++)
					x.b2[r][h] = bload1(rB, b2off[r][h], 0);
		}
#pragma unroll
		for (int h = 0; h < H; h++)
			x.b1[h] = bloadq<QN>(rB, b1off[h]);
	};
	// Partial last block of the matrix: rows at or past b_end read row b_begin instead.
	auto load_tail = [&](int bb, blk &x) {
#pragma unroll
		for (int r = 0; r < 4; r++) {
			const int b = bb + 4 * g + r;
			const int bs = b < b_end ? b : b_begin;
			x.u[r] = *reinterpret_cast<const float4 *>(U + (size_t)bs * ldu + ucol);
			const float *Br = B + (size_t)bs * KT + c2;
			if constexpr (G2N)
				load_b2(x, r, loadq<QN>(Br));
			else
#pragma unroll
				for (int h = 0; h < H; h++)
					x.b2[r][h] = Br[16 * h];
		}
		const int b = bb + j;
		const float *Br = B + (size_t)(b < b_end ? b : b_begin) * KT + QN * g;
#pragma unroll
		for (int h = 0; h < H; h++)
			x.b1[h] = loadq<QN>(Br + 16 * h);
	};
	// The three stages of one block.  MASK zeroes rows at or past b_end.
	auto gemm1 = [&](const blk &x, floatx4 (&s)[4], bool b1ok) {
#pragma unroll
		for (int h = 0; h < H; h++)
#pragma unroll
			for (int qq = 0; qq < QN; qq++) {
				const float4 b1 = x.b1[h];
				const float bq = qq == 0 ? b1.x : qq == 1 ? b1.y : qq == 2 ? b1.z : b1.w;
				const float bv = b1ok ? bq : 0.0f;
#pragma unroll
				for (int t = 0; t < 4; t++)
					s[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(
						bv, aop[t][h][qq], (h == 0 && qq == 0) ? floatx4{0.f, 0.f, 0.f, 0.f} : s[t],
						0, 0, 0);
			}
	};
	auto ratio = [&](const blk &x, const floatx4 (&s)[4], float (&rt)[4][4], int bb, auto mask_tag) {
		constexpr bool MASK = decltype(mask_tag)::value;
#pragma unroll
		for (int r = 0; r < 4; r++) {
			const bool rok = !MASK || bb + 4 * g + r < b_end;
			const float uv[4] = {x.u[r].x, x.u[r].y, x.u[r].z, x.u[r].w};
#pragma unroll
			for (int t = 0; t < 4; t++) {
				const float v = uv[t] * __builtin_amdgcn_rcpf(s[t][r]);
				rt[t][r] = rok ? v : 0.0f;
			}
		}
	};
	auto gemm2 = [&](const blk &x, const float (&rt)[4][4]) {
		if constexpr (G2N) {
#pragma unroll
			for (int r = 0; r < 4; r++)
#pragma unroll
				for (int cg = 0; cg < CW; cg++)
#pragma unroll
					for (int t = 0; t < 4; t++)
						acc[t][cg] = __builtin_amdgcn_mfma_f32_4x4x1f32(rt[t][r], x.b2[r][cg], acc[t][cg], 0,
												0, 0);
			return;
		}
#pragma unroll
		for (int h = 0; h < H; h++)
#pragma unroll
			for (int r = 0; r < 4; r++)
#pragma unroll
				for (int t = 0; t < 4; t++)
					acc[t][h] = __builtin_amdgcn_mfma_f32_16x16x4f32(rt[t][r], x.b2[r][h], acc[t][h], 0,
											 0, 0);
	};
	auto compute = [&](const blk &x, int bb, auto mask_tag) {
		constexpr bool MASK = decltype(mask_tag)::value;
		floatx4 s[4];
		float rt[4][4];
		gemm1(x, s, !MASK || bb + j < b_end);
		ratio(x, s, rt, bb, mask_tag);
		gemm2(x, rt);
	};

	// Full blocks stream through a D-deep register ring (D blocks in flight, slot indices
	// compile-time after unrolling); a partial last block of the matrix is handled masked.
	// Ring depth 2 up to KT 32: at KT 16 it fits 120 VGPRs (4 waves per SIMD, 9% faster than
	// a 3-deep ring at 3 waves), at KT 32 166 VGPRs with the kernel's 3-wave bound (7% faster
	// than 3 deep at 2 waves); KT 64 keeps one block in flight, deeper rings cost occupancy.
	constexpr int D = KT <= 32 ? 2 : 1;
	const int nfull = (b_end - b_begin) / 16;
	// this wave group's full blocks: kb = wb, wb + NB, ...
	const int mine = idle ? 0 : (nfull > wb ? (nfull - wb + NB - 1) / NB : 0);
	// Refills are unconditional (a refill past the end re-reads this wave's last full block
	// and is never consumed), so the number of loads in flight is the same on every path and
	// each slot waits only for its own loads, not for the whole ring.
	auto blk_row = [&](int i) { return b_begin + 16 * (wb + NB * min(i, mine - 1)); };
	blk ring[D];
	if (mine > 0) {
#pragma unroll
		for (int d = 0; d < D; d++)
			load_full(blk_row(d), ring[d]);
	}
	if constexpr (D >= 2) {
		// Software pipeline across blocks: GEMM1 of the next block is issued next to GEMM2 of
		// this one (independent MFMA chains), and the next block's ratios fill the gaps, so
		// the matrix pipe never waits on the reciprocal of its own block.  Past the last
		// block the "next" stages run on already-loaded data and their result is unused.
		float rt[4][4];
		if (mine > 0) {
			floatx4 s0[4];
			gemm1(ring[0], s0, true);
			ratio(ring[0], s0, rt, 0, std::false_type{});
		}
		for (int i = 0; i < mine; i += D) {
#pragma unroll
			for (int d = 0; d < D; d++) {
				if (i + d < mine) {
					floatx4 sn[4];
					gemm1(ring[(d + 1) % D], sn, true);
					gemm2(ring[d], rt);
					load_full(blk_row(i + d + D), ring[d]);
					ratio(ring[(d + 1) % D], sn, rt, 0, std::false_type{});
				}
			}
		}
	} else {
		for (int i = 0; i < mine; i += D) {
#pragma unroll
			for (int d = 0; d < D; d++) {
				if (i + d < mine)
					compute(ring[d], 0, std::false_type{});
				load_full(blk_row(i + d + D), ring[d]);
			}
		}
	}
	if (!idle && (b_end - b_begin) % 16 != 0 && wb == nfull % NB) {
		const int bb = b_begin + 16 * nfull;
		blk x;
		load_tail(bb, x);
		compute(x, bb, std::true_type{});
	}

	if constexpr (G2N) {
		// acc[t][cg][i] in lane (g, 4jj + n): rows 4g..4g+3 of each block, a = a0 + 16jj + 4i + t,
		// c = QN n + cg.  Sum over g, two exchanges (lane groups g ^ 2, then g ^ 1), leaving
		// component i = g in lane group g: o[t][cg] for a = a0 + 16jj + 4g + t.
		const bool hi = (g & 2) != 0, odd = (g & 1) != 0;
		float o[4][CW];
#pragma unroll
		for (int t = 0; t < 4; t++)
#pragma unroll
			for (int cg = 0; cg < CW; cg++) {
				const floatx4 v = acc[t][cg];
				const float k0 = hi ? v[2] : v[0], k1 = hi ? v[3] : v[1];
				const float s0 = hi ? v[0] : v[2], s1 = hi ? v[1] : v[3];
				const float e0 = k0 + __shfl_xor(s0, 32), e1 = k1 + __shfl_xor(s1, 32);
				o[t][cg] = (odd ? e1 : e0) + __shfl_xor(odd ? e0 : e1, 16);
			}
		if constexpr (NB > 1) {
			float (*red)[NA][4][CW][64] = reinterpret_cast<float (*)[NA][4][CW][64]>(smem);
			if (wb > 0) {
#pragma unroll
				for (int t = 0; t < 4; t++)
#pragma unroll
					for (int cg = 0; cg < CW; cg++)
						red[wb - 1][wa][t][cg][lane] = o[t][cg];
			}
			__syncthreads();
			if (wb > 0 || idle)
				return;
#pragma unroll
			for (int w = 0; w < NB - 1; w++)
#pragma unroll
				for (int t = 0; t < 4; t++)
#pragma unroll
					for (int cg = 0; cg < CW; cg++)
						o[t][cg] += red[w][wa][t][cg][lane];
		}
		if (idle)
			return;
		float *dst = ha.part + (size_t)by * p * KT;
#pragma unroll
		for (int t = 0; t < 4; t++) {
			const int a = a0 + 16 * (j >> 2) + 4 * g + t;
			if (a < p)
#pragma unroll
				for (int cg = 0; cg < CW; cg++)
					dst[(size_t)a * KT + c2 + cg] = o[t][cg];
		}
		return;
	}
	// acc[t][h][r] = partial for a = a0 + 4*(4g + r) + t, c = 16h + j.
	if constexpr (NB > 1) {
		float (*red)[NA][4][H][4][64] = reinterpret_cast<float (*)[NA][4][H][4][64]>(smem);
		if (wb > 0) {
#pragma unroll
			for (int t = 0; t < 4; t++)
#pragma unroll
				for (int h = 0; h < H; h++)
#pragma unroll
					for (int r = 0; r < 4; r++)
						red[wb - 1][wa][t][h][r][lane] = acc[t][h][r];
		}
		__syncthreads();
		if (wb > 0 || idle)
			return;
#pragma unroll
		for (int w = 0; w < NB - 1; w++)
#pragma unroll
			for (int t = 0; t < 4; t++)
#pragma unroll
				for (int h = 0; h < H; h++)
#pragma unroll
					for (int r = 0; r < 4; r++)
						acc[t][h][r] += red[w][wa][t][h][r][lane];
	}
	if (idle)
		return;
	float *dst = ha.part + (size_t)by * p * KT;
#pragma unroll
	for (int t = 0; t < 4; t++)
#pragma unroll
		for (int r = 0; r < 4; r++) {
			const int a = a0 + 4 * (4 * g + r) + t;
			if (a < p)
#pragma unroll
				for (int h = 0; h < H; h++)
					dst[(size_t)a * KT + 16 * h + j] = acc[t][h][r];
		}
}

template <int KT, int NA, int NB, int KE = KT>
__global__ __launch_bounds__(64 * NA * NB, (KT == 32 ? 3 : 1)) void k_partial_mfma(half_args h, const nmfk_state *__restrict__ st,
										       int S)
{
	__shared__ float smem[smem_mfma(KT, NA, NB)];
	if (st != nullptr && st->converged)
		return;
	int bx, by;
	if (!xcd_tile(h.nab + 1, (h.nab + 1) * S, bx, by))
		return;
	partial_mfma_body<KT, NA, NB, KE>(bx, by, h, smem);
}

// --------------------------------------------------------------------------------------
// The half-step with both contractions in BF16 on the matrix cores (MFMA 16x16x32, FP32
// accumulation); the ratio U / (A . B) and everything after it stay in FP32.
//
// Same wave ownership as the FP32 MFMA kernel (4 tiles t of 16 rows a per wave, row
// a = a0 + 4j + t), but U is walked in 32-row blocks made of two 16-row halves hb:
//   GEMM1  S[hb][t] = B[bb+16hb .. +16][:] . A[tile t][:]^T   (K = the rank, 32 per MFMA);
//          lane (g, j) then holds S at rows b = bb + 16hb + 4g + r, r = 0..3;
//   ratio  R = U / S in FP32 (hardware reciprocal, one product), rounded to BF16;
//   GEMM2  acc[t][h] += R^T . B   (K = the block's 32 rows): lane (g, j)'s 8 GEMM1 outputs
//          (both halves) are its A-operand elements, i.e. K element i of lane group g is
//          row b = bb + 16*(i/4) + 4g + (i%4); the B operand takes B at those same rows.
// B comes from the row-major BF16 copy Bb (q x KT) the finish kernel keeps beside the FP32
// factor, staged per block through LDS (see the body).
// --------------------------------------------------------------------------------------

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// LDS floats of the BF16 body: the wave groups' sums (smem_mfma), and before them each
// wave's 32-row B slot (32 x (kt + 4) BF16).
__host__ __device__ constexpr int smem_bf16(int kt, int na, int nb)
{
	return smem_mfma(kt, na, nb) > na * nb * 16 * (kt + 4) ? smem_mfma(kt, na, nb) : na * nb * 16 * (kt + 4);
}

__device__ __forceinline__ bf16x8 to_bf16x8(float4 lo, float4 hi)
{
	bf16x8 r;
	r[0] = (__bf16)lo.x, r[1] = (__bf16)lo.y, r[2] = (__bf16)lo.z, r[3] = (__bf16)lo.w;
	r[4] = (__bf16)hi.x, r[5] = (__bf16)hi.y, r[6] = (__bf16)hi.z, r[7] = (__bf16)hi.w;
	return r;
}

struct bf16_args {
	const __bf16 *Bb;    // q x KT, row-major
};

template <int KT, int NA, int NB>
__device__ __forceinline__ void partial_bf16_body(int bx, int by, const half_args &ha, const bf16_args &bb16,
						  float *__restrict__ smem)
{
	static_assert(KT % 32 == 0, "BF16 path needs KT % 32 == 0");
	constexpr int H = KT / 16;     // 16-column groups (GEMM2 outputs)
	constexpr int CH = KT / 32;    // 32-column chunks (GEMM1 K steps)
	const float *__restrict__ U = ha.U;
	const size_t ldu = ha.ldu;
	const float *__restrict__ A = ha.A;
	const __bf16 *__restrict__ Bb = bb16.Bb;
	const int p = ha.p, q = ha.q, bchunk = ha.bchunk;
	if (bx == ha.nab) {
		if (by == 0)
			den_reduce<KT>(ha.colB, ha.ncolB, ha.den, smem);
		return;
	}
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int wa = wave % NA, wb = wave / NA;
	const int j = lane & 15, g = lane >> 4;
	const int a0 = (bx * NA + wa) * 64;
	const bool idle = a0 >= p;
	const int aj = a0 + 4 * j;
	const int avail = p - aj;

	// GEMM1 B operand: lane (g, j) supplies A[aj + t][32 ch + 8 g + i] as BF16.
	bf16x8 aop[4][CH];
#pragma unroll
	for (int t = 0; t < 4; t++) {
		const bool ok = t < avail;
		const float *Ar = A + (size_t)(ok ? aj + t : 0) * KT + 8 * g;
#pragma unroll
		for (int ch = 0; ch < CH; ch++) {
			float4 lo = *reinterpret_cast<const float4 *>(Ar + 32 * ch);
			float4 hi = *reinterpret_cast<const float4 *>(Ar + 32 * ch + 4);
			if (!ok)
				lo = hi = make_float4(0.f, 0.f, 0.f, 0.f);
			aop[t][ch] = to_bf16x8(lo, hi);
		}
	}
	floatx4 acc[4][H];
#pragma unroll
	for (int t = 0; t < 4; t++)
#pragma unroll
		for (int h = 0; h < H; h++)
			acc[t][h] = floatx4{0.f, 0.f, 0.f, 0.f};

	const int b_begin = by * bchunk;
	const int b_end = min(q, b_begin + bchunk);
	const int ucol = avail > 0 ? aj : 0;

	// B reaches the matrix cores through this wave's own LDS slot, one 32-row block at a time:
	// the block's rows are fetched with whole-line 16-byte loads (B rows are KT BF16 =
	// 2 KT bytes, so NCH chunks per row and 64 / NCH rows per load), written to the slot
	// (row stride BS = KT + 4 elements, which spreads a column's rows over the banks), and
	// both operand layouts are read from there: GEMM1's (row bb + 16hb + j, 8 consecutive
	// columns) as row segments, GEMM2's (column 16h + j, rows bb + 16(i / 4) + 4g + i % 4) as
	// a gather.  Read straight from memory, the two layouts touch 16 partly used cache lines
	// per load, 3x the line traffic of U itself; staged, B costs whole lines and LDS reads.
	// Each wave uses only its own slot and LDS executes a wave's accesses in order, so no
	// barrier is needed.
	constexpr int BS = KT + 4;
	constexpr int NCH = KT / 8;             // 16-byte chunks per B row
	constexpr int RPL = 64 / NCH;           // rows one load instruction covers
	constexpr int NLD = 32 / RPL;           // loads per 32-row block
	static_assert(32 % RPL == 0, "B staging: a block is whole loads");
	__bf16 *__restrict__ slot = reinterpret_cast<__bf16 *>(smem) + (size_t)wave * 32 * BS;
	const int lrow = lane / NCH, lchunk = lane % NCH;
	uint4 braw[NLD];
	// Rows at or past b_end (the partial last block of a split) read row b_begin instead.
	auto fetch_b = [&](int bb, bool full) {
#pragma unroll
		for (int i = 0; i < NLD; i++) {
			int b = bb + lrow + RPL * i;
			if (!full)
				b = b < b_end ? b : b_begin;
			braw[i] = *reinterpret_cast<const uint4 *>(Bb + (size_t)b * KT + 8 * lchunk);
		}
	};
	auto stage_b = [&]() {
#pragma unroll
		for (int i = 0; i < NLD; i++) {
			uint2 *d = reinterpret_cast<uint2 *>(slot + (lrow + RPL * i) * BS + 8 * lchunk);
			d[0] = make_uint2(braw[i].x, braw[i].y);
			d[1] = make_uint2(braw[i].z, braw[i].w);
		}
	};
	bf16x8 b1[2][CH], b2[H];
	auto read_b1 = [&]() {
#pragma unroll
		for (int hb = 0; hb < 2; hb++)
#pragma unroll
			for (int ch = 0; ch < CH; ch++) {
				const uint2 *src = reinterpret_cast<const uint2 *>(slot + (16 * hb + j) * BS + 32 * ch + 8 * g);
				const uint2 lo = src[0], hi = src[1];
				const uint4 w = make_uint4(lo.x, lo.y, hi.x, hi.y);
				__builtin_memcpy(&b1[hb][ch], &w, 16);
			}
	};
	auto read_b2 = [&]() {
#pragma unroll
		for (int h = 0; h < H; h++)
#pragma unroll
			for (int i = 0; i < 8; i++)
				b2[h][i] = slot[(16 * (i / 4) + 4 * g + i % 4) * BS + 16 * h + j];
	};

	// One block: GEMM1 S = B . A^T on both 16-row halves, the ratios U / S (FP32, hardware
	// reciprocal, one product), GEMM2 acc += R^T . B.  MASK zeroes the ratios of rows at or
	// past b_end.
	auto gemm1_ratio = [&](const float4 (&u)[2][4], bf16x8 (&rt)[4], int bb, auto mask_tag) {
		constexpr bool MASK = decltype(mask_tag)::value;
#pragma unroll
		for (int hb = 0; hb < 2; hb++) {
			floatx4 sv[4];
#pragma unroll
			for (int t = 0; t < 4; t++) {
				sv[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
				for (int ch = 0; ch < CH; ch++)
					sv[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b1[hb][ch], aop[t][ch], sv[t], 0, 0, 0);
			}
#pragma unroll
			for (int r = 0; r < 4; r++) {
				const bool rok = !MASK || bb + 16 * hb + 4 * g + r < b_end;
				const float uv[4] = {u[hb][r].x, u[hb][r].y, u[hb][r].z, u[hb][r].w};
#pragma unroll
				for (int t = 0; t < 4; t++) {
					const float v = uv[t] * __builtin_amdgcn_rcpf(sv[t][r]);
					rt[t][4 * hb + r] = (__bf16)(rok ? v : 0.0f);
				}
			}
		}
	};
	auto gemm2 = [&](const bf16x8 (&rt)[4]) {
#pragma unroll
		for (int t = 0; t < 4; t++)
#pragma unroll
			for (int h = 0; h < H; h++)
				acc[t][h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rt[t], b2[h], acc[t][h], 0, 0, 0);
	};
	auto load_u = [&](int bb, float4 (&u)[2][4], bool full) {
#pragma unroll
		for (int hb = 0; hb < 2; hb++)
#pragma unroll
			for (int r = 0; r < 4; r++) {
				int b = bb + 16 * hb + 4 * g + r;
				if (!full)
					b = b < b_end ? b : b_begin;
				u[hb][r] = *reinterpret_cast<const float4 *>(U + (size_t)b * ldu + ucol);
			}
	};

	const int nfull = (b_end - b_begin) / 32;
	const int mine = idle ? 0 : (nfull > wb ? (nfull - wb + NB - 1) / NB : 0);
	auto blk_row = [&](int i) { return b_begin + 32 * (wb + NB * min(i, mine - 1)); };
	if (mine > 0) {
		// U (the HBM stream) has a two-block ring; B's next block is fetched into registers
		// while this one is in the slot.  Refills past the end re-read this wave's last block
		// and are never consumed.
		float4 ur[2][2][4];
		auto step = [&](float4 (&u)[2][4], int next_b, int next_u) {
			stage_b();
			read_b1();
			bf16x8 rt[4];
			gemm1_ratio(u, rt, 0, std::false_type{});
			read_b2();
			fetch_b(next_b, true);
			load_u(next_u, u, true);
			gemm2(rt);
		};
		load_u(blk_row(0), ur[0], true);
		load_u(blk_row(1), ur[1], true);
		fetch_b(blk_row(0), true);
		for (int i = 0; i < mine; i += 2) {
			step(ur[0], blk_row(i + 1), blk_row(i + 2));
			if (i + 1 < mine)
				step(ur[1], blk_row(i + 2), blk_row(i + 3));
		}
	}
	if (!idle && (b_end - b_begin) % 32 != 0 && wb == nfull % NB) {
		const int bb = b_begin + 32 * nfull;
		float4 u[2][4];
		load_u(bb, u, false);
		fetch_b(bb, false);
		stage_b();
		read_b1();
		bf16x8 rt[4];
		gemm1_ratio(u, rt, bb, std::true_type{});
		read_b2();
		gemm2(rt);
	}

	if constexpr (NB > 1) {
		__syncthreads();   // every wave is done with its B slot, which the sums below reuse
		float (*red)[NA][4][H][4][64] = reinterpret_cast<float (*)[NA][4][H][4][64]>(smem);
		if (wb > 0) {
#pragma unroll
			for (int t = 0; t < 4; t++)
#pragma unroll
				for (int h = 0; h < H; h++)
#pragma unroll
					for (int r = 0; r < 4; r++)
						red[wb - 1][wa][t][h][r][lane] = acc[t][h][r];
		}
		__syncthreads();
		if (wb > 0 || idle)
			return;
#pragma unroll
		for (int w = 0; w < NB - 1; w++)
#pragma unroll
			for (int t = 0; t < 4; t++)
#pragma unroll
				for (int h = 0; h < H; h++)
#pragma unroll
					for (int r = 0; r < 4; r++)
						acc[t][h][r] += red[w][wa][t][h][r][lane];
	}
	if (idle)
		return;
	float *dst = ha.part + (size_t)by * p * KT;
#pragma unroll
	for (int t = 0; t < 4; t++)
#pragma unroll
		for (int r = 0; r < 4; r++) {
			const int a = a0 + 4 * (4 * g + r) + t;
			if (a < p)
#pragma unroll
				for (int h = 0; h < H; h++)
					dst[(size_t)a * KT + 16 * h + j] = acc[t][h][r];
		}
}

template <int KT, int NA, int NB>
__global__ __launch_bounds__(64 * NA * NB) void k_partial_bf16(half_args h, bf16_args b, const nmfk_state *__restrict__ st,
							     int S)
{
	__shared__ __align__(16) float smem[smem_bf16(KT, NA, NB)];
	if (st != nullptr && st->converged)
		return;
	int bx, by;
	if (!xcd_tile(h.nab + 1, (h.nab + 1) * S, bx, by))
		return;
	partial_bf16_body<KT, NA, NB>(bx, by, h, b, smem);
}

// A[a][c] = (A[a][c] * sum_s part[s][a][c]) / den[c], optional clamp, then the column sums
// of the new A per block of 64 rows (the next half-step's denominators).  Kernel_mul_div
// order: multiply first, then divide (src/GPU_kernels.cu:2029-2033).  Block bx covers rows
// [64 bx, 64 bx + 64); smem holds NMFK_FIN_ROWS x (KT + 1) floats.
__host__ __device__ constexpr int smem_fin(int kt)
{
	return NMFK_FIN_ROWS * (kt + 1);
}

// Q consecutive floats at a Q-aligned float offset (Q = KT/4: 1, 2, or a multiple of 4).
template <int Q>
__device__ __forceinline__ void ldq(const float *__restrict__ p, float (&v)[Q])
{
	if constexpr (Q % 4 == 0) {
#pragma unroll
		for (int i = 0; i < Q; i += 4) {
			const float4 t = *reinterpret_cast<const float4 *>(p + i);
			v[i] = t.x, v[i + 1] = t.y, v[i + 2] = t.z, v[i + 3] = t.w;
		}
	} else if constexpr (Q == 2) {
		const float2 t = *reinterpret_cast<const float2 *>(p);
		v[0] = t.x, v[1] = t.y;
	} else {
#pragma unroll
		for (int i = 0; i < Q; i++)
			v[i] = p[i];
	}
}

template <int Q>
__device__ __forceinline__ void stq(float *__restrict__ p, const float (&v)[Q])
{
	if constexpr (Q % 4 == 0) {
#pragma unroll
		for (int i = 0; i < Q; i += 4)
			*reinterpret_cast<float4 *>(p + i) = make_float4(v[i], v[i + 1], v[i + 2], v[i + 3]);
	} else if constexpr (Q == 2) {
		*reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
	} else {
#pragma unroll
		for (int i = 0; i < Q; i++)
			p[i] = v[i];
	}
}

// The BF16 copy the finish kernel keeps of the factor it updates (null: none).
struct shadow_args {
	__bf16 *Ab;
};

// BF16 copy of a factor A (p x KT), row-major like A: thread (row a, quarter c0) writes its
// KT / 4 consecutive values, 16 bytes at a time.
template <int KT>
__device__ __forceinline__ void store_shadow(int a, int c0, const float (&x)[KT / 4], __bf16 *__restrict__ Ab)
{
	constexpr int Q = KT / 4;
	static_assert(Q % 8 == 0, "BF16 copies: kt a multiple of 32");
#pragma unroll
	for (int i = 0; i < Q; i += 8) {
		bf16x8 v;
#pragma unroll
		for (int e = 0; e < 8; e++)
			v[e] = (__bf16)x[i + e];
		*reinterpret_cast<bf16x8 *>(Ab + (size_t)a * KT + c0 + i) = v;
	}
}

template <int KT>
__global__ __launch_bounds__(NMFK_WG) void k_shadow(const float *__restrict__ A, int p, __bf16 *__restrict__ Ab)
{
	constexpr int Q = KT / 4;
	const int a = blockIdx.x * NMFK_FIN_ROWS + threadIdx.x / 4;
	const int c0 = (threadIdx.x % 4) * Q;
	if (a >= p)
		return;
	float x[Q];
	ldq<Q>(A + (size_t)a * KT + c0, x);
	store_shadow<KT>(a, c0, x, Ab);
}

template <int KT>
__device__ __forceinline__ void finish_body(int bx, const float *__restrict__ part, int S, float *__restrict__ A,
					    int p, int k, const float *__restrict__ den, float *__restrict__ colA,
					    int clamp, float *__restrict__ smem, const shadow_args &sh = shadow_args{})
{
	constexpr int TPR = fin_tpr(KT), Q = fin_q(KT);
	const int a = bx * NMFK_FIN_ROWS + threadIdx.x / TPR;
	const int c0 = (threadIdx.x % TPR) * Q;
	float x[Q];
	if (a < p) {
		const size_t pst = (size_t)p * KT;
		const float *pa = part + (size_t)a * KT + c0;
		float s[Q];
		ldq<Q>(pa, s);
		int sp = 1;
		// 8 splits' loads in flight per step; the additions keep ascending split order.
		for (; sp + 8 <= S; sp += 8) {
			float v[8][Q];
#pragma unroll
			for (int u = 0; u < 8; u++)
				ldq<Q>(pa + (sp + u) * pst, v[u]);
#pragma unroll
			for (int u = 0; u < 8; u++)
#pragma unroll
				for (int i = 0; i < Q; i++)
					s[i] += v[u][i];
		}
		// The last S - sp < 8 splits.  Up to KT 32 all their loads go out at once before the
		// same ascending additions (c1 H finish, 30 splits: 4.02 -> 3.76 us); at KT 64 the
		// registers that takes halve the occupancy and the W finish at 50000 x 2000 slows from
		// 11.4 to 12.7 us, so there they stay one at a time.
		if constexpr (KT <= 32) {
			if (sp < S) {
				float v[7][Q];
#pragma unroll
				for (int u = 0; u < 7; u++)
					if (sp + u < S)
						ldq<Q>(pa + (sp + u) * pst, v[u]);
#pragma unroll
				for (int u = 0; u < 7; u++)
					if (sp + u < S)
#pragma unroll
						for (int i = 0; i < Q; i++)
							s[i] += v[u][i];
			}
		} else {
			for (; sp < S; sp++) {
				float v[Q];
				ldq<Q>(pa + sp * pst, v);
#pragma unroll
				for (int i = 0; i < Q; i++)
					s[i] += v[i];
			}
		}
		float *Ar = A + (size_t)a * KT + c0;
		float ar[Q];
		ldq<Q>(Ar, ar);
#pragma unroll
		for (int i = 0; i < Q; i++) {
			float v = 0.0f;
			if (c0 + i < k) {
				v = ar[i] * s[i];
				v = v / den[c0 + i];
				if (clamp)
					v = clamp_rmin(v);
			}
			x[i] = v;
		}
		stq<Q>(Ar, x);
		if constexpr (KT % 32 == 0) {
			if (sh.Ab != nullptr)
				store_shadow<KT>(a, c0, x, sh.Ab);
		}
	} else {
#pragma unroll
		for (int i = 0; i < Q; i++)
			x[i] = 0.0f;
	}
	block_colsum_store<KT>(x, colA + (size_t)bx * KT, smem);
}

template <int KT>
__global__ __launch_bounds__(NMFK_WG) void k_finish(const float *__restrict__ part, int S, float *__restrict__ A,
						    int p, int k, const float *__restrict__ den,
						    float *__restrict__ colA, int clamp,
						    const nmfk_state *__restrict__ st, shadow_args sh)
{
	__shared__ float smem[smem_fin(KT)];
	if (st != nullptr && st->converged)
		return;
	finish_body<KT>(blockIdx.x, part, S, A, p, k, den, colA, clamp, smem, sh);
}

template <int KT>
__global__ __launch_bounds__(NMFK_WG) void k_colsum_parts(const float *__restrict__ A, int p,
							  float *__restrict__ colA)
{
	__shared__ float smem[smem_fin(KT)];
	constexpr int TPR = fin_tpr(KT), Q = fin_q(KT);
	const int a = blockIdx.x * NMFK_FIN_ROWS + threadIdx.x / TPR;
	const int c0 = (threadIdx.x % TPR) * Q;
	float x[Q];
#pragma unroll
	for (int i = 0; i < Q; i++)
		x[i] = (a < p) ? A[(size_t)a * KT + c0 + i] : 0.0f;
	block_colsum_store<KT>(x, colA + (size_t)blockIdx.x * KT, smem);
}

template <int KT>
__global__ __launch_bounds__(NMFK_WG) void k_den(const float *__restrict__ parts, int nparts, float *__restrict__ den)
{
	__shared__ float smem[4 * KT];
	den_reduce<KT>(parts, nparts, den, smem);
}

// --------------------------------------------------------------------------------------
// One pass over V per iteration (FP32 matrix cores, 16x16x4).
//
// With H(t) fixed, the W half-step of iteration t and the sums of the H half-step of
// iteration t + 1 both run over the rows of V, and both are row-separable:
//   W(t)[i]    = W(t-1)[i] * (sum_j H(t)[j] V[i][j] / (W(t-1)[i] . H(t)[j])) / colsum(H(t))
//   sums[j]   += W(t)[i] V[i][j] / (W(t)[i] . H(t)[j])          over the rows i of V
// so one workgroup can take a block of 16 whole rows, finish their new W rows, and add
// their share of the next H update while the block is still on chip: V is read from HBM
// once per iteration instead of once per half-step (src/NMF_routines.c:704-1000 then
// :1344-1636, in the same order; the stream path of nmfk_stream.c does the same per block).
//
// Workgroup = NWV waves; wave w owns the 16 JT columns j0 = 16 JT w .. of V as JT tiles jt
// of 16, and holds H at those rows in registers for the whole pass, in the two operand
// layouts the products need.  Each phase runs over all the wave's tiles without branches
// (tiles past m hold zero rows of H and are masked), so the tiles' independent matrix-core
// chains interleave.  For each row block (rows i0 .. i0 + 15):
//   phase A (W rows):  GEMM1 S[j][i] = H[j] . W[i] (M = j, N = i, K = rank), ratio
//            R = V / S with V[i0 + l16][j0 + 16jt + 4g + r] (one float4 per lane and tile),
//            GEMM2 Waux[i][c] += sum_j R[j][i] H[j][c] (K = j: GEMM1's output register r is
//            the A operand of the K chunk {4g + r}, so no data moves between the products);
//            the waves' Waux are added through LDS in wave order, and the new rows
//            W[i] * Waux[i] / colsum(H) (multiply, then divide) are written out;
//   phase B (H sums):  GEMM1 S'[i][j] = W_new[i] . H[j] (M = i, N = j), ratio with
//            V[i0 + 4g + r][j0 + 16jt + l16], i.e. the block transposed, which the wave
//            reads back from its own LDS copy of the tiles it loaded in phase A; GEMM2
//            sums[j][c] += sum_i R'[i][j] W_new[i][c] (K = i) in registers across blocks.
// A workgroup takes a contiguous range of row blocks and writes its sums as one slice of
// part (G x m x KT) plus the column sums of its new W rows; k_fused_finish adds the slices
// in workgroup order.  colsum(H(t)) is formed in the kernel from the H it holds.
//
// Test passes (the last iteration of a test interval) clamp the new W rows to R_MIN before
// phase B and contract phase B against R_MIN-clamped H, as the reference clamps both before
// its test (src/NMF_GPU.c:611-638) after update_W used the unclamped H; the test kernel then
// clamps H in place before k_fused_finish multiplies it.
// --------------------------------------------------------------------------------------

// x where ok, else +0 (a bit mask, not a branch: the operands are always computed, and the
// compiler does not move their loads behind a branch).
__device__ __forceinline__ float masked(bool ok, float x)
{
	return __uint_as_float(__float_as_uint(x) & (ok ? 0xffffffffu : 0u));
}

#define NMFK_FUSED_RB 16      // rows of V per block
#define NMFK_FUSED_VRS 20     // LDS row stride (floats) of a wave's 16 x 16 V tiles
#define NMFK_FUSED_COLS 128   // columns of V per wave (8 tiles)
#define NMFK_FUSED_MAXW 8     // waves per workgroup at most (m <= 1024)

struct fused_args {
	const float *V;
	size_t ldv;
	float *W;             // n x KT: W(t-1) in, W(t) out, rows of this workgroup only
	const float *H;       // m x KT: H(t), read only
	float *part;          // G x m x KT: this pass's sums of the next H half-step
	float *colW;          // G x KT: column sums of the new W rows of each workgroup
	int n, m, k;
	int nblocks;          // ceil(n / 16)
	int G;                // workgroups; workgroup g takes blocks [g nblocks / G, (g+1) nblocks / G)
	int upd_w;            // 0: W stays (the first pass only adds the sums of H(1) against W(0))
	int clamp;            // test pass
};

template <int KT, int NWV, int JT>
__global__ __launch_bounds__(64 * NWV, 2) void k_fused_mfma(fused_args fa, const nmfk_state *__restrict__ st)
{
	static_assert(KT % 16 == 0, "matrix-core pass needs KT % 16 == 0");
	static_assert(NWV * 64 >= NMFK_FUSED_RB * KT, "one thread per new W element of a block");
	constexpr int HN = KT / 16;   // 16-column groups of the rank
	constexpr int KS = KT / 4;    // K steps of GEMM1 (K = the rank, 4 per MFMA)
	constexpr int NA = 1;         // phase A accumulators (tiles jt add into accA[jt % NA])
	constexpr int RB = NMFK_FUSED_RB;
	constexpr int VRS = NMFK_FUSED_VRS;
	__shared__ float red[NWV][RB][KT];          // per-wave W-half sums of the block
	__shared__ float wnl[RB][KT + 1];           // the block's new W rows
	__shared__ float wol[RB][KT + 1];           // the block's W(t-1) rows
	__shared__ float cws[RB][KT + 1];           // end: column sums of the new W, by row slot
	__shared__ float hsum[NWV][KT];             // per-wave column sums of H(t)
	__shared__ float den_s[KT];                 // colsum(H(t))
	__shared__ __align__(16) float vt[NWV][JT][RB][VRS];   // each wave's V tiles, for the transposed read
	__shared__ __align__(16) float hbl[NWV][JT][64][4 * HN]; // each lane's GEMM2 operand of H in phase A
	if (st != nullptr && st->converged)
		return;

	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int l16 = lane & 15, g = lane >> 4;
	const int n = fa.n, m = fa.m, k = fa.k;
	const size_t ldv = fa.ldv;
	const int j0 = wave * 16 * JT;
	const bool active = j0 < m;               // wave-uniform: the wave has columns
	const int gid = blockIdx.x;
	const int b0 = (int)(((long long)gid * fa.nblocks) / fa.G);
	const int b1 = (int)(((long long)(gid + 1) * fa.nblocks) / fa.G);
	const float *__restrict__ H = fa.H;
	const int t = threadIdx.x;
	const int wi = t / KT, wc = t % KT;       // W element of the block this thread finishes (t < RB*KT)
	const bool wown = t < RB * KT;

	// Loads of a block, issued one block ahead of its phase A: V rows through a buffer
	// descriptor based at the block's first row and sized to the rows that exist (rows past n
	// read zero; offsets past a row's columns read other rows or zero, and are masked), and
	// W(t-1) rows likewise: wa[s] = W[i0 + l16][4s + g], GEMM1's N operand in phase A, which
	// wave 0 also leaves in LDS (wol) for the new-row step.  Nothing reads these registers
	// before phase A, so their loads stay in flight across the barriers in between.
	const int voff = (int)(((size_t)l16 * ldv + (size_t)(j0 + 4 * g)) * 4);   // tile jt: + 64 jt bytes
	float4 vr[JT];
	float wa[KS];
	auto rows_of = [&](int blk) { return min(RB, n - blk * RB); };
	auto load_blk = [&](int blk) {
		const int i0 = blk * RB;
		const int rows = rows_of(blk);
		const __amdgpu_buffer_rsrc_t rv = make_rsrc(fa.V + (size_t)i0 * ldv, (size_t)rows * ldv * 4);
#pragma unroll
		for (int jt = 0; jt < JT; jt++)
			vr[jt] = bload4(rv, voff + 64 * jt, 0);
		const __amdgpu_buffer_rsrc_t rw = make_rsrc(fa.W + (size_t)i0 * KT, (size_t)rows * KT * 4);
#pragma unroll
		for (int s = 0; s < KS; s++)
			wa[s] = bload1(rw, (l16 * KT + 4 * s + g) * 4, 0);
	};
	auto zero4 = [](floatx4 (&a)[HN]) {
#pragma unroll
		for (int h = 0; h < HN; h++)
			a[h] = floatx4{0.f, 0.f, 0.f, 0.f};
	};
	// The first block's loads are in flight while H is set up.
	if (b0 < b1)
		load_blk(b0);

	// H (m x KT, contiguous) is staged in LDS through the V-tile area, which is free until the
	// first phase A, with coalesced 16-byte loads by the whole workgroup.  From there each
	// wave takes H at its columns, in registers: ha[jt][s] = H[j0 + 16jt + l16][4s + g] (the
	// K = rank operand of both phases), and in the lane's own LDS slots:
	// hbl[w][jt][lane][4h + r] = H[j0 + 16jt + 4g + r][16h + l16] (phase A's K = j operand).
	// Rows past m are zero, so past-the-end tiles contribute nothing.
	static_assert(sizeof(vt) >= (size_t)NWV * JT * 16 * KT * sizeof(float), "H staging fits the V tiles");
	float *hs = &vt[0][0][0][0];
	{
		const int nq = m * KT / 4;   // float4s of H
		const float4 *H4 = reinterpret_cast<const float4 *>(H);
		for (int q = t; q < nq; q += 64 * NWV)
			*reinterpret_cast<float4 *>(hs + 4 * q) = H4[q];
	}
	__syncthreads();
	float ha[JT][KS];
	float hcs[HN];
#pragma unroll
	for (int h = 0; h < HN; h++)
		hcs[h] = 0.0f;
#pragma unroll
	for (int jt = 0; jt < JT; jt++) {
		const int ja = j0 + 16 * jt + l16;
#pragma unroll
		for (int s = 0; s < KS; s++)
			ha[jt][s] = ja < m ? hs[ja * KT + 4 * s + g] : 0.0f;
#pragma unroll
		for (int h = 0; h < HN; h++) {
			float4 q;
			float *qv = reinterpret_cast<float *>(&q);
#pragma unroll
			for (int r = 0; r < 4; r++) {
				const int jb = j0 + 16 * jt + 4 * g + r;
				qv[r] = jb < m ? hs[jb * KT + 16 * h + l16] : 0.0f;
				hcs[h] += qv[r];
			}
			*reinterpret_cast<float4 *>(&hbl[wave][jt][lane][4 * h]) = q;
		}
	}
	// colsum(H(t)): the wave's rows in a fixed order (tiles, rows r, then the lane groups by
	// a xor tree), then the waves in order.
#pragma unroll
	for (int h = 0; h < HN; h++) {
		float s = hcs[h];
		s += __shfl_xor(s, 16, 64);
		s += __shfl_xor(s, 32, 64);
		if (g == 0)
			hsum[wave][16 * h + l16] = s;
	}
	__syncthreads();   // also: every wave has read hs before the V tiles overwrite it
	if (t < KT) {
		float s = 0.0f;
#pragma unroll
		for (int w = 0; w < NWV; w++)
			s += hsum[w][t];
		den_s[t] = s;
	}
	// (den_s is read after the first barrier of the block loop)

	// Phase B contracts against H clamped to R_MIN on test passes (columns < k only).

	floatx4 accB[JT][HN];
#pragma unroll
	for (int jt = 0; jt < JT; jt++)
#pragma unroll
		for (int h = 0; h < HN; h++)
			accB[jt][h] = floatx4{0.f, 0.f, 0.f, 0.f};
	float cw = 0.0f;

	// Each phase is two straight-line stages over all tiles, so that the tiles' independent
	// matrix chains interleave: GEMM1, then the ratios and GEMM2.  The loop body orders them
	// GEMM1(B), GEMM1(A), ratios + GEMM2(B), ratios + GEMM2(A), so that each phase's ratios
	// (vector ALU work) run while the other phase's products occupy the matrix cores.
	// Phase A, GEMM1: S[jt][r] = S[j0 + 16jt + 4g + r][i0 + l16] = H[j] . W(t-1)[i].
	auto gemm1_a = [&](floatx4 (&S)[JT]) {
#pragma unroll
		for (int jt = 0; jt < JT; jt++)
			S[jt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
		for (int s = 0; s < KS; s++)
#pragma unroll
			for (int jt = 0; jt < JT; jt++)
				S[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ha[jt][s], wa[s], S[jt], 0, 0, 0);
	};
	// Phase A, ratios and GEMM2: accA[jt % NA][h][q] = Waux[i0 + 4g + q][16h + l16] over tile jt.
	// FULL: the block has 16 rows and the wave's columns are all < m (no masks).
	auto gemm2_a = [&](int rows, const floatx4 (&S)[JT], floatx4 (&accA)[NA][HN], auto full_tag) {
		constexpr bool FULL = decltype(full_tag)::value;
		const bool iok = l16 < rows;
#pragma unroll
		for (int jt = 0; jt < JT; jt++) {
			const float v[4] = {vr[jt].x, vr[jt].y, vr[jt].z, vr[jt].w};
			float R[4];
#pragma unroll
			for (int r = 0; r < 4; r++) {
				const float x = v[r] * __builtin_amdgcn_rcpf(S[jt][r]);
				R[r] = FULL ? x : masked(iok && j0 + 16 * jt + 4 * g + r < m, x);
			}
#pragma unroll
			for (int h = 0; h < HN; h++) {
				const float4 q = *reinterpret_cast<const float4 *>(&hbl[wave][jt][lane][4 * h]);
				const float hv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
				for (int r = 0; r < 4; r++)
					accA[jt % NA][h] = __builtin_amdgcn_mfma_f32_16x16x4f32(R[r], hv[r], accA[jt % NA][h], 0, 0, 0);
			}
		}
	};
	// The V tiles a wave keeps for phase B (vt[w][jt][row][col] = V[i0 + row][j0 + 16jt + col]),
	// and wave 0's copy of the W(t-1) rows for the new-row step.
	auto store_vt = [&]() {
#pragma unroll
		for (int jt = 0; jt < JT; jt++)
			*reinterpret_cast<float4 *>(&vt[wave][jt][l16][4 * g]) = vr[jt];
		if (wave == 0)
#pragma unroll
			for (int s = 0; s < KS; s++)
				wol[l16][4 * s + g] = wa[s];
	};
	// The wave's W-half sums of the block.
	auto store_red = [&](const floatx4 (&accA)[NA][HN]) {
#pragma unroll
		for (int h = 0; h < HN; h++)
#pragma unroll
			for (int q = 0; q < 4; q++)
				red[wave][4 * g + q][16 * h + l16] = accA[0][h][q];
	};
	// Phase B (the next H half-step's sums against the new rows), GEMM1:
	// S[jt][r] = S'[i0 + 4g + r][j0 + 16jt + l16] = W(t)[i] . H[j], H clamped on test passes.
	auto gemm1_b = [&](const float (&wna)[KS], floatx4 (&S)[JT], auto clamp_tag) {
		constexpr bool CLAMP = decltype(clamp_tag)::value;
#pragma unroll
		for (int jt = 0; jt < JT; jt++)
			S[jt] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
		for (int s = 0; s < KS; s++)
#pragma unroll
			for (int jt = 0; jt < JT; jt++) {
				const float hv = (CLAMP && 4 * s + g < k) ? clamp_rmin(ha[jt][s]) : ha[jt][s];
				S[jt] = __builtin_amdgcn_mfma_f32_16x16x4f32(wna[s], hv, S[jt], 0, 0, 0);
			}
	};
	// Phase B, ratios (V read back transposed from the wave's LDS tiles) and GEMM2.
	auto gemm2_b = [&](int rows, const floatx4 (&S)[JT], const float (&wnb)[4][HN], auto full_tag) {
		constexpr bool FULL = decltype(full_tag)::value;
#pragma unroll
		for (int jt = 0; jt < JT; jt++) {
			const bool jok = j0 + 16 * jt + l16 < m;
			float v[4], R[4];
#pragma unroll
			for (int r = 0; r < 4; r++)
				v[r] = vt[wave][jt][4 * g + r][l16];
#pragma unroll
			for (int r = 0; r < 4; r++) {
				const float x = v[r] * __builtin_amdgcn_rcpf(S[jt][r]);
				R[r] = FULL ? x : masked(jok && 4 * g + r < rows, x);
			}
#pragma unroll
			for (int r = 0; r < 4; r++)
#pragma unroll
				for (int h = 0; h < HN; h++)
					accB[jt][h] = __builtin_amdgcn_mfma_f32_16x16x4f32(R[r], wnb[r][h], accB[jt][h], 0, 0, 0);
		}
	};
	// One loop body: phase B of block b and (WITH_A) phase A of block b + 1, interleaved;
	// FULL when both blocks have 16 rows and the wave's columns are all < m, CLAMP on test
	// passes (compiled apart, so the common body carries no masks and no clamp).
	auto block_body = [&](int rows, int rows_n, floatx4 (&accA)[NA][HN], auto with_a_tag, auto full_tag,
			      auto clamp_tag) {
		constexpr bool WITH_A = decltype(with_a_tag)::value;
		float wna[KS], wnb[4][HN];
#pragma unroll
		for (int s = 0; s < KS; s++)
			wna[s] = wnl[l16][4 * s + g];
#pragma unroll
		for (int r = 0; r < 4; r++)
#pragma unroll
			for (int h = 0; h < HN; h++)
				wnb[r][h] = wnl[4 * g + r][16 * h + l16];
		floatx4 SB[JT], SA[JT];
		gemm1_b(wna, SB, clamp_tag);
		if constexpr (WITH_A)
			gemm1_a(SA);
		gemm2_b(rows, SB, wnb, full_tag);
		if constexpr (WITH_A)
			gemm2_a(rows_n, SA, accA, full_tag);
	};
	const bool cols_full = j0 + 16 * JT <= m;   // wave-uniform

	// Software pipeline over the workgroup's row blocks: the loop body finishes block b's new
	// rows, runs phase B of block b, then phase A of block b + 1 (whose V tiles replace block
	// b's in LDS once phase B has read them), while the loads of block b + 2 are in flight.
	// Wave 0 is always active (m >= 1), so wol is written for every block.
	if (b0 < b1) {
		if (active) {
			floatx4 accA[NA][HN];
#pragma unroll
			for (int a = 0; a < NA; a++)
				zero4(accA[a]);
			if (fa.upd_w) {
				floatx4 SA[JT];
				gemm1_a(SA);
				gemm2_a(rows_of(b0), SA, accA, std::false_type{});
			}
			store_vt();
			store_red(accA);
		} else {
			floatx4 accA[NA][HN];
#pragma unroll
			for (int a = 0; a < NA; a++)
				zero4(accA[a]);
			store_red(accA);
		}
		if (b0 + 1 < b1)
			load_blk(b0 + 1);
	}

	for (int b = b0; b < b1; b++) {
		const int i0 = b * RB;
		const int rows = rows_of(b);
		__syncthreads();
		// New W rows of block b (kernel_mul_div order: multiply, then divide).
		if (wown) {
			const float wold = wol[wi][wc];
			float wn = 0.0f;
			if (wi < rows) {
				if (fa.upd_w) {
					if (wc < k) {
						float waux = 0.0f;
#pragma unroll
						for (int w = 0; w < NWV; w++)
							waux += red[w][wi][wc];
						wn = wold * waux;
						wn = wn / den_s[wc];
						if (fa.clamp)
							wn = clamp_rmin(wn);
					}
					fa.W[(size_t)(i0 + wi) * KT + wc] = wn;
				} else {
					wn = wold;
				}
			}
			wnl[wi][wc] = wn;
			cw += wn;
		}
		__syncthreads();
		const bool next = b + 1 < b1;
		floatx4 accA[NA][HN];
#pragma unroll
		for (int a = 0; a < NA; a++)
			zero4(accA[a]);
		if (active) {
			const std::true_type T;
			const std::false_type F;
			if (next && fa.upd_w) {
				const bool full = cols_full && rows == RB && rows_of(b + 1) == RB;
				if (full && !fa.clamp)
					block_body(rows, RB, accA, T, T, F);
				else if (full)
					block_body(rows, RB, accA, T, T, T);
				else if (!fa.clamp)
					block_body(rows, rows_of(b + 1), accA, T, F, F);
				else
					block_body(rows, rows_of(b + 1), accA, T, F, T);
			} else if (fa.clamp) {
				block_body(rows, 0, accA, F, F, T);
			} else {
				block_body(rows, 0, accA, F, F, F);
			}
			if (next)
				store_vt();
		}
		if (next)
			store_red(accA);
		if (b + 2 < b1)
			load_blk(b + 2);
		// red and wnl are rewritten only after the next barrier, when every wave has read them.
	}

	// This workgroup's slice of the sums: accB[jt][h][q] = sums[j0 + 16jt + 4g + q][16h + l16].
	float *dst = fa.part + (size_t)gid * m * KT;
#pragma unroll
	for (int jt = 0; jt < JT; jt++)
#pragma unroll
		for (int q = 0; q < 4; q++) {
			const int j = j0 + 16 * jt + 4 * g + q;
			if (j < m)
#pragma unroll
				for (int h = 0; h < HN; h++)
					dst[(size_t)j * KT + 16 * h + l16] = accB[jt][h][q];
		}
	if (wown)
		cws[wi][wc] = cw;
	__syncthreads();
	if (t < KT) {
		float s = 0.0f;
#pragma unroll
		for (int i = 0; i < RB; i++)
			s += cws[i][t];
		fa.colW[(size_t)gid * KT + t] = s;
	}
}

// H(t+1) = H(t) * (sum over the G slices of part, in slice order) / colsum(W(t)), with
// colsum(W(t)) from the G column-sum rows (den_reduce).  One output element per lane; the
// 16 waves of a workgroup take the slices g = w, w + 16, ... and are added in wave order.
#define NMFK_FFIN_WAVES 16

template <int KT>
__global__ __launch_bounds__(64 * NMFK_FFIN_WAVES) void k_fused_finish(const float *__restrict__ part, int G,
									 const float *__restrict__ colW,
									 float *__restrict__ H, int m, int k,
									 const nmfk_state *__restrict__ st)
{
	__shared__ float red[NMFK_FFIN_WAVES][64];
	__shared__ float s4[4 * KT];
	__shared__ float den[KT];
	if (st != nullptr && st->converged)
		return;
	den_reduce<KT>(colW, G, den, s4);
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const size_t tot = (size_t)m * KT;
	const size_t e = (size_t)blockIdx.x * 64 + lane;
	float s = 0.0f;
	if (e < tot) {
		const float *p = part + e;
		int gg = w;
		// 8 slices' loads in flight per step; the additions keep ascending slice order.
		for (; gg + 7 * NMFK_FFIN_WAVES < G; gg += 8 * NMFK_FFIN_WAVES) {
			float v[8];
#pragma unroll
			for (int u = 0; u < 8; u++)
				v[u] = p[(size_t)(gg + u * NMFK_FFIN_WAVES) * tot];
#pragma unroll
			for (int u = 0; u < 8; u++)
				s += v[u];
		}
		for (; gg < G; gg += NMFK_FFIN_WAVES)
			s += p[(size_t)gg * tot];
	}
	red[w][lane] = s;
	__syncthreads();
	if (w == 0 && e < tot) {
		float sum = 0.0f;
#pragma unroll
		for (int u = 0; u < NMFK_FFIN_WAVES; u++)
			sum += red[u][lane];
		const int c = (int)(e % KT);
		float v = 0.0f;
		if (c < k) {
			v = H[e] * sum;
			v = v / den[c];
		}
		H[e] = v;
	}
}

// --------------------------------------------------------------------------------------
// Argmax with the reference's thread layout (src/GPU_kernels.cu:2592-2830,
// src/matrix_operations.c:190-203): bw "threads" (lanes) each scan columns t, t+bw, t+2bw
// with a strict '>' starting from their first column (or from (-1, 0) without one); a
// tree over halves then keeps the lower lane unless lower < upper.  bw = 32 while the
// reference pitch (width rounded up to 32) is <= 96, else next_pow2(ceil(pitch/3)).
// One wave per row; widths up to 192 (bw <= 64).
// --------------------------------------------------------------------------------------

__host__ __device__ __forceinline__ int ref_block_width(int width)
{
	const int kp = (width + 31) & ~31;
	if (kp <= 96)
		return 32;
	const int need = (kp + 2) / 3;
	int bw = 1;
	while (bw < need)
		bw <<= 1;
	return bw;
}

// Lane t (t < bw) holds v[i] = row[t + i*bw] for i < 3 (bw >= ceil(pitch/3) >= width/3,
// so three items per lane always cover the row, as in the reference's 3 items per thread).
// Returns the winning column in every lane.
__device__ __forceinline__ int wave_argmax_ref(const float (&v)[3], int width, int bw, int lane)
{
	float mv = -1.0f;
	int mi = 0;
	if (lane < bw && lane < width) {
		mv = v[0];
		mi = lane;
		if (lane + bw < width && v[1] > mv) {
			mv = v[1];
			mi = lane + bw;
		}
		if (lane + 2 * bw < width && v[2] > mv) {
			mv = v[2];
			mi = lane + 2 * bw;
		}
	}
	for (int half = bw >> 1; half >= 1; half >>= 1) {
		const float ov = __shfl_down(mv, half, 64);
		const int oi = __shfl_down(mi, half, 64);
		if (lane < half && mv < ov) {
			mv = ov;
			mi = oi;
		}
	}
	return __shfl(mi, 0, 64);
}

__device__ __forceinline__ void load_row_items(const float *__restrict__ row, int width, int bw, int lane,
					       float (&v)[3])
{
#pragma unroll
	for (int i = 0; i < 3; i++) {
		const int c = lane + i * bw;
		v[i] = (lane < bw && c < width) ? row[c] : 0.0f;
	}
}

__global__ __launch_bounds__(256) void k_idx_max(const float *__restrict__ A, int height, int width,
						 int pitch, uint32_t *__restrict__ idx)
{
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
	if (row >= height)
		return;
	const int bw = ref_block_width(width);
	float v[3];
	load_row_items(A + (size_t)row * pitch, width, bw, lane, v);
	const int w = wave_argmax_ref(v, width, bw, lane);
	if (lane == 0)
		idx[row] = (uint32_t)w;
}

// --------------------------------------------------------------------------------------
// Convergence test: adjust(H) + classification + connectivity difference + stop rule
// (src/NMF_GPU.c:611-708), as two passes.
//   rows pass: each block clamps NMFK_TEST_ROWS rows of H (written back) and classifies
//              them (8 rows per wave, loaded together) into cls;
//   final pass (one block per job): the contingency table n_ab = #{j : c_j == a, l_j == b}
//              in LDS, then the difference  #{i<j : [c_i==c_j] != [l_i==l_j]}  as
//              sum_a C(r_a,2) + sum_b C(s_b,2) - 2 sum_ab C(n_ab,2)  (integer, exact),
//              last <- cls, and the stop rule.  A job whose test budget (max_tests) is
//              spent without converging is marked converged = 2 (done, not converged).
// --------------------------------------------------------------------------------------

__device__ __forceinline__ void test_rows_body(int bx, float *__restrict__ H, int m, int k, int kt,
					       uint32_t *__restrict__ cls)
{
	constexpr int R = NMFK_TEST_ROWS / (NMFK_WG / 64);
	const int lane = threadIdx.x & 63;
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
	const int bw = ref_block_width(k);
	const int j0 = bx * NMFK_TEST_ROWS + wave * R;
	float v[R][3];
#pragma unroll
	for (int i = 0; i < R; i++) {
		if (j0 + i < m)
			load_row_items(H + (size_t)(j0 + i) * kt, k, bw, lane, v[i]);
	}
#pragma unroll
	for (int i = 0; i < R; i++) {
		const int j = j0 + i;
		if (j >= m)
			break;
		float *row = H + (size_t)j * kt;
#pragma unroll
		for (int t = 0; t < 3; t++) {
			const int cc = lane + t * bw;
			if (lane < bw && cc < k) {
				v[i][t] = clamp_rmin(v[i][t]);
				row[cc] = v[i][t];
			}
		}
		const int c = wave_argmax_ref(v[i], k, bw, lane);
		if (lane == 0)
			cls[j] = (uint32_t)c;
	}
}

// tab: NMFK_KMAX * NMFK_KMAX uint32 of LDS; red: 2 * (NMFK_WG / 64) uint64 of LDS.
__device__ __forceinline__ void test_final_body(int m, const uint32_t *__restrict__ cls, uint32_t *__restrict__ last,
						nmfk_state *__restrict__ st, uint32_t *__restrict__ tab,
						unsigned long long *__restrict__ red)
{
	constexpr int NWV = NMFK_WG / 64;
	for (int e = threadIdx.x; e < NMFK_KMAX * NMFK_KMAX; e += NMFK_WG)
		tab[e] = 0;
	__syncthreads();
	for (int j = threadIdx.x; j < m; j += NMFK_WG) {
		const uint32_t c = cls[j], l = last[j];
		last[j] = c;
		atomicAdd(&tab[c * NMFK_KMAX + l], 1u);
	}
	__syncthreads();
	unsigned long long both = 0, t = 0;
	for (int e = threadIdx.x; e < NMFK_KMAX * NMFK_KMAX; e += NMFK_WG) {
		const unsigned long long v = tab[e];
		both += v * (v - (v > 0)) / 2;
	}
	if (threadIdx.x < NMFK_KMAX) {
		unsigned long long r = 0, c = 0;
		for (int b = 0; b < NMFK_KMAX; b++) {
			r += tab[threadIdx.x * NMFK_KMAX + b];
			c += tab[b * NMFK_KMAX + threadIdx.x];
		}
		t = r * (r - (r > 0)) / 2 + c * (c - (c > 0)) / 2;
	}
	for (int off = 32; off >= 1; off >>= 1) {
		both += __shfl_xor(both, off, 64);
		t += __shfl_xor(t, off, 64);
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	if (lane == 0) {
		red[wave] = both;
		red[NWV + wave] = t;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned long long B = 0, Tm = 0;
		for (int w = 0; w < NWV; w++) {
			B += red[w];
			Tm += red[NWV + w];
		}
		const unsigned long long diff = Tm - 2ull * B;
		st->last_diff = diff;
		st->ntests += 1;
		if (diff)
			st->inc = 0;
		else if (st->inc < st->stop_threshold)
			st->inc += 1;
		else {
			st->converged = 1;
			st->stop_test = st->ntests;
		}
		if (!st->converged && st->ntests >= st->max_tests)
			st->converged = 2;
	}
}

__global__ __launch_bounds__(NMFK_WG) void k_test_rows(float *__restrict__ H, int m, int k, int kt,
						       uint32_t *__restrict__ cls, const nmfk_state *__restrict__ st)
{
	if (st->converged)
		return;
	test_rows_body(blockIdx.x, H, m, k, kt, cls);
}

__global__ __launch_bounds__(NMFK_WG) void k_test_final(int m, const uint32_t *__restrict__ cls,
							uint32_t *__restrict__ last, nmfk_state *__restrict__ st)
{
	__shared__ uint32_t tab[NMFK_KMAX * NMFK_KMAX];
	__shared__ unsigned long long red[2 * (NMFK_WG / 64)];
	if (st->converged)
		return;
	test_final_body(m, cls, last, st, tab, red);
}

// --------------------------------------------------------------------------------------
// Batched launches (continuous batching and rank sweeps): one grid serves every active
// job.  Workgroups walk a tile table; a tile is one workgroup (bx, by) of the job's own
// single-job grid, run by the same body, so the job's arithmetic does not depend on what
// else is in the batch.  Bodies: VALU <KT, 4, 4> for KT 4/8, MFMA <KT, 1, 4> for KT 16/32
// (the single-job planner's defaults; the host checks every job's plan before batching).
// --------------------------------------------------------------------------------------

__device__ __forceinline__ half_args batch_half(const nmfk_bjob &J, int half)
{
	half_args h;
	const bool hh = half == 0;
	const nmfk_hplan &pl = hh ? J.planH : J.planW;
	h.U = hh ? J.V : J.Vt;
	h.ldu = hh ? J.ldv : J.ldvt;
	h.A = hh ? J.H : J.W;
	h.B = hh ? J.W : J.H;
	h.p = hh ? J.m : J.n;
	h.q = hh ? J.n : J.m;
	h.bchunk = pl.bchunk;
	h.nab = pl.nab;
	h.part = J.part;
	h.colB = hh ? J.colW : J.colH;
	h.ncolB = hh ? J.ncolW : J.ncolH;
	h.den = hh ? J.denW : J.denH;
	h.lb = (int)pl.lb;
	return h;
}

// VALU body of a batched half-step with lane groups (2, 4 or 8, from its plan).
template <int KT>
__device__ __forceinline__ void batch_valu_body(int bx, int by, const half_args &h, float *__restrict__ smem)
{
	switch (h.lb) {
	case 8:
		partial_valu_body<KT, 4, 4, 8>(bx, by, h, smem);
		break;
	case 4:
		partial_valu_body<KT, 4, 4, 4>(bx, by, h, smem);
		break;
	default:
		partial_valu_body<KT, 4, 4, 2>(bx, by, h, smem);
		break;
	}
}

// Six batched partial kernels, one per body class, so each runs at its own occupancy: the
// VALU bodies (KT 4/8) of half-steps without lane groups, those with lane groups (few rows
// of the updated factor: their variants need more registers), the matrix-core bodies of KT 16
// and of KT 32 in 1 x 4 waves, and the same in 2 x 2 waves (the plan of a lone job over a
// short contraction range); a batch launches each over its own tile table.
__host__ __device__ constexpr int batch_cls_kt(int cls)
{
	return (cls == 2 || cls == 4) ? 16 : 32;
}
__host__ __device__ constexpr int batch_cls_na(int cls)
{
	return cls >= 4 ? 2 : 1;
}

template <int CLS>
constexpr int smem_batch_partial()
{
	return CLS <= 1 ? smem_valu(8, 4, 4)
			: smem_mfma(batch_cls_kt(CLS), batch_cls_na(CLS), 4 / batch_cls_na(CLS));
}

template <int CLS>
__global__ __launch_bounds__(NMFK_WG, (CLS == 0 || CLS == 2 || CLS == 4 ? 4 : 1)) void k_batch_partial(const nmfk_bjob *__restrict__ jobs,
							   const nmfk_tile *__restrict__ tiles, int ntiles, int half)
{
	__shared__ float smem[smem_batch_partial<CLS>()];
	for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
		const nmfk_tile tl = tiles[t];
		const nmfk_bjob &J = jobs[tl.job];
		// a half-step of another class would not fit this kernel's body: never run it here.  The
		// 1 x 4 classes leave the shape to the tables (the check would cost the kt-16 body a
		// spill at its 128-register bound).
		bool fits;
		if constexpr (CLS <= 1)
			fits = J.kt <= 8 && ((half == 0 ? J.planH.lb : J.planW.lb) > 1) == (CLS == 1);
		else if constexpr (CLS <= 3)
			fits = J.kt == (uint32_t)batch_cls_kt(CLS);
		else
			fits = J.kt == (uint32_t)batch_cls_kt(CLS) && (half == 0 ? J.planH.ap : J.planW.ap) == 2u;
		if (!fits || !J.live || J.st->converged)
			continue;
		const half_args h = batch_half(J, half);
		if constexpr (CLS == 0) {
			if (J.kt == 4)
				partial_valu_body<4, 4, 4, 1>(tl.bx, tl.by, h, smem);
			else
				partial_valu_body<8, 4, 4, 1>(tl.bx, tl.by, h, smem);
		} else if constexpr (CLS == 1) {
			if (J.kt == 4)
				batch_valu_body<4>(tl.bx, tl.by, h, smem);
			else
				batch_valu_body<8>(tl.bx, tl.by, h, smem);
		} else {
			constexpr int KT = batch_cls_kt(CLS), NA = batch_cls_na(CLS), NB = 4 / NA;
			if constexpr (KT == 16) {
				// a rank up to 4 or 8 sums only its columns in GEMM1 (nmfk_plan_set_rank)
				const uint32_t ke = half == 0 ? J.planH.ke : J.planW.ke;
				if (ke == 4)
					partial_mfma_body<16, NA, NB, 4>(tl.bx, tl.by, h, smem);
				else if (ke == 8)
					partial_mfma_body<16, NA, NB, 8>(tl.bx, tl.by, h, smem);
				else
					partial_mfma_body<16, NA, NB>(tl.bx, tl.by, h, smem);
			} else {
				partial_mfma_body<KT, NA, NB>(tl.bx, tl.by, h, smem);
			}
		}
		__syncthreads();   // the next tile reuses the LDS
	}
}

__global__ __launch_bounds__(NMFK_WG) void k_batch_finish(const nmfk_bjob *__restrict__ jobs,
							  const nmfk_tile *__restrict__ tiles, int ntiles, int half,
							  int clamp)
{
	__shared__ float smem[smem_fin(NMFK_BATCH_KTMAX)];
	for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
		const nmfk_tile tl = tiles[t];
		const nmfk_bjob &J = jobs[tl.job];
		if (!J.live || J.st->converged)
			continue;
		const bool hh = half == 0;
		const float *part = J.part;
		const int S = hh ? J.planH.S : J.planW.S;
		float *A = hh ? J.H : J.W;
		const int p = hh ? J.m : J.n;
		const float *den = hh ? J.denW : J.denH;
		float *colA = hh ? J.colH : J.colW;
		const int cl = hh ? 0 : clamp;
		switch (J.kt) {
		case 4:
			finish_body<4>(tl.bx, part, S, A, p, J.k, den, colA, cl, smem);
			break;
		case 8:
			finish_body<8>(tl.bx, part, S, A, p, J.k, den, colA, cl, smem);
			break;
		case 16:
			finish_body<16>(tl.bx, part, S, A, p, J.k, den, colA, cl, smem);
			break;
		case 32:
			finish_body<32>(tl.bx, part, S, A, p, J.k, den, colA, cl, smem);
			break;
		}
		__syncthreads();
	}
}

__global__ __launch_bounds__(NMFK_WG) void k_batch_test_rows(const nmfk_bjob *__restrict__ jobs,
							     const nmfk_tile *__restrict__ tiles, int ntiles)
{
	for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
		const nmfk_tile tl = tiles[t];
		const nmfk_bjob &J = jobs[tl.job];
		if (!J.live || J.st->converged)
			continue;
		test_rows_body(tl.bx, J.H, J.m, J.k, J.kt, J.cls);
	}
}

__global__ __launch_bounds__(NMFK_WG) void k_batch_test_final(const nmfk_bjob *__restrict__ jobs, int njobs)
{
	__shared__ uint32_t tab[NMFK_KMAX * NMFK_KMAX];
	__shared__ unsigned long long red[2 * (NMFK_WG / 64)];
	for (int t = blockIdx.x; t < njobs; t += gridDim.x) {
		const nmfk_bjob &J = jobs[t];
		if (!J.live || J.st->converged)
			continue;
		test_final_body(J.m, J.cls, J.last, J.st, tab, red);
		__syncthreads();
	}
}

// out[t] = state of job t (zeros for an empty slot), for one host copy per poll.
__global__ void k_batch_gather_states(const nmfk_bjob *__restrict__ jobs, int njobs, nmfk_state *__restrict__ out)
{
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < njobs) {
		nmfk_state s = {};
		if (jobs[t].live)
			s = *jobs[t].st;
		out[t] = s;
	}
}

// C[i][j] += (cls[i] == cls[j]); integer adds, so the order of jobs does not matter.
__global__ __launch_bounds__(256) void k_consensus_add(const uint32_t *__restrict__ cls, int m, int32_t *__restrict__ C)
{
	const int j = blockIdx.x * 64 + (threadIdx.x & 63);
	const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
	if (i < m && j < m)
		C[(size_t)i * m + j] += (cls[i] == cls[j]) ? 1 : 0;
}

__global__ void k_state_reset(nmfk_state *st, uint32_t stop_threshold, uint32_t max_tests, uint32_t *last, int m)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i == 0) {
		st->inc = 0;
		st->converged = 0;
		st->ntests = 0;
		st->stop_test = 0;
		st->last_diff = 0;
		st->stop_threshold = stop_threshold;
		st->max_tests = max_tests;
	}
	for (int j = i; j < m; j += gridDim.x * blockDim.x)
		last[j] = 0;
}

// --------------------------------------------------------------------------------------
// Initialisation and synthetic inputs.
// --------------------------------------------------------------------------------------

__global__ void k_uniform(float *__restrict__ A, uint32_t height, uint32_t width, uint32_t pitch,
			  uint64_t seed, uint32_t tag)
{
	const size_t total = (size_t)height * pitch;
	for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
	     e += (size_t)gridDim.x * blockDim.x) {
		const uint32_t i = (uint32_t)(e / pitch), j = (uint32_t)(e % pitch);
		A[e] = (j < width) ? uniform01(seed, tag, (uint64_t)i * width + j) : 0.0f;
	}
}

__global__ void k_gen_lognormal(float *__restrict__ V, uint32_t n, uint32_t m, uint32_t ldv,
				uint64_t seed, double mu, double sigma)
{
#pragma clang fp contract(off)
	const size_t total = (size_t)n * ldv;
	for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
	     e += (size_t)gridDim.x * blockDim.x) {
		const uint32_t i = (uint32_t)(e / ldv), j = (uint32_t)(e % ldv);
		V[e] = (j < m) ? (float)exp(mu + sigma * normal01(seed, TAG_NOISE, (uint64_t)i * m + j)) : 0.0f;
	}
}

__global__ void k_gen_exp1(double *__restrict__ out, size_t count, uint64_t seed, uint32_t tag)
{
	for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count;
	     e += (size_t)gridDim.x * blockDim.x) {
		uint32_t o[4];
		philox_block(seed, tag, e, o);
		out[e] = -log(u53(o[0], o[1]));
	}
}

__global__ void k_gen_planted(float *__restrict__ V, uint32_t n, uint32_t m, uint32_t ldv, uint32_t kstar,
			      const double *__restrict__ Ws, const double *__restrict__ Hs,
			      uint64_t seed, double sigma)
{
#pragma clang fp contract(off)
	const size_t total = (size_t)n * ldv;
	for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
	     e += (size_t)gridDim.x * blockDim.x) {
		const uint32_t i = (uint32_t)(e / ldv), j = (uint32_t)(e % ldv);
		if (j >= m) {
			V[e] = 0.0f;
			continue;
		}
		double s = 0.0;
		for (uint32_t c = 0; c < kstar; c++)
			s += Ws[(size_t)i * kstar + c] * Hs[(size_t)c * m + j];
		const double z = normal01(seed, TAG_NOISE, (uint64_t)i * m + j);
		V[e] = (float)(s * exp(sigma * z));
	}
}

// 64x64 tiles through LDS.
__global__ __launch_bounds__(256) void k_transpose(const float *__restrict__ V, int n, int m, size_t ldv,
						   float *__restrict__ Vt, size_t ldvt)
{
	__shared__ float tile[64][65];
	const int j0 = blockIdx.x * 64;
	const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
	// 64-row tiles strided over the grid's y extent (at most 65535 blocks)
	for (int i0 = blockIdx.y * 64; i0 < n; i0 += gridDim.y * 64) {
		for (int r = ty; r < 64; r += 4) {
			const int i = i0 + r, j = j0 + tx;
			tile[r][tx] = (i < n && j < m) ? V[(size_t)i * ldv + j] : 0.0f;
		}
		__syncthreads();
		for (int r = ty; r < 64; r += 4) {
			const int j = j0 + r, i = i0 + tx;
			if (j < m && i < n)
				Vt[(size_t)j * ldvt + i] = tile[tx][r];
		}
		__syncthreads();   // the tile is rewritten by the next row block
	}
}

// --------------------------------------------------------------------------------------
// FP64 Frobenius and KL sums (dot_product_VWH, src/NMF_routines.c:1758-2794, plus the
// generalized KL objective).  Lane = column j; the block's row range is strided over.
// --------------------------------------------------------------------------------------

#define DOT_ROWSPLIT 64

template <int KT>
__global__ __launch_bounds__(256) void k_dot_kl(const float *__restrict__ V, size_t ldv,
						const float *__restrict__ W, const float *__restrict__ H,
						int n, int m, double *__restrict__ parts)
{
	const int j = blockIdx.x * 256 + threadIdx.x;
	const int rows_per = (n + gridDim.y - 1) / gridDim.y;
	const int i0 = blockIdx.y * rows_per, i1 = min(n, i0 + rows_per);
	double hj[KT];
#pragma unroll
	for (int c = 0; c < KT; c++)
		hj[c] = (j < m) ? (double)H[(size_t)j * KT + c] : 0.0;
	double sv = 0.0, sd = 0.0, sk = 0.0;
	if (j < m) {
		for (int i = i0; i < i1; i++) {
			const float *Wi = W + (size_t)i * KT;
			double wh = 0.0;
#pragma unroll
			for (int c = 0; c < KT; c++)
				wh += (double)Wi[c] * hj[c];
			const double v = (double)V[(size_t)i * ldv + j];
			const double d = v - wh;
			sv += v * v;
			sd += d * d;
			sk += (v > 0.0 ? v * log(v / wh) : 0.0) - v + wh;
		}
	}
	__shared__ double red[3][4];
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	sv = wave_sum_d(sv);
	sd = wave_sum_d(sd);
	sk = wave_sum_d(sk);
	if (lane == 0) {
		red[0][wave] = sv;
		red[1][wave] = sd;
		red[2][wave] = sk;
	}
	__syncthreads();
	if (threadIdx.x < 3) {
		const int q = threadIdx.x;
		parts[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 3 + q] =
			((red[q][0] + red[q][1]) + red[q][2]) + red[q][3];
	}
}

// --------------------------------------------------------------------------------------
// Elementwise and row kernels of the reference's matrix_operations.h.
// --------------------------------------------------------------------------------------

__global__ void k_div_sub(float *__restrict__ A, const float *__restrict__ B, size_t nitems, int div_op)
{
	for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < nitems;
	     e += (size_t)gridDim.x * blockDim.x) {
		const float a = A[e], b = B[e];
		A[e] = div_op ? (b / a) : (b - a);   // A = B / A  or  A = B - A
	}
}

__global__ void k_mul_div(float *__restrict__ A, const float *__restrict__ Aux, const float *__restrict__ acc,
			  uint32_t height, uint32_t width, uint32_t pitch)
{
	const size_t total = (size_t)height * pitch;
	for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
	     e += (size_t)gridDim.x * blockDim.x) {
		const uint32_t c = (uint32_t)(e % pitch);
		if (c < width) {
			const float t = A[e] * Aux[e];
			A[e] = t / acc[c];
		}
	}
}

__global__ void k_adjust(float *__restrict__ A, uint32_t height, uint32_t width, uint32_t pitch)
{
	const size_t total = (size_t)height * pitch;
	for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
	     e += (size_t)gridDim.x * blockDim.x) {
		const uint32_t c = (uint32_t)(e % pitch);
		if (c < width)
			A[e] = clamp_rmin(A[e]);
	}
}

// Column sums of an arbitrary height x width matrix: parts[split][c], rows of split
// summed ascending.
__global__ void k_colsum_generic(const float *__restrict__ A, uint32_t height, uint32_t width, uint32_t pitch,
				 uint32_t rows_per, float *__restrict__ parts)
{
	const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= width)
		return;
	const uint32_t i0 = blockIdx.y * rows_per;
	const uint32_t i1 = min(height, i0 + rows_per);
	float s = 0.0f;
	for (uint32_t i = i0; i < i1; i++)
		s += A[(size_t)i * pitch + c];
	parts[(size_t)blockIdx.y * width + c] = s;
}

// --------------------------------------------------------------------------------------
// Host-side launchers (extern "C").
// --------------------------------------------------------------------------------------

static inline unsigned grid1d(size_t total, unsigned block)
{
	size_t g = (total + block - 1) / block;
	if (g > 65535u * 16u)
		g = 65535u * 16u;
	return (unsigned)(g ? g : 1);
}

#define NMFK_KT_LIST(X) X(4) X(8) X(16) X(32) X(64)
#define NMFK_KT_MFMA_LIST(X) X(16) X(32) X(64)
#define NMFK_KT_BF16_LIST(X) X(32) X(64)

extern "C" uint32_t nmfk_kt(uint32_t k)
{
	// k <= 8: 4 or 8 columns (the VALU kernel's widths); 9..16 use one 16-column group of
	// the FP32 MFMA kernel; larger ranks are padded to the 32-column chunks of the BF16 MFMA
	// kernel (32 or 64), so every rank above 32 can run in BF16.
	if (k <= 8)
		return (k + 3u) & ~3u;
	if (k <= 16)
		return 16u;
	return (k + 31u) & ~31u;
}

// A job's padded rank.  Ranks up to 8 are padded to 16 for the FP32 MFMA kernel, unless V
// has fewer than 32 rows or columns: there the VALU kernel's lane groups (several rows of U
// per wave) fit the narrow factor and keep its 4 or 8 columns.  The MFMA kernel does 2-4x the
// padded flops but runs far more waves per SIMD than the VALU kernel's plans allow.  Per
// iteration, VALU kt 4/8 against MFMA kt 16: 20000 x 1000, k = 8: 96.1 against 46.8 us; k = 3:
// 75.8 / 46.3; 10000 x 500, k = 8: 57.6 / 28.5; 5000 x 200, k = 6: 37.1 / 21.8; 1000 x 60000,
// k = 5: 265 / 131; 5000 x 38, k = 3: 20.3 / 19.1; but 60000 x 20, k = 4: 27.8 / 33.3.
extern "C" uint32_t nmfk_job_kt(uint32_t n, uint32_t m, uint32_t k)
{
	if (k <= 8 && n >= 32u && m >= 32u)
		return 16u;
	return nmfk_kt(k);
}

static inline uint32_t default_ap(uint32_t kt)
{
	return kt <= 8 ? 4u : (kt <= 32 ? 2u : 1u);
}

// Launch-shape tunables (workgroups the partial kernels aim for, minimum steps per wave,
// most splits, and the MFMA workgroup shape NA x NB).
// The BF16 kernel (kt 64) reads at the memory system's rate for its access pattern with
// one four-wave workgroup per CU as well as with two.  At 50000 x 2000, k = 64 its H half
// takes 101 us with the 8 splits a 256-workgroup target gives, 105 us with 16 (512) and
// 115 us with the 32 of the FP32 target; 12 splits (384) take 127 us.
static int g_tune[NMFK_TUNE_COUNT] = {512, 1024, 8, 64, 1, 4, 256, 2048};

static bool mfma_shape_ok(int na, int nb)
{
	return (na == 1 && nb == 4) || (na == 2 && nb == 2) || (na == 4 && nb == 1) || (na == 2 && nb == 4) ||
	       (na == 4 && nb == 2);
}

extern "C" int nmfk_tune_set(int key, int value)
{
	if (key < 0 || key >= NMFK_TUNE_COUNT || value < 1)
		return (int)hipErrorInvalidValue;
	g_tune[key] = value;
	return 0;
}

extern "C" int nmfk_tune_get(int key)
{
	return (key >= 0 && key < NMFK_TUNE_COUNT) ? g_tune[key] : -1;
}

// Rows of U per split for S splits of q rows, whole wave steps.
static uint32_t split_rows(uint32_t q, uint32_t S, uint32_t rows_per_step)
{
	const uint32_t bc = (q + S - 1u) / S;
	return (bc + rows_per_step - 1u) / rows_per_step * rows_per_step;
}

// Split count of a lone FP32 matrix-core job.  These launches are bound by their matrix work
// per SIMD, so a launch lasts about as long as its busiest CU needs: the workgroups that CU
// runs times the rows each streams (q / S).  The XCD-aware order gives each of the 8 XCDs
// ceil(tiles / 8) workgroups over its 32 CUs, of which at most R four-wave workgroups are
// resident per CU (4 at kt 16, 3 at kt 32, by their register use); beyond that they queue
// and even out.  Among S0 / 2 .. S0 splits this takes the least such load, the fewest splits
// on a tie (a shorter finish), keeping at least 240 workgroups (about one per CU).  At
// 20000 x 1000, k = 16 the H half's 40 splits become 30: partial 20.7-21.5 -> 18.6-19.1 us.
static uint32_t balanced_splits(uint32_t q, uint32_t nab, uint32_t S0, uint32_t rows_per_step, uint32_t kt)
{
	const uint32_t R = kt <= 16u ? 4u : 3u;
	auto load = [&](uint32_t s) {
		const uint32_t per = ((nab + 1u) * s + 7u) / 8u;
		const double w = per <= 32u * R ? (double)((per + 31u) / 32u) : per / 32.0;
		return w / s;
	};
	uint32_t best = S0;
	double bl = load(S0);
	for (uint32_t s = S0; s >= 1u && 2u * s >= S0; s--) {
		const uint32_t rows = split_rows(q, s, rows_per_step);
		const uint32_t se = (q + rows - 1u) / rows;   // splits that whole wave steps give
		if (se != S0 && (nab + 1u) * se < 240u)
			continue;
		const double l = load(se);
		if (l < bl * (1.0 - 1e-9) || (l <= bl * (1.0 + 1e-9) && se < best))
			best = se, bl = l;
	}
	return best;
}

extern "C" void nmfk_plan_half_path(uint32_t p, uint32_t q, uint32_t kt, int path, nmfk_hplan *pl)
{
	nmfk_plan_half_target(p, q, kt, path, 0, pl);
}

extern "C" void nmfk_plan_half_target(uint32_t p, uint32_t q, uint32_t kt, int path, uint32_t target_wgs,
				      nmfk_hplan *pl)
{
	if (path == NMFK_PATH_AUTO)
		path = kt >= 64 ? NMFK_PATH_MFMA_BF16 : (kt % 16 == 0) ? NMFK_PATH_MFMA32 : NMFK_PATH_VALU;
	if (path == NMFK_PATH_MFMA_BF16 && kt % 32 != 0)
		path = NMFK_PATH_MFMA32;
	if (path != NMFK_PATH_VALU && kt % 16 != 0)
		path = NMFK_PATH_VALU;
	// rows a per workgroup: VALU lanes own ap rows each, in 64 / lb lanes per group (the lb
	// groups split the rows of U); the MFMA kernel's NA side-by-side waves own 64 rows each
	// (ap = NA), its NB wave groups split the rows of U.
	int na = g_tune[NMFK_TUNE_MFMA_NA], nb = g_tune[NMFK_TUNE_MFMA_NB];
	if (!mfma_shape_ok(na, nb))
		na = 1, nb = 4;
	// BF16 over a short contraction range (fewer than 4096 rows of U, e.g. the W half of a
	// wide V): two waves side by side and two splitting the rows, unless the shape tunables
	// were changed.  At 50000 x 2000, k = 64 the W half takes 125.5 us against 132.7 us.
	if (path == NMFK_PATH_MFMA_BF16 && na == 1 && nb == 4 && q < 4096u)
		na = 2, nb = 2;
	// Three rules for an FP32 job at kt 16 or 32 planned on its own (no workgroup target, as
	// batched replans give) with the matrix-core tunables at their defaults.  First, over
	// 1000-4095 rows of U (the W half of V 20000 x 1000): 2 x 2 waves with at least 12 steps
	// per wave group, where that still gives 240 workgroups or more.  At 20000 x 1000, k = 16
	// the W partial takes 18.5-18.7 us in 3 splits against 20.6-21.8 us in 2; shorter ranges
	// (600 rows) lose 1-2 us, and so do jobs too small to fill the GPU either way (V 2500 x
	// 1000: 63 workgroups against 85).  Second, the split count is balanced over the CUs
	// (balanced_splits).  Third, a job with fewer than 240 workgroups takes 4 steps (below).
	const bool lone_fp32 = path == NMFK_PATH_MFMA32 && kt <= 32u && target_wgs == 0 && na == 1 && nb == 4 &&
			       g_tune[NMFK_TUNE_MIN_STEPS] == 8 && g_tune[NMFK_TUNE_WGS_MFMA] == 1024 &&
			       g_tune[NMFK_TUNE_MAX_SPLITS] == 64;
	const bool valu = path == NMFK_PATH_VALU;
	auto fill = [&](int na_, int nb_, int min_steps) {
		const uint32_t ap = valu ? default_ap(kt) : (uint32_t)na_;
		const uint32_t nw = valu ? 4u : (uint32_t)(na_ * nb_);
		// VALU lane groups: as many as leave every lane a row of A (p <= 64 ap / lb)
		uint32_t lb = 1;
		if (valu)
			while (lb < 8 && p <= 64u * ap / (2u * lb))
				lb *= 2;
		const uint32_t rows_per_step = valu ? lb : path == NMFK_PATH_MFMA_BF16 ? 32u : 16u;   // rows of U per wave step
		const uint32_t steppers = valu ? nw : (uint32_t)nb_;       // wave groups splitting the rows
		const uint32_t rows_per_wg = valu ? 64u / lb * ap : 64u * ap;   // rows a per workgroup
		const uint32_t nab = (p + rows_per_wg - 1u) / rows_per_wg;
		const uint32_t target =
			target_wgs ? target_wgs
				   : (uint32_t)g_tune[valu ? NMFK_TUNE_WGS_VALU
							   : path == NMFK_PATH_MFMA_BF16 ? NMFK_TUNE_WGS_BF16 : NMFK_TUNE_WGS_MFMA];
		uint32_t S = (target + nab - 1u) / nab;
		const uint32_t min_rows = steppers * rows_per_step * (uint32_t)min_steps;
		const uint32_t smax = (q + min_rows - 1u) / min_rows;
		if (S > smax)
			S = smax;
		if (S > (uint32_t)g_tune[NMFK_TUNE_MAX_SPLITS])
			S = (uint32_t)g_tune[NMFK_TUNE_MAX_SPLITS];
		if (S < 1)
			S = 1;
		uint32_t bchunk = split_rows(q, S, rows_per_step);
		S = (q + bchunk - 1u) / bchunk;
		if (lone_fp32) {
			S = balanced_splits(q, nab, S, rows_per_step, kt);
			bchunk = split_rows(q, S, rows_per_step);
			S = (q + bchunk - 1u) / bchunk;
		}
		pl->nab = nab;
		pl->S = S;
		pl->bchunk = bchunk;
		pl->ap = ap;
		pl->nw = nw;
		pl->path = (uint32_t)path;
		pl->lb = lb;
		pl->ke = kt;
	};
	if (lone_fp32 && q >= 1000u && q < 4096u) {
		fill(2, 2, 12);
		if ((pl->nab + 1u) * pl->S >= 240u)
			return;
	}
	fill(na, nb, g_tune[NMFK_TUNE_MIN_STEPS]);
	// Third, a lone FP32 job too small for 240 workgroups (the balanced plans' floor) at 8 steps
	// per wave group takes 4: twice the splits over the idle CUs.  Per iteration: V 5000 x 38,
	// k = 3 14.8 against 15.8 us; 2500 x 1000, k = 16 20.7-20.9 against 26.4; 3000 x 300,
	// k = 24 26.7 against 31.8; 1000 x 500, k = 6 18.9 against 22.5; 10000 x 500, k = 8 even.
	// A plan balanced to 240-255 workgroups (about one per CU) keeps its 8 steps.
	if (lone_fp32 && (pl->nab + 1u) * pl->S < 240u)
		fill(na, nb, 4);
}

extern "C" void nmfk_plan_set_rank(nmfk_hplan *pl, uint32_t kt, uint32_t k)
{
	pl->ke = kt;
	if (pl->path == NMFK_PATH_MFMA32 && kt == 16u && k <= 8u)
		pl->ke = k <= 4u ? 4u : 8u;
}

extern "C" void nmfk_plan_half(uint32_t p, uint32_t q, uint32_t kt, nmfk_hplan *pl)
{
	nmfk_plan_half_path(p, q, kt, NMFK_PATH_AUTO, pl);
}

// (nab + 1) x S tiles (+ the denominator column) as a 1-D grid padded to a multiple of 8
// (see xcd_tile).
static inline dim3 xcd_grid(const nmfk_hplan *pl)
{
	const unsigned total = (pl->nab + 1) * pl->S;
	return dim3((total + 7u) & ~7u);
}

template <int KT, int AP>
static int launch_partial_t(const float *U, size_t ldu, const float *A, const float *B, uint32_t p,
			    uint32_t q, const nmfk_hplan *pl, float *part, const float *colB, uint32_t ncolB,
			    float *den, const nmfk_state *st, hipStream_t stream)
{
	const dim3 grid = xcd_grid(pl);
	if (pl->nw != 4)
		return (int)hipErrorInvalidValue;
	const half_args h = {U, ldu, A, B, (int)p, (int)q, (int)pl->bchunk, (int)pl->nab, part, colB, (int)ncolB, den,
			     (int)pl->lb};
	switch (pl->lb) {
	case 1:
		hipLaunchKernelGGL((k_partial<KT, AP, 4, 1>), grid, dim3(256), 0, stream, h, st, (int)pl->S);
		break;
	case 2:
		hipLaunchKernelGGL((k_partial<KT, AP, 4, 2>), grid, dim3(256), 0, stream, h, st, (int)pl->S);
		break;
	case 4:
		hipLaunchKernelGGL((k_partial<KT, AP, 4, 4>), grid, dim3(256), 0, stream, h, st, (int)pl->S);
		break;
	case 8:
		hipLaunchKernelGGL((k_partial<KT, AP, 4, 8>), grid, dim3(256), 0, stream, h, st, (int)pl->S);
		break;
	default:
		return (int)hipErrorInvalidValue;
	}
	return (int)hipGetLastError();
}

template <int KT>
static int launch_partial_mfma_t(const float *U, size_t ldu, const float *A, const float *B, uint32_t p,
				 uint32_t q, const nmfk_hplan *pl, float *part, const float *colB, uint32_t ncolB,
				 float *den, const nmfk_state *st, hipStream_t stream)
{
	const dim3 grid = xcd_grid(pl);
	const int na = (int)pl->ap, nb = pl->ap ? (int)(pl->nw / pl->ap) : 0;
	// byte offsets within one 16-row block of U are 32-bit
	if (pl->bchunk % 16 != 0 || !mfma_shape_ok(na, nb) || (size_t)16 * ldu * 4 > 0x7fffffffull)
		return (int)hipErrorInvalidValue;
	const half_args h = {U, ldu, A, B, (int)p, (int)q, (int)pl->bchunk, (int)pl->nab, part, colB, (int)ncolB, den};
	// A rank up to 4 or 8 at KT 16 sums only its columns in GEMM1 (the default shapes have
	// these bodies; any other shape sums all 16, which is as correct).
	if constexpr (KT == 16) {
		const int ke = (pl->ke == 4 || pl->ke == 8) ? (int)pl->ke : 16;
#define NMFK_MFMA_KE(NA_, NB_, KE_)                                                                      \
	if (na == NA_ && nb == NB_ && ke == KE_) {                                                     \
		hipLaunchKernelGGL((k_partial_mfma<KT, NA_, NB_, KE_>), grid, dim3(64 * NA_ * NB_), 0, stream, h, st, \
				   (int)pl->S);                                                              \
		return (int)hipGetLastError();                                                         \
	}
		NMFK_MFMA_KE(1, 4, 4)
		NMFK_MFMA_KE(1, 4, 8)
		NMFK_MFMA_KE(2, 2, 4)
		NMFK_MFMA_KE(2, 2, 8)
#undef NMFK_MFMA_KE
	}
#define NMFK_MFMA_SHAPE(NA_, NB_)                                                                        \
	if (na == NA_ && nb == NB_) {                                                                  \
		hipLaunchKernelGGL((k_partial_mfma<KT, NA_, NB_>), grid, dim3(64 * NA_ * NB_), 0, stream, h, st, (int)pl->S); \
		return (int)hipGetLastError();                                                         \
	}
	NMFK_MFMA_SHAPE(1, 4)
	NMFK_MFMA_SHAPE(2, 2)
	NMFK_MFMA_SHAPE(4, 1)
	NMFK_MFMA_SHAPE(2, 4)
	NMFK_MFMA_SHAPE(4, 2)
#undef NMFK_MFMA_SHAPE
	return (int)hipErrorInvalidValue;
}

extern "C" int nmfk_launch_partial(const float *U, size_t ldu, const float *A, const float *B, uint32_t p,
				   uint32_t q, uint32_t kt, const nmfk_hplan *pl, float *part,
				   const float *colB, uint32_t ncolB, float *den,
				   const nmfk_state *st, hipStream_t stream)
{
	if (ldu % 4 != 0 || pl->path == NMFK_PATH_MFMA_BF16)   // BF16: nmfk_launch_partial_bf16
		return (int)hipErrorInvalidValue;
	if (pl->path == NMFK_PATH_MFMA32) {
		switch (kt) {
#define X(K)                                                                                          \
	case K:                                                                                       \
		return launch_partial_mfma_t<K>(U, ldu, A, B, p, q, pl, part, colB, ncolB, den, st, stream);
			NMFK_KT_MFMA_LIST(X)
#undef X
		default:
			return (int)hipErrorInvalidValue;
		}
	}
	if (pl->ap != default_ap(kt))
		return (int)hipErrorInvalidValue;
	switch (kt) {
#define X(K)                                                                                          \
	case K:                                                                                       \
		return launch_partial_t<K, (K <= 8 ? 4 : (K <= 32 ? 2 : 1))>(U, ldu, A, B, p, q, pl, part, \
									      colB, ncolB, den, st, stream);
		NMFK_KT_LIST(X)
#undef X
	default:
		return (int)hipErrorInvalidValue;
	}
}

extern "C" int nmfk_launch_finish_shadow(const float *part, uint32_t S, float *A, uint32_t p, uint32_t k,
					 uint32_t kt, const float *den, float *colA, int clamp, const nmfk_state *st,
					 uint16_t *Ab, hipStream_t stream)
{
	const unsigned g = (p + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS;
	if (Ab != nullptr && kt % 32 != 0)
		return (int)hipErrorInvalidValue;
	const shadow_args sh = {(__bf16 *)Ab};
	switch (kt) {
#define X(K)                                                                                           \
	case K:                                                                                        \
		hipLaunchKernelGGL((k_finish<K>), dim3(g), dim3(NMFK_WG), 0, stream, part, (int)S, A, (int)p, \
				   (int)k, den, colA, clamp, st, sh);                                   \
		return (int)hipGetLastError();
		NMFK_KT_LIST(X)
#undef X
	default:
		return (int)hipErrorInvalidValue;
	}
}

extern "C" int nmfk_fused_plan(uint32_t n, uint32_t m, uint32_t kt, uint32_t *G, uint32_t *nwv)
{
	if (kt != 16 || n == 0 || m == 0 || m > NMFK_FUSED_MAXW * NMFK_FUSED_COLS)
		return -1;
	const uint32_t nw = m <= 4 * NMFK_FUSED_COLS ? 4 : 8;
	const uint32_t nb = (n + NMFK_FUSED_RB - 1) / NMFK_FUSED_RB;
	const uint32_t target = std::max(1u, (uint32_t)g_tune[NMFK_TUNE_FUSED_WAVES] / nw);
	const uint32_t bpw = (nb + target - 1) / target;   // row blocks per workgroup, balanced
	*G = (nb + bpw - 1) / bpw;
	*nwv = nw;
	return 0;
}

extern "C" int nmfk_launch_fused(const float *V, size_t ldv, float *W, const float *H, uint32_t n, uint32_t m,
				 uint32_t k, uint32_t kt, uint32_t G, uint32_t nwv, float *part, float *colW, int upd_w,
				 int clamp, const nmfk_state *st, hipStream_t stream)
{
	if (kt != 16 || ldv < m || ldv % 4 != 0 || G == 0)
		return (int)hipErrorInvalidValue;
	fused_args fa;
	fa.V = V, fa.ldv = ldv, fa.W = W, fa.H = H, fa.part = part, fa.colW = colW;
	fa.n = (int)n, fa.m = (int)m, fa.k = (int)k;
	fa.nblocks = (int)((n + NMFK_FUSED_RB - 1) / NMFK_FUSED_RB);
	fa.G = (int)G, fa.upd_w = upd_w, fa.clamp = clamp;
	if (m > nwv * NMFK_FUSED_COLS)
		return (int)hipErrorInvalidValue;
	constexpr int JT = NMFK_FUSED_COLS / 16;
	if (nwv == 4)
		hipLaunchKernelGGL((k_fused_mfma<16, 4, JT>), dim3(G), dim3(64 * 4), 0, stream, fa, st);
	else if (nwv == 8)
		hipLaunchKernelGGL((k_fused_mfma<16, 8, JT>), dim3(G), dim3(64 * 8), 0, stream, fa, st);
	else
		return (int)hipErrorInvalidValue;
	return (int)hipGetLastError();
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_fused_finish(const float *part, uint32_t G, const float *colW, float *H, uint32_t m,
					uint32_t k, uint32_t kt, const nmfk_state *st, hipStream_t stream)
{
	if (kt != 16)
		return (int)hipErrorInvalidValue;
	const unsigned blocks = (unsigned)(((size_t)m * kt + 63) / 64);
	hipLaunchKernelGGL((k_fused_finish<16>), dim3(blocks), dim3(64 * NMFK_FFIN_WAVES), 0, stream, part, (int)G,
			   colW, H, (int)m, (int)k, st);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_finish(const float *part, uint32_t S, float *A, uint32_t p, uint32_t k, uint32_t kt,
				  const float *den, float *colA, int clamp, const nmfk_state *st,
				  hipStream_t stream)
{
	return nmfk_launch_finish_shadow(part, S, A, p, k, kt, den, colA, clamp, st, nullptr, stream);
}

extern "C" int nmfk_launch_shadow(const float *A, uint32_t p, uint32_t kt, uint16_t *Ab, hipStream_t stream)
{
	const unsigned g = (p + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS;
	switch (kt) {
#define X(K)                                                                                           \
	case K:                                                                                        \
		hipLaunchKernelGGL((k_shadow<K>), dim3(g), dim3(NMFK_WG), 0, stream, A, (int)p, (__bf16 *)Ab); \
		return (int)hipGetLastError();
		NMFK_KT_BF16_LIST(X)
#undef X
	default:
		return (int)hipErrorInvalidValue;
	}
}

template <int KT>
static int launch_partial_bf16_t(const float *U, size_t ldu, const float *A, const uint16_t *Bb, uint32_t p,
				 uint32_t q, const nmfk_hplan *pl, float *part, const float *colB, uint32_t ncolB,
				 float *den, const nmfk_state *st, hipStream_t stream)
{
	const dim3 grid = xcd_grid(pl);
	const int na = (int)pl->ap, nb = pl->ap ? (int)(pl->nw / pl->ap) : 0;
	if (pl->bchunk % 32 != 0 || !mfma_shape_ok(na, nb))
		return (int)hipErrorInvalidValue;
	const half_args h = {U, ldu, A, nullptr, (int)p, (int)q, (int)pl->bchunk, (int)pl->nab, part, colB, (int)ncolB, den};
	const bf16_args b = {(const __bf16 *)Bb};
#define NMFK_BF16_SHAPE(NA_, NB_)                                                                        \
	if (na == NA_ && nb == NB_) {                                                                  \
		hipLaunchKernelGGL((k_partial_bf16<KT, NA_, NB_>), grid, dim3(64 * NA_ * NB_), 0, stream, h, b, st, (int)pl->S); \
		return (int)hipGetLastError();                                                         \
	}
	NMFK_BF16_SHAPE(1, 4)
	NMFK_BF16_SHAPE(2, 2)
	NMFK_BF16_SHAPE(4, 1)
	NMFK_BF16_SHAPE(2, 4)
	NMFK_BF16_SHAPE(4, 2)
#undef NMFK_BF16_SHAPE
	return (int)hipErrorInvalidValue;
}

extern "C" int nmfk_launch_partial_bf16(const float *U, size_t ldu, const float *A, const uint16_t *Bb, uint32_t p,
					uint32_t q, uint32_t kt, const nmfk_hplan *pl, float *part, const float *colB,
					uint32_t ncolB, float *den, const nmfk_state *st, hipStream_t stream)
{
	if (ldu % 4 != 0 || pl->path != NMFK_PATH_MFMA_BF16)
		return (int)hipErrorInvalidValue;
	switch (kt) {
#define X(K)                                                                                          \
	case K:                                                                                       \
		return launch_partial_bf16_t<K>(U, ldu, A, Bb, p, q, pl, part, colB, ncolB, den, st, stream);
		NMFK_KT_BF16_LIST(X)
#undef X
	default:
		return (int)hipErrorInvalidValue;
	}
}

extern "C" int nmfk_launch_colsum_parts(const float *A, uint32_t p, uint32_t kt, float *colA, hipStream_t stream)
{
	const unsigned g = (p + NMFK_FIN_ROWS - 1) / NMFK_FIN_ROWS;
	switch (kt) {
#define X(K)                                                                                      \
	case K:                                                                                   \
		hipLaunchKernelGGL((k_colsum_parts<K>), dim3(g), dim3(NMFK_WG), 0, stream, A, (int)p, colA); \
		return (int)hipGetLastError();
		NMFK_KT_LIST(X)
#undef X
	default:
		return (int)hipErrorInvalidValue;
	}
}

extern "C" int nmfk_launch_den(const float *parts, uint32_t nparts, uint32_t kt, float *den, hipStream_t stream)
{
	switch (kt) {
#define X(K)                                                                                      \
	case K:                                                                                   \
		hipLaunchKernelGGL((k_den<K>), dim3(1), dim3(256), 0, stream, parts, (int)nparts, den);  \
		return (int)hipGetLastError();
		NMFK_KT_LIST(X)
#undef X
	default:
		return (int)hipErrorInvalidValue;
	}
}

extern "C" int nmfk_launch_test(float *H, uint32_t m, uint32_t k, uint32_t kt, uint32_t *cls, uint32_t *last,
				nmfk_state *st, hipStream_t stream)
{
	if (k > NMFK_KMAX || k == 0)
		return (int)hipErrorInvalidValue;
	const unsigned g = (m + NMFK_TEST_ROWS - 1) / NMFK_TEST_ROWS;
	hipLaunchKernelGGL(k_test_rows, dim3(g), dim3(NMFK_WG), 0, stream, H, (int)m, (int)k, (int)kt, cls,
			   (const nmfk_state *)st);
	hipLaunchKernelGGL(k_test_final, dim3(1), dim3(NMFK_WG), 0, stream, (int)m, (const uint32_t *)cls, last, st);
	return (int)hipGetLastError();
}

static inline unsigned batch_grid(uint32_t ntiles, uint32_t grid)
{
	if (grid == 0 || grid > ntiles)
		grid = ntiles;
	return grid ? grid : 1;
}

extern "C" int nmfk_launch_batch_partial(const nmfk_bjob *jobs, const nmfk_tile *tiles, uint32_t ntiles, int half,
					 int cls, uint32_t grid, hipStream_t stream)
{
	if (ntiles == 0)
		return 0;
	const dim3 g(batch_grid(ntiles, grid));
	switch (cls) {
	case 0:
		hipLaunchKernelGGL((k_batch_partial<0>), g, dim3(NMFK_WG), 0, stream, jobs, tiles, (int)ntiles, half);
		break;
	case 1:
		hipLaunchKernelGGL((k_batch_partial<1>), g, dim3(NMFK_WG), 0, stream, jobs, tiles, (int)ntiles, half);
		break;
	case 2:
		hipLaunchKernelGGL((k_batch_partial<2>), g, dim3(NMFK_WG), 0, stream, jobs, tiles, (int)ntiles, half);
		break;
	case 3:
		hipLaunchKernelGGL((k_batch_partial<3>), g, dim3(NMFK_WG), 0, stream, jobs, tiles, (int)ntiles, half);
		break;
	case 4:
		hipLaunchKernelGGL((k_batch_partial<4>), g, dim3(NMFK_WG), 0, stream, jobs, tiles, (int)ntiles, half);
		break;
	case 5:
		hipLaunchKernelGGL((k_batch_partial<5>), g, dim3(NMFK_WG), 0, stream, jobs, tiles, (int)ntiles, half);
		break;
	default:
		return (int)hipErrorInvalidValue;
	}
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_batch_finish(const nmfk_bjob *jobs, const nmfk_tile *tiles, uint32_t ntiles, int half,
					int clamp, uint32_t grid, hipStream_t stream)
{
	if (ntiles == 0)
		return 0;
	hipLaunchKernelGGL(k_batch_finish, dim3(batch_grid(ntiles, grid)), dim3(NMFK_WG), 0, stream, jobs, tiles,
			   (int)ntiles, half, clamp);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_batch_test(const nmfk_bjob *jobs, const nmfk_tile *row_tiles, uint32_t nrow_tiles,
				      uint32_t njobs, uint32_t grid, hipStream_t stream)
{
	if (njobs == 0)
		return 0;
	hipLaunchKernelGGL(k_batch_test_rows, dim3(batch_grid(nrow_tiles, grid)), dim3(NMFK_WG), 0, stream, jobs,
			   row_tiles, (int)nrow_tiles);
	hipLaunchKernelGGL(k_batch_test_final, dim3(njobs), dim3(NMFK_WG), 0, stream, jobs, (int)njobs);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_batch_gather_states(const nmfk_bjob *jobs, uint32_t njobs, nmfk_state *out,
					       hipStream_t stream)
{
	if (njobs == 0)
		return 0;
	hipLaunchKernelGGL(k_batch_gather_states, dim3((njobs + 255) / 256), dim3(256), 0, stream, jobs, (int)njobs, out);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_consensus_add(const uint32_t *cls, uint32_t m, int32_t *C, hipStream_t stream)
{
	if (m == 0)
		return 0;
	hipLaunchKernelGGL(k_consensus_add, dim3((m + 63) / 64, (m + 3) / 4), dim3(256), 0, stream, cls, (int)m, C);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_state_reset(nmfk_state *st, uint32_t stop_threshold, uint32_t max_tests, uint32_t *last,
				       uint32_t m, hipStream_t stream)
{
	hipLaunchKernelGGL(k_state_reset, dim3(grid1d(m, 256)), dim3(256), 0, stream, st, stop_threshold, max_tests,
			   last, (int)m);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_uniform(float *A, uint32_t height, uint32_t width, uint32_t pitch, uint64_t seed,
				   uint32_t tag, hipStream_t stream)
{
	hipLaunchKernelGGL(k_uniform, dim3(grid1d((size_t)height * pitch, 256)), dim3(256), 0, stream, A, height,
			   width, pitch, seed, tag);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_gen_lognormal(float *V, uint32_t n, uint32_t m, uint32_t ldv, uint64_t seed, double mu,
					 double sigma, hipStream_t stream)
{
	hipLaunchKernelGGL(k_gen_lognormal, dim3(grid1d((size_t)n * ldv, 256)), dim3(256), 0, stream, V, n, m, ldv,
			   seed, mu, sigma);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_gen_planted(float *V, uint32_t n, uint32_t m, uint32_t ldv, uint32_t kstar,
				       uint64_t seed, double sigma, double *scratch, hipStream_t stream)
{
	double *Ws = scratch, *Hs = scratch + (size_t)n * kstar;
	hipLaunchKernelGGL(k_gen_exp1, dim3(grid1d((size_t)n * kstar, 256)), dim3(256), 0, stream, Ws,
			   (size_t)n * kstar, seed, TAG_WSTAR);
	hipLaunchKernelGGL(k_gen_exp1, dim3(grid1d((size_t)kstar * m, 256)), dim3(256), 0, stream, Hs,
			   (size_t)kstar * m, seed, TAG_HSTAR);
	hipLaunchKernelGGL(k_gen_planted, dim3(grid1d((size_t)n * ldv, 256)), dim3(256), 0, stream, V, n, m, ldv,
			   kstar, Ws, Hs, seed, sigma);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_transpose(const float *V, uint32_t n, uint32_t m, size_t ldv, float *Vt, size_t ldvt,
				     hipStream_t stream)
{
	const uint32_t rows = (n + 63) / 64;
	dim3 grid((m + 63) / 64, rows < 65535u ? rows : 65535u);
	hipLaunchKernelGGL(k_transpose, grid, dim3(256), 0, stream, V, (int)n, (int)m, ldv, Vt, ldvt);
	return (int)hipGetLastError();
}

extern "C" uint32_t nmfk_dot_kl_blocks(uint32_t n, uint32_t m)
{
	(void)n;
	return ((m + 255) / 256) * DOT_ROWSPLIT;
}

extern "C" int nmfk_launch_dot_kl(const float *V, size_t ldv, const float *W, const float *H, uint32_t n,
				  uint32_t m, uint32_t k, uint32_t kt, double *parts, hipStream_t stream)
{
	(void)k;
	dim3 grid((m + 255) / 256, DOT_ROWSPLIT);
	switch (kt) {
#define X(K)                                                                                      \
	case K:                                                                                   \
		hipLaunchKernelGGL((k_dot_kl<K>), grid, dim3(256), 0, stream, V, ldv, W, H, (int)n, (int)m, \
				   parts);                                                        \
		return (int)hipGetLastError();
		NMFK_KT_LIST(X)
#undef X
	default:
		return (int)hipErrorInvalidValue;
	}
}

extern "C" int nmfk_launch_div_sub(float *A, const float *B, size_t nitems, int div_operator, hipStream_t stream)
{
	hipLaunchKernelGGL(k_div_sub, dim3(grid1d(nitems, 256)), dim3(256), 0, stream, A, B, nitems, div_operator);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_mul_div(float *A, const float *Aux, const float *acc, uint32_t height, uint32_t width,
				   uint32_t pitch, hipStream_t stream)
{
	hipLaunchKernelGGL(k_mul_div, dim3(grid1d((size_t)height * pitch, 256)), dim3(256), 0, stream, A, Aux, acc,
			   height, width, pitch);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_adjust(float *A, uint32_t height, uint32_t width, uint32_t pitch, hipStream_t stream)
{
	hipLaunchKernelGGL(k_adjust, dim3(grid1d((size_t)height * pitch, 256)), dim3(256), 0, stream, A, height,
			   width, pitch);
	return (int)hipGetLastError();
}

extern "C" int nmfk_launch_idx_max(const float *A, uint32_t height, uint32_t width, uint32_t pitch, uint32_t *idx,
				   hipStream_t stream)
{
	if (width == 0 || width > 192 || width > pitch)
		return (int)hipErrorInvalidValue;
	hipLaunchKernelGGL(k_idx_max, dim3((height + 3) / 4), dim3(256), 0, stream, A, (int)height, (int)width,
			   (int)pitch, idx);
	return (int)hipGetLastError();
}

__global__ void k_sum_rows(const float *__restrict__ parts, uint32_t nparts, uint32_t width, float *__restrict__ out)
{
	const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= width)
		return;
	float s = 0.0f;
	for (uint32_t g = 0; g < nparts; g++)
		s += parts[(size_t)g * width + c];
	out[c] = s;
}

extern "C" int nmfk_launch_colsum_any(const float *A, uint32_t height, uint32_t width, uint32_t pitch, float *tmp,
				      size_t tmp_items, float *acc, hipStream_t stream)
{
	// Rows are cut into S splits of consecutive rows (S limited by the scratch <tmp>), each
	// split summed ascending by one thread per column, then the S partial rows summed
	// ascending: deterministic for a given (height, scratch size).
	uint32_t S = 1;
	if (tmp && width)
		S = (uint32_t)(tmp_items / width);
	const uint32_t want = (height + 255) / 256;
	if (S > want)
		S = want;
	if (S <= 1) {
		hipLaunchKernelGGL(k_colsum_generic, dim3((width + 63) / 64, 1), dim3(64), 0, stream, A, height, width,
				   pitch, height ? height : 1, acc);
		return (int)hipGetLastError();
	}
	const uint32_t rows_per = (height + S - 1) / S;
	S = (height + rows_per - 1) / rows_per;
	hipLaunchKernelGGL(k_colsum_generic, dim3((width + 63) / 64, S), dim3(64), 0, stream, A, height, width, pitch,
			   rows_per, tmp);
	hipLaunchKernelGGL(k_sum_rows, dim3((width + 63) / 64), dim3(64), 0, stream, tmp, S, width, acc);
	return (int)hipGetLastError();
}

// --------------------------------------------------------------------------------------
// Streaming probes: the measured HBM ceiling the roofline is quoted against, and the
// known-size reference the traffic counters are scaled by.
// --------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_stream_read(const float4 *__restrict__ src, size_t n4, float *__restrict__ sink)
{
	float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
	for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) {
		const float4 v = src[i];
		acc.x += v.x, acc.y += v.y, acc.z += v.z, acc.w += v.w;
	}
	const float s = (acc.x + acc.y) + (acc.z + acc.w);
	if (s == 1234.5f)   // never true for the zero-filled probe buffer; keeps the loads live
		sink[0] = s;
}

__global__ __launch_bounds__(256) void k_stream_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n4)
{
	for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256)
		dst[i] = src[i];
}

extern "C" int nmfk_launch_stream(int kind, const float *src, float *dst, size_t nfloats, unsigned blocks,
				  hipStream_t stream)
{
	const size_t n4 = nfloats / 4;
	if (kind == 0)
		hipLaunchKernelGGL(k_stream_read, dim3(blocks), dim3(256), 0, stream, (const float4 *)src, n4, dst);
	else
		hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, stream, (const float4 *)src, (float4 *)dst,
				   n4);
	return (int)hipGetLastError();
}
