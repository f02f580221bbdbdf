// nmfk_noalloc.hip -- nmfk_dot_kl_noalloc: the FP64 Frobenius / KL sums of nmfk_job_dot_kl
// on caller-owned device buffers, with no device allocation.
//
// The per-block FP64 partials that a job keeps in its own buffer live here in one
// file-scope device array (`scratch`, NMFK_NOALLOC_SCRATCH_BYTES per device, held until
// the process ends).  Host code never uses the array's name as an address: on the first
// call on a device, with that device current, hipGetSymbolAddress gives that device's
// copy, and that pointer is what the kernels receive.  The partials are summed on the
// device, block by block in index order, which is the order nmfk_job_dot_kl sums them on
// the host, so both entry points return the same bits.
//
// Every call on one device shares that device's array: calls that overlap in time on one
// device (different streams or host threads) race and must be serialised by the caller.
// Each call writes every partial it later reads, so nothing has to start at zero.

#include <hip/hip_runtime.h>
#include <mutex>
#include <stdint.h>

#include "nmfk_kernels.h"

extern "C" void nmfk_set_error(const char *fmt, ...);

#define NMFK_NOALLOC_SCRATCH_BYTES ((size_t)1 << 20)
#define NMFK_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[NMFK_NOALLOC_SCRATCH_BYTES / sizeof(double)];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NMFK_NOALLOC_MAX_DEVICES];

// Device address of `scratch` on the current device, looked up once per device.
static int scratch_on_current_device(double **out)
{
	int dev = -1;
	hipError_t e = hipGetDevice(&dev);
	if (e != hipSuccess) {
		nmfk_set_error("nmfk_dot_kl_noalloc: hipGetDevice: %s", hipGetErrorString(e));
		return 3;
	}
	if (dev < 0 || dev >= NMFK_NOALLOC_MAX_DEVICES) {
		nmfk_set_error("nmfk_dot_kl_noalloc: device %d beyond the %d supported", dev, NMFK_NOALLOC_MAX_DEVICES);
		return 2;
	}
	std::lock_guard<std::mutex> guard(g_scratch_mu);
	if (g_scratch_ptr[dev] == nullptr) {
		void *p = nullptr;
		e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || p == nullptr) {
			nmfk_set_error("nmfk_dot_kl_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
				       hipGetErrorString(e));
			return 3;
		}
		g_scratch_ptr[dev] = static_cast<double *>(p);
	}
	*out = g_scratch_ptr[dev];
	return 0;
}

// out[q] = sum over blocks b = 0, 1, ... of parts[3 b + q], one thread per sum, in block order.
__global__ __launch_bounds__(64) void k_sum_parts_in_order(const double *__restrict__ parts, uint32_t nblocks,
							   double *__restrict__ out)
{
	const int q = threadIdx.x;
	if (q >= 3)
		return;
	double s = 0.0;
	for (uint32_t b = 0; b < nblocks; b++)
		s += parts[3 * (size_t)b + q];
	out[q] = s;
}

extern "C" size_t nmfk_dot_kl_noalloc_max_bytes(void)
{
	return NMFK_NOALLOC_SCRATCH_BYTES;
}

extern "C" int nmfk_dot_kl_noalloc(const float *V, size_t ldv, const float *W, const float *H, uint32_t n,
				   uint32_t m, uint32_t k, double *out, void *stream)
{
	if (V == nullptr || W == nullptr || H == nullptr || out == nullptr) {
		nmfk_set_error("nmfk_dot_kl_noalloc: null device pointer");
		return 2;
	}
	if (n == 0 || m == 0 || k == 0 || k > NMFK_KMAX || ldv < m) {
		nmfk_set_error("nmfk_dot_kl_noalloc: invalid shape n=%u m=%u k=%u ldv=%zu (1 <= k <= %d, ldv >= m)", n,
			       m, k, ldv, NMFK_KMAX);
		return 2;
	}
	const uint32_t nblocks = nmfk_dot_kl_blocks(n, m);
	const size_t need = (size_t)nblocks * 3 * sizeof(double);
	if (need > NMFK_NOALLOC_SCRATCH_BYTES) {
		nmfk_set_error("nmfk_dot_kl_noalloc: m=%u needs %zu bytes of scratch, more than the %zu reserved", m,
			       need, (size_t)NMFK_NOALLOC_SCRATCH_BYTES);
		return 2;
	}
	double *parts = nullptr;
	int rc = scratch_on_current_device(&parts);
	if (rc)
		return rc;
	hipStream_t s = static_cast<hipStream_t>(stream);
	int e = nmfk_launch_dot_kl(V, ldv, W, H, n, m, k, nmfk_job_kt(n, m, k), parts, s);
	if (e == 0) {
		hipLaunchKernelGGL(k_sum_parts_in_order, dim3(1), dim3(64), 0, s, parts, nblocks, out);
		e = (int)hipGetLastError();
	}
	if (e != 0) {
		nmfk_set_error("nmfk_dot_kl_noalloc: launch: %s", hipGetErrorString((hipError_t)e));
		return 3;
	}
	return 0;
}
