"""nmfk_dot_kl_noalloc through the C-ABI: the FP64 Frobenius / KL sums of caller-owned
device buffers, computed with the library's static device scratch array instead of any
allocation, must equal nmfk_job_dot_kl on the same buffers bit for bit.  Two calls in a
row, the second on a smaller input, so that the second reads none of the first's partials;
then the sizes the entry point must refuse."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


def _noalloc_sums(nmfk, hip, job, d_out):
    lib = nmfk.load()
    inf = job.info()
    job.sync()
    # a stale value in the output must not survive: fill it with NaN bytes first
    assert hip.hipMemset(d_out, 0xFF, 24) == 0
    assert hip.hipDeviceSynchronize() == 0
    rc = lib.nmfk_dot_kl_noalloc(inf.dV, inf.ldv, inf.dW, inf.dH, inf.n, inf.m, inf.k, d_out, inf.stream)
    assert rc == 0, nmfk.last_error()
    assert hip.hipStreamSynchronize(inf.stream) == 0
    out = np.empty(3, np.float64)
    assert hip.hipMemcpy(out.ctypes.data, d_out, 24, 2) == 0
    return out


@pytest.mark.parametrize("shapes", [((3000, 700, 12, 21), (900, 130, 5, 22))], ids=["large_then_small"])
def test_noalloc_dot_kl_equals_job_dot_kl(nmfk, shapes):
    hip = _hip()
    d_out = C.c_void_p()
    assert hip.hipMalloc(C.byref(d_out), 24) == 0
    try:
        for n, m, k, seed in shapes:
            print(f"nmfk_dot_kl_noalloc n={n} m={m} k={k} seed={seed}", flush=True)
            with nmfk.Job(n, m, k) as job:
                job.gen_V_lognormal(seed, 0.0, 1.0)
                job.init_random(seed + 100)
                job.run(40, 10, 40)                      # factors past their initial values
                got = _noalloc_sums(nmfk, hip, job, d_out)
                want = np.array(job.dot_kl(), np.float64)
            assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (n, m, k, got, want)
            assert got[0] > 0 and got[1] > 0 and got[2] > 0
    finally:
        hip.hipFree(d_out)


def test_noalloc_dot_kl_refuses_what_does_not_fit(nmfk):
    lib = nmfk.load()
    assert lib.nmfk_dot_kl_noalloc_max_bytes() == 1 << 20
    dummy = C.c_void_p(256)  # never dereferenced: every call below is refused before any launch
    # 256-column blocks x 64 row splits x 3 doubles must fit in 1 MiB: m <= 682 * 256
    assert lib.nmfk_dot_kl_noalloc(dummy, 682 * 256 + 1, dummy, dummy, 10, 682 * 256 + 1, 4, dummy, None) != 0
    assert "scratch" in nmfk.last_error()
    assert lib.nmfk_dot_kl_noalloc(None, 100, dummy, dummy, 10, 100, 4, dummy, None) != 0
    assert lib.nmfk_dot_kl_noalloc(dummy, 99, dummy, dummy, 10, 100, 4, dummy, None) != 0
    assert lib.nmfk_dot_kl_noalloc(dummy, 100, dummy, dummy, 10, 100, 65, dummy, None) != 0
