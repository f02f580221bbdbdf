"""The drop-in API (GPU_setup.h / NMF_routines.h / matrix_operations.h / common.h) driven
exactly as the reference's nmf() drives it (src/NMF_GPU.c:312-924), through the C-ABI,
against the oracle; and the standalone matrix operations against their oracle forms."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

u32 = C.c_uint32
F32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
U32 = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
vp = C.c_void_p


@pytest.fixture(scope="module")
def lib(nmfk):
    L = nmfk.load()
    sigs = {
        "initialize_GPU": (C.c_size_t, [u32, u32]),
        "setup_GPU": (C.c_int, [C.c_size_t, C.c_bool]),
        "finalize_GPU_device": (C.c_int, []),
        "init_random": (C.c_int, [u32]),
        "destroy_random": (C.c_int, []),
        "set_random_values": (C.c_int, [vp, u32, u32, u32, C.c_char_p, vp, vp]),
        "update_H": (C.c_int, []),
        "update_W": (C.c_int, []),
        "matrix_adjust": (C.c_int, [vp, u32, u32, C.c_char_p, vp, vp]),
        "get_classification": (C.c_int, [vp, U32]),
        "get_difference": (C.c_size_t, [U32, U32, u32]),
        "dot_product_VWH": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "upload_matrix_partial": (C.c_int, [F32, u32, u32, u32, u32, C.c_char_p, u32, vp, vp, vp]),
        "upload_matrix": (C.c_int, [F32, u32, u32, vp, C.c_char_p, vp, vp]),
        "download_matrix": (C.c_int, [vp, C.c_size_t, C.c_size_t, vp, C.c_char_p, vp]),
        "matrix_to_row": (C.c_int, [vp, u32, u32, C.c_char_p, vp, vp, vp]),
        "matrix_div_sub": (C.c_int, [vp, vp, u32, u32, C.c_char_p, C.c_char_p, C.c_bool, vp, vp]),
        "matrix_mul_div": (C.c_int, [vp, vp, vp, u32, u32, C.c_char_p, vp]),
        "matrix_idx_max": (C.c_int, [vp, u32, u32, u32, C.c_char_p, vp, vp]),
        "sync_GPU": (C.c_int, [vp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


def g(lib, name, typ=C.c_uint32):
    return typ.in_dll(lib, name)


def gp(lib, name):
    return C.c_void_p.in_dll(lib, name).value


class BlockT(C.Structure):
    """block_t of GPU_setup.h: rows or columns of V per block on the device."""
    _fields_ = [("BL", u32 * 2), ("BLp", u32 * 2), ("num_steps", u32 * 2)]


def reference_style_nmf(lib, V, k, seed, nIters, T, S, mem_size=None, nmfk=None):
    """src/NMF_GPU.c:312-795 written against the drop-in API.  With <mem_size>, the host
    matrix Vrow (pinned, pitch Mp) is set before setup_GPU as init_V() does
    (src/NMF_GPU.c:212-303), and the device budget decides the blockwise plan."""
    n, m = V.shape
    free_mem = lib.initialize_GPU(0, k)
    assert free_mem > 0
    g(lib, "N").value, g(lib, "M").value = n, m
    host = None
    if mem_size is not None:
        Mp = (m + 3) // 4 * 4
        host = nmfk.PinnedArray((n, Mp))
        host.array[...] = 0
        host.array[:, :m] = V
        g(lib, "Mp").value = Mp
        C.c_void_p.in_dll(lib, "Vrow").value = host.ptr
    try:
        return _reference_style_loop(lib, V, k, seed, nIters, T, S, free_mem if mem_size is None else mem_size,
                                     host)
    finally:
        if host is not None:
            C.c_void_p.in_dll(lib, "Vrow").value = None
            host.close()


def _reference_style_loop(lib, V, k, seed, nIters, T, S, mem_size, host):
    n, m = V.shape
    assert lib.setup_GPU(mem_size, True) == 0
    Kp, Mp = g(lib, "Kp").value, g(lib, "Mp").value
    BLN = BlockT.in_dll(lib, "block_N").BL[0]
    stream_W, stream_H = gp(lib, "stream_W"), gp(lib, "stream_H")
    ev_W = C.addressof(C.c_void_p.in_dll(lib, "event_W"))
    try:
        assert lib.init_random(seed) == 0
        assert lib.set_random_values(gp(lib, "d_W"), n, k, Kp, b"d_W", stream_W, ev_W) == 0
        assert lib.set_random_values(gp(lib, "d_H"), m, k, Kp, b"d_H", stream_H, None) == 0
        assert lib.destroy_random() == 0
        if host is None:
            Vrow = np.zeros((n, Mp), np.float32)
            Vrow[:, :m] = V
        else:
            Vrow = host.array
        ev_Vrow = C.c_void_p.in_dll(lib, "event_Vrow").value
        # the driver uploads the first block of rows (src/NMF_GPU.c:494)
        assert lib.upload_matrix_partial(Vrow, BLN, Mp, 0, 0, b"Vrow", Mp, gp(lib, "d_Vrow"),
                                         gp(lib, "stream_Vrow"), ev_Vrow) == 0
        cls = np.zeros(m, np.uint32)
        last = np.zeros(m, np.uint32)
        inc, it, converged = 0, 0, False
        niter_conv, niter_rem = nIters // T, nIters % T
        while it < niter_conv:
            for _ in range(T):
                assert lib.update_H() == 0
                assert lib.update_W() == 0
            assert lib.matrix_adjust(gp(lib, "d_H"), m, Kp, b"d_H", stream_H, None) == 0
            assert lib.matrix_adjust(gp(lib, "d_W"), n, Kp, b"d_W", stream_W, ev_W) == 0
            assert lib.get_classification(gp(lib, "d_classification"), cls) == 0
            diff = lib.get_difference(cls, last, m)
            cls, last = last, cls
            if diff:
                inc = 0
            elif inc < S:
                inc += 1
            else:
                it += 1
                converged = True
                break
            it += 1
        if not converged:
            for _ in range(niter_rem):
                assert lib.update_H() == 0
                assert lib.update_W() == 0
        iters = it * T if converged else nIters
        H = np.zeros((m, Kp), np.float32)
        W = np.zeros((n, Kp), np.float32)
        assert lib.download_matrix(H.ctypes.data, m * Kp, 4, gp(lib, "d_H"), b"d_H", stream_H) == 0
        assert lib.download_matrix(W.ctypes.data, n * Kp, 4, gp(lib, "d_W"), b"d_W", stream_W) == 0
        dv, dvwh = C.c_float(), C.c_float()
        assert lib.dot_product_VWH(C.byref(dv), C.byref(dvwh)) == 0
        return {"W": W[:, :k], "H": H[:, :k], "iterations": iters, "converged": converged,
                "cls": last.copy(), "dist": float(np.sqrt(dvwh.value) / np.sqrt(dv.value)), "BLN": BLN}
    finally:
        assert lib.finalize_GPU_device() == 0


@pytest.mark.parametrize("n,m,k", [(800, 60, 3), (1500, 90, 12)])
def test_reference_style_driver_matches_oracle_and_job_api(lib, nmfk, orc, n, m, k):
    V = orc.gen_lognormal(n, m, 31, 0.0, 1.0)
    r = reference_style_nmf(lib, V, k, seed=5, nIters=500, T=10, S=6)
    W0, H0 = orc.init_WH(n, m, k, 5)
    o = orc.nmf(V, W0, H0, 500, 10, 6, fp64=True)
    assert r["iterations"] == o["iterations"] and r["converged"] == o["converged"]
    assert np.array_equal(r["cls"], orc.idx_max(o["H"]))
    assert rel_err(r["H"], o["H"]) < 1e-3 and rel_err(r["W"], o["W"]) < 1e-3
    dv, dvwh = orc.dot_VWH(V, o["W"], o["H"])
    assert abs(r["dist"] - np.sqrt(dvwh / dv)) < 1e-4
    with nmfk.Job(n, m, k) as j:
        j.set_V(V)
        j.init_random(5)
        jr = j.run(500, 10, 6)
        assert jr["iterations"] == r["iterations"]
        assert np.array_equal(j.get_H(), r["H"]) and np.array_equal(j.get_W(), r["W"])


def test_reference_style_driver_blockwise(lib, nmfk, orc):
    # A device budget below the resident job's size: setup_GPU() takes the largest block of
    # rows that fits and update_H/update_W stream the host matrix Vrow, as the reference's
    # blockwise mode.  Same stop iteration and classes as the FP64 oracle and the resident
    # drop-in run; factors within 1e-3.
    n, m, k = 1500, 90, 5
    V = orc.gen_lognormal(n, m, 32, 0.0, 1.0)
    budget = nmfk.load().nmfk_streamed_bytes(n, m, k, 256)
    r = reference_style_nmf(lib, V, k, seed=7, nIters=400, T=10, S=6, mem_size=budget, nmfk=nmfk)
    assert 64 <= r["BLN"] <= 256
    res = reference_style_nmf(lib, V, k, seed=7, nIters=400, T=10, S=6)
    assert res["BLN"] == n
    W0, H0 = orc.init_WH(n, m, k, 7)
    o = orc.nmf(V, W0, H0, 400, 10, 6, fp64=True)
    for x in (r, res):
        assert x["iterations"] == o["iterations"] and x["converged"] == o["converged"]
        assert np.array_equal(x["cls"], orc.idx_max(o["H"]))
    assert rel_err(r["W"], res["W"]) < 1e-3 and rel_err(r["H"], res["H"]) < 1e-3
    assert abs(r["dist"] - res["dist"]) < 1e-5


def test_matrix_operations_match_oracle(lib, orc):
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    rng = np.random.default_rng(11)
    h, p = 3000, 48

    def dev(a):
        d = C.c_void_p()
        assert hip.hipMalloc(C.byref(d), a.nbytes) == 0
        assert hip.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0
        return d

    def host(d, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(out.ctypes.data, d, out.nbytes, 2) == 0
        return out

    free_mem = lib.initialize_GPU(0, 4)
    assert free_mem > 0
    A = (rng.random((h, p), dtype=np.float32) + 0.5).astype(np.float32)
    B = (rng.random((h, p), dtype=np.float32) + 0.5).astype(np.float32)
    dA, dB = dev(A), dev(B)
    acc = dev(np.zeros(p, np.float32))
    tmp = dev(np.zeros((h // 4) * p, np.float32))
    assert lib.matrix_to_row(dA, h, p, b"A", tmp, acc, None) == 0
    assert rel_err(host(acc, p), orc.colsum(A)) < 1e-5
    assert lib.matrix_div_sub(dA, dB, h, p, b"A", b"B", True, None, None) == 0
    assert np.array_equal(host(dA, (h, p)), orc.div_sub(A, B, True, fp64=False))
    assert lib.matrix_div_sub(dB, dA, h, p, b"B", b"A", False, None, None) == 0
    assert np.array_equal(host(dB, (h, p)), orc.div_sub(B, B / A, False, fp64=False))
    A2 = host(dA, (h, p))
    Aux = (rng.random((h, p), dtype=np.float32) + 0.5).astype(np.float32)
    ac = (rng.random(p, dtype=np.float32) + 1.0).astype(np.float32)
    dAux, dac = dev(Aux), dev(ac)
    assert lib.matrix_mul_div(dA, dAux, dac, h, p, b"A", None) == 0
    assert np.array_equal(host(dA, (h, p)), orc.mul_div(A2, Aux, ac, fp64=False))
    X = rng.random((h, p)).astype(np.float32)
    X[::7, 3] = np.nan
    X[::5, 4] = -1.0
    X[::3, 5] = 1e-39
    dX = dev(X)
    assert lib.matrix_adjust(dX, h, p, b"X", None, None) == 0
    assert np.array_equal(host(dX, (h, p)), orc.adjust(X, fp64=False))
    py = 128
    for width in (3, 17, 32, 40, 64, 100, 128):
        Y = rng.integers(0, 5, (h, py)).astype(np.float32)          # many ties
        Y[::11, 1] = np.nan
        dY, dI = dev(Y), dev(np.zeros(h, np.uint32))
        assert lib.matrix_idx_max(dY, width, py, h, b"Y", None, dI) == 0
        got = host(dI, h, np.uint32)
        assert np.array_equal(got, orc.idx_max(np.ascontiguousarray(Y[:, :width]), fp64=False)), width
        hip.hipFree(dY), hip.hipFree(dI)
    for d in (dA, dB, acc, tmp, dAux, dac, dX):
        hip.hipFree(d)
