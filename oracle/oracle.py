"""ctypes binding of the CPU oracle (liboracle.so) and of the reference's own common.c
built in place under oracle/_ref/.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py, always as the checker, never as the thing measured.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_COMMON = HERE / "_ref" / "libref_common.so"
CPU_BASELINE = HERE / "nmf_cpu_baseline"

_F32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_F64 = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_U32 = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_SZ = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(LIB))
        u32, u64, dbl = C.c_uint32, C.c_uint64, C.c_double
        L.orc_philox4x32_10.argtypes = [_U32, _U32, _U32]
        L.orc_uniform01.restype = C.c_float
        L.orc_uniform01.argtypes = [u64, u32, u64]
        L.orc_uniform_fill.argtypes = [_F32, u32, u32, u32, u64, u32]
        L.orc_normal.restype = dbl
        L.orc_normal.argtypes = [u64, u32, u64]
        L.orc_gen_lognormal.argtypes = [_F32, u32, u32, u32, u64, dbl, dbl]
        L.orc_gen_planted.argtypes = [_F32, u32, u32, u32, u32, u64, dbl]
        L.orc_get_difference.restype = C.c_size_t
        L.orc_get_difference.argtypes = [_U32, _U32, u32]
        L.orc_stop_rule.restype = u32
        L.orc_stop_rule.argtypes = [_SZ, u32, u32]
        for sfx, P in (("f64", _F64), ("f32", _F32)):
            g = lambda n: getattr(L, f"{n}_{sfx}")
            g("orc_colsum").argtypes = [P, u32, u32, u32, P]
            g("orc_update_H").argtypes = [P, u32, u32, u32, P, P, u32, u32]
            g("orc_update_W").argtypes = [P, u32, u32, u32, P, P, u32, u32]
            g("orc_adjust").argtypes = [P, u32, u32, u32]
            g("orc_idx_max").argtypes = [P, u32, u32, u32, _U32]
            g("orc_mul_div").argtypes = [P, P, P, u32, u32, u32]
            g("orc_div_sub").argtypes = [P, P, u32, u32, u32, C.c_int]
            g("orc_dot_VWH").argtypes = [P, u32, u32, u32, P, P, u32, u32,
                                         C.POINTER(dbl), C.POINTER(dbl)]
            g("orc_kl_div").restype = dbl
            g("orc_kl_div").argtypes = [P, u32, u32, u32, P, P, u32, u32]
            g("orc_nmf").restype = C.c_int
            g("orc_nmf").argtypes = [P, u32, u32, u32, P, P, u32, u32, u32, u32, u32,
                                     _U32, C.POINTER(u32), C.POINTER(C.c_int),
                                     C.c_void_p, C.c_void_p, u32, C.POINTER(u32)]
        _lib = L
    return _lib


def _dt(fp64: bool):
    return np.float64 if fp64 else np.float32


def _sfx(fp64: bool):
    return "f64" if fp64 else "f32"


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def uniform_fill(height, width, seed, tag, pitch=None):
    pitch = width if pitch is None else pitch
    A = np.zeros((height, pitch), np.float32)
    lib().orc_uniform_fill(A, height, width, pitch, seed, tag)
    return A


def init_WH(n, m, k, seed):
    """W0 (tag 0) and H0 (tag 1), exactly as the GPU library draws them."""
    return uniform_fill(n, k, seed, 0), uniform_fill(m, k, seed, 1)


def gen_lognormal(n, m, seed, mu, sigma):
    V = np.zeros((n, m), np.float32)
    lib().orc_gen_lognormal(V, n, m, m, seed, mu, sigma)
    return V


def gen_planted(n, m, kstar, seed, sigma):
    V = np.zeros((n, m), np.float32)
    lib().orc_gen_planted(V, n, m, m, kstar, seed, sigma)
    return V


def get_difference(cls, last):
    cls = np.ascontiguousarray(cls, np.uint32)
    last = np.ascontiguousarray(last, np.uint32)
    return int(lib().orc_get_difference(cls, last, len(cls)))


def stop_rule(diffs, stop_threshold):
    d = np.ascontiguousarray(diffs, np.uint64)
    return int(lib().orc_stop_rule(d, len(d), stop_threshold))


def colsum(A, fp64=True):
    A = np.ascontiguousarray(A, _dt(fp64))
    acc = np.zeros(A.shape[1], _dt(fp64))
    getattr(lib(), f"orc_colsum_{_sfx(fp64)}")(A, A.shape[0], A.shape[1], A.shape[1], acc)
    return acc


def update_H(V, W, H, fp64=True):
    dt = _dt(fp64)
    V, W = np.ascontiguousarray(V, dt), np.ascontiguousarray(W, dt)
    H = np.array(H, dtype=dt, copy=True, order="C")
    n, m = V.shape
    k = W.shape[1]
    getattr(lib(), f"orc_update_H_{_sfx(fp64)}")(V, n, m, m, W, H, k, k)
    return H


def update_W(V, W, H, fp64=True):
    dt = _dt(fp64)
    V, H = np.ascontiguousarray(V, dt), np.ascontiguousarray(H, dt)
    W = np.array(W, dtype=dt, copy=True, order="C")
    n, m = V.shape
    k = W.shape[1]
    getattr(lib(), f"orc_update_W_{_sfx(fp64)}")(V, n, m, m, W, H, k, k)
    return W


def adjust(A, fp64=True):
    A = np.array(A, dtype=_dt(fp64), copy=True, order="C")
    getattr(lib(), f"orc_adjust_{_sfx(fp64)}")(A, A.shape[0], A.shape[1], A.shape[1])
    return A


def idx_max(A, fp64=True):
    A = np.ascontiguousarray(A, _dt(fp64))
    out = np.zeros(A.shape[0], np.uint32)
    getattr(lib(), f"orc_idx_max_{_sfx(fp64)}")(A, A.shape[0], A.shape[1], A.shape[1], out)
    return out


def mul_div(A, Aux, acc, fp64=True):
    dt = _dt(fp64)
    A = np.array(A, dtype=dt, copy=True, order="C")
    getattr(lib(), f"orc_mul_div_{_sfx(fp64)}")(A, np.ascontiguousarray(Aux, dt),
                                                np.ascontiguousarray(acc, dt),
                                                A.shape[0], A.shape[1], A.shape[1])
    return A


def div_sub(A, B, div, fp64=True):
    dt = _dt(fp64)
    A = np.array(A, dtype=dt, copy=True, order="C")
    getattr(lib(), f"orc_div_sub_{_sfx(fp64)}")(A, np.ascontiguousarray(B, dt),
                                                A.shape[0], A.shape[1], A.shape[1], int(div))
    return A


def dot_VWH(V, W, H, fp64=True):
    dt = _dt(fp64)
    V, W, H = (np.ascontiguousarray(x, dt) for x in (V, W, H))
    a, b = C.c_double(), C.c_double()
    n, m = V.shape
    k = W.shape[1]
    getattr(lib(), f"orc_dot_VWH_{_sfx(fp64)}")(V, n, m, m, W, H, k, k, C.byref(a), C.byref(b))
    return a.value, b.value


def kl_div(V, W, H, fp64=True):
    dt = _dt(fp64)
    V, W, H = (np.ascontiguousarray(x, dt) for x in (V, W, H))
    n, m = V.shape
    k = W.shape[1]
    return getattr(lib(), f"orc_kl_div_{_sfx(fp64)}")(V, n, m, m, W, H, k, k)


def nmf(V, W0, H0, nIters=2000, niter_test_conv=10, stop_threshold=40, fp64=True,
        max_trace=0):
    """The reference nmf() loop.  Returns a dict with W, H, cls, iterations, converged,
    ntests, and (if max_trace) per-test class vectors and differences."""
    dt = _dt(fp64)
    V = np.ascontiguousarray(V, dt)
    W = np.array(W0, dtype=dt, copy=True, order="C")
    H = np.array(H0, dtype=dt, copy=True, order="C")
    n, m = V.shape
    k = W.shape[1]
    cls = np.zeros(m, np.uint32)
    it, conv, nt = C.c_uint32(), C.c_int(), C.c_uint32()
    ctr = np.zeros((max(max_trace, 1), m), np.uint32)
    dtr = np.zeros(max(max_trace, 1), np.uint64)
    rc = getattr(lib(), f"orc_nmf_{_sfx(fp64)}")(
        V, n, m, m, W, H, k, k, nIters, niter_test_conv, stop_threshold, cls,
        C.byref(it), C.byref(conv), ctr.ctypes.data if max_trace else None,
        dtr.ctypes.data if max_trace else None, max_trace, C.byref(nt))
    if rc != 0:
        raise RuntimeError("orc_nmf failed")
    out = {"W": W, "H": H, "cls": cls, "iterations": it.value, "converged": bool(conv.value),
           "ntests": nt.value}
    if max_trace:
        t = min(nt.value, max_trace)
        out["cls_trace"] = ctr[:t].copy()
        out["diff_trace"] = dtr[:t].copy()
    return out


def ref_common():
    """The reference's src/common.c compiled unmodified (None when not built here)."""
    if not REF_COMMON.exists():
        return None
    L = C.CDLL(str(REF_COMMON))
    L.get_difference.restype = C.c_size_t
    L.get_difference.argtypes = [_U32, _U32, C.c_uint32]
    return L


REF_MATRIX_IO = pathlib.Path(__file__).resolve().parent / "_ref" / "libref_matrix_io.so"


class _RefTag(C.Structure):
    _fields_ = [("tokens", C.c_char_p), ("ptokens", C.POINTER(C.c_char_p))]


class _RefTags(C.Structure):
    _fields_ = [("name", C.c_char_p), ("headers", _RefTag), ("labels", _RefTag)]


def ref_matrix_io():
    """The reference's src/matrix_io/*.c + src/common.c compiled unmodified (None when not
    built here).  Returns (load, save): load(path, fmt, numeric_hdrs, numeric_lbls) ->
    (matrix float32 rows x cols, name, headers, labels) or None on failure; save(path, fmt,
    A, transpose, name, headers, labels) -> rc."""
    if not REF_MATRIX_IO.exists():
        return None
    L = C.CDLL(str(REF_MATRIX_IO))
    L.set_default_matrix_limits()
    u32 = C.c_uint32
    L.matrix_load.argtypes = [C.c_char_p, C.c_bool, C.c_bool, C.c_int, C.POINTER(C.POINTER(C.c_float)),
                              C.POINTER(u32), C.POINTER(u32), C.POINTER(u32), C.POINTER(_RefTags)]
    L.matrix_save.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_float), u32, u32, u32, C.c_bool,
                              C.POINTER(_RefTags), C.c_bool]

    def toks(tag, n):
        if not tag.ptokens:
            return None
        return [tag.ptokens[i].decode() if tag.ptokens[i] else "" for i in range(n)]

    def load(path, fmt, numeric_hdrs=False, numeric_lbls=False):
        A = C.POINTER(C.c_float)()
        r, c_, pitch = u32(), u32(), u32()
        mt = _RefTags()
        rc = L.matrix_load(str(path).encode(), numeric_hdrs, numeric_lbls, fmt, C.byref(A), C.byref(r),
                           C.byref(c_), C.byref(pitch), C.byref(mt))
        if rc != 0:
            return None
        M = np.ctypeslib.as_array(A, shape=(r.value, pitch.value))[:, :c_.value].copy()
        name = mt.name.decode() if mt.name else None
        return M, name, toks(mt.headers, c_.value), toks(mt.labels, r.value)

    def save(path, fmt, A, transpose=False, name=None, headers=None, labels=None):
        A = np.ascontiguousarray(A, np.float32)
        keep = []

        def mk(lst):
            if lst is None:
                return _RefTag()
            arr = (C.c_char_p * len(lst))(*[x.encode() for x in lst])
            keep.append(arr)
            blob = C.create_string_buffer(b"\0".join(x.encode() for x in lst) + b"\0")
            keep.append(blob)
            return _RefTag(C.cast(blob, C.c_char_p), C.cast(arr, C.POINTER(C.c_char_p)))
        nm = C.c_char_p(name.encode()) if name is not None else None
        mt = _RefTags(nm, mk(headers), mk(labels))
        r, c_ = A.shape
        return L.matrix_save(str(path).encode(), fmt, A.ctypes.data_as(C.POINTER(C.c_float)), r, c_, c_,
                             transpose, C.byref(mt), False)
    return load, save
