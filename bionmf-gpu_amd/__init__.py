"""bionmf-gpu_amd -- Python mirror of the MI355X KL-NMF C-ABI (libnmfgpu_amd.so).

The product is the C library: C host code (src/NMF_GPU.c-style driver, the
NMF_routines.h / matrix_operations.h / GPU_setup.h entry points and a job-handle API)
over hand-written gfx950 HIP kernels.  This module only binds it with ctypes so that
tests and the benchmark can drive it with numpy arrays; it never computes anything
itself and raises if the library is missing.

The package directory name contains a hyphen, so it is loaded by path:

    import importlib.util, pathlib
    spec = importlib.util.spec_from_file_location(
        "bionmf_gpu_amd", pathlib.Path("bionmf-gpu_amd/__init__.py"))
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib
import weakref

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libnmfgpu_amd.so"
DRIVER_PATH = _HERE / "NMF_GPU"
INCLUDE_DIR = _HERE.parent / "include"

_lib = None


class NmfkError(RuntimeError):
    pass


class Options(C.Structure):
    """nmfk_options (include/nmfk.h): the reference's -k, -i, -j, -t flags."""

    _fields_ = [("k", C.c_uint32), ("nIters", C.c_uint32),
                ("niter_test_conv", C.c_uint32), ("stop_threshold", C.c_uint32)]


class Result(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("converged", C.c_uint32),
                ("ntests", C.c_uint32), ("pad_", C.c_uint32),
                ("last_diff", C.c_uint64), ("seconds", C.c_double)]

    def as_dict(self):
        return {"iterations": self.iterations, "converged": bool(self.converged),
                "ntests": self.ntests, "last_diff": self.last_diff,
                "seconds": self.seconds}


class Timing(C.Structure):
    _fields_ = [("ms_per_iter", C.c_double), ("ms_partial_H", C.c_double),
                ("ms_finish_H", C.c_double), ("ms_partial_W", C.c_double),
                ("ms_finish_W", C.c_double)]


class TimingOnePass(C.Structure):
    _fields_ = [("ms_per_iter", C.c_double), ("ms_pass", C.c_double), ("ms_finish", C.c_double),
                ("workgroups", C.c_uint32), ("waves_per_workgroup", C.c_uint32)]


class JobInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32), ("k", C.c_uint32), ("kt", C.c_uint32),
                ("ldv", C.c_size_t), ("ldvt", C.c_size_t),
                ("dV", C.c_void_p), ("dVt", C.c_void_p), ("dW", C.c_void_p), ("dH", C.c_void_p),
                ("stream", C.c_void_p)]


class BatchTuning(C.Structure):
    """nmfk_batch_tuning (include/nmfk.h)."""

    _fields_ = [("grid_partial", C.c_uint32), ("grid_finish", C.c_uint32), ("order", C.c_uint32),
                ("chunks_per_poll", C.c_uint32), ("compact_pct", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SweepConfig(C.Structure):
    _fields_ = [("ks", C.POINTER(C.c_uint32)), ("nk", C.c_uint32), ("restarts", C.c_uint32),
                ("base_seed", C.c_uint64), ("nIters", C.c_uint32), ("niter_test_conv", C.c_uint32),
                ("stop_threshold", C.c_uint32), ("mode", C.c_uint32), ("plan_target", C.c_uint32),
                ("autotune", C.c_uint32), ("autotune_iters", C.c_uint32)]


class SweepOut(C.Structure):
    _fields_ = [("res", C.POINTER(Result)), ("cls", C.POINTER(C.c_uint32)),
                ("consensus", C.POINTER(C.c_int32)), ("W_H_out", C.POINTER(C.c_float)),
                ("seconds", C.c_double), ("tune_seconds", C.c_double), ("tuning", BatchTuning),
                ("plan_target", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32), ("k", C.c_uint32), ("restarts", C.c_uint32),
                ("seed", C.c_uint64), ("arrival", C.c_double)]


class RequestResult(C.Structure):
    _fields_ = [("admitted", C.c_double), ("finished", C.c_double), ("latency", C.c_double),
                ("iterations_max", C.c_uint32), ("iterations_sum", C.c_uint32),
                ("converged", C.c_uint32), ("best_restart", C.c_uint32), ("kl_best", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RestartResult(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("converged", C.c_uint32), ("ntests", C.c_uint32), ("pad_", C.c_uint32),
                ("last_diff", C.c_uint64), ("kl", C.c_double), ("cls", C.c_void_p), ("W", C.c_void_p),
                ("H", C.c_void_p)]


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("capacity", C.c_uint32), ("nIters", C.c_uint32),
                ("niter_test_conv", C.c_uint32), ("stop_threshold", C.c_uint32),
                ("chunks_per_poll", C.c_uint32), ("plan_target", C.c_uint32), ("progress_every", C.c_uint32),
                ("mu", C.c_double), ("sigma", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("requests", C.c_uint32), ("peak_jobs", C.c_uint32),
                ("requests_per_s", C.c_double), ("factorizations_per_s", C.c_double),
                ("p50", C.c_double), ("p99", C.c_double), ("mean_latency", C.c_double),
                ("peak_bytes", C.c_size_t), ("step_s", C.c_double), ("admit_s", C.c_double),
                ("wait_s", C.c_double), ("retire_s", C.c_double), ("chunks", C.c_uint32),
                ("pad_", C.c_uint32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Tags(C.Structure):
    """nmfk_tags (include/nmfk_io.h)."""

    _fields_ = [("name", C.c_char_p), ("headers", C.POINTER(C.c_char_p)), ("labels", C.POINTER(C.c_char_p)),
                ("nheaders", C.c_uint32), ("nlabels", C.c_uint32)]


_F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_U32P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def load(path: os.PathLike | str | None = None):
    """Loads libnmfgpu_amd.so (built by __graft_entry__.build() or `make`)."""
    global _lib
    if _lib is not None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise NmfkError(f"{p} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p))
    vp = C.c_void_p
    u32, u64, sz, dbl, i32 = C.c_uint32, C.c_uint64, C.c_size_t, C.c_double, C.c_int
    _sig(lib, "nmfk_last_error", C.c_char_p, [])
    _sig(lib, "nmfk_set_device", i32, [i32])
    _sig(lib, "nmfk_default_options", None, [C.POINTER(Options)])
    _sig(lib, "nmfk_job_create", vp, [u32, u32, u32])
    _sig(lib, "nmfk_job_destroy", None, [vp])
    _sig(lib, "nmfk_job_bytes", sz, [u32, u32, u32])
    _sig(lib, "nmfk_job_set_V", i32, [vp, _F32P, sz])
    _sig(lib, "nmfk_job_set_V_device", i32, [vp, vp, sz])
    _sig(lib, "nmfk_job_gen_V_lognormal", i32, [vp, u64, dbl, dbl])
    _sig(lib, "nmfk_job_gen_V_planted", i32, [vp, u64, u32, dbl])
    _sig(lib, "nmfk_job_get_V", i32, [vp, _F32P, sz])
    _sig(lib, "nmfk_job_init_random", i32, [vp, u64])
    _sig(lib, "nmfk_job_set_WH", i32, [vp, _F32P, sz, _F32P, sz])
    _sig(lib, "nmfk_job_get_W", i32, [vp, _F32P, sz])
    _sig(lib, "nmfk_job_get_H", i32, [vp, _F32P, sz])
    _sig(lib, "nmfk_job_run", i32, [vp, C.POINTER(Options), C.POINTER(Result)])
    _sig(lib, "nmfk_job_run_traced", i32, [vp, C.POINTER(Options), C.POINTER(Result), C.c_void_p, C.c_uint32,
                                           C.POINTER(C.c_uint32)])
    _sig(lib, "nmfk_job_update_H", i32, [vp])
    _sig(lib, "nmfk_job_update_W", i32, [vp])
    _sig(lib, "nmfk_job_adjust", i32, [vp])
    _sig(lib, "nmfk_job_get_classification", i32, [vp, _U32P])
    _sig(lib, "nmfk_job_dot_kl", i32, [vp, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)])
    _sig(lib, "nmfk_job_sync", i32, [vp])
    _sig(lib, "nmfk_dot_kl_noalloc", i32, [vp, sz, vp, vp, u32, u32, u32, vp, vp])
    _sig(lib, "nmfk_dot_kl_noalloc_max_bytes", sz, [])
    _sig(lib, "nmfk_job_info_get", i32, [vp, C.POINTER(JobInfo)])
    _sig(lib, "nmfk_job_time_iterations", i32, [vp, u32, C.POINTER(Timing)])
    _sig(lib, "nmfk_job_set_kernel_path", i32, [vp, i32])
    _sig(lib, "nmfk_job_set_schedule", i32, [vp, i32])
    _sig(lib, "nmfk_job_schedule", i32, [vp])
    _sig(lib, "nmfk_job_time_one_pass", i32, [vp, u32, C.POINTER(TimingOnePass)])
    _sig(lib, "nmfk_job_kernel_path", i32, [vp])
    _sig(lib, "nmfk_tune_set", i32, [i32, i32])
    _sig(lib, "nmfk_tune_get", i32, [i32])
    _sig(lib, "nmfk_job_kt", u32, [u32, u32, u32])
    _sig(lib, "nmfk_measure_stream", i32, [sz, u32, u32, C.POINTER(dbl), C.POINTER(dbl)])
    _sig(lib, "nmfk_job_create_shared", vp, [vp, u32])
    _sig(lib, "nmfk_job_replan", i32, [vp, u32])
    _sig(lib, "nmfk_job_plan_get", i32, [vp, _U32P])
    _sig(lib, "nmfk_job_create_streamed", vp, [u32, u32, u32, _F32P, sz, u32, i32])
    _sig(lib, "nmfk_job_streamed_info", i32, [vp, C.POINTER(u32), C.POINTER(u32), C.POINTER(dbl)])
    _sig(lib, "nmfk_streamed_bytes", sz, [u32, u32, u32, u32])
    _sig(lib, "nmfk_host_alloc", vp, [sz])
    _sig(lib, "nmfk_host_free", None, [vp])
    _sig(lib, "nmfk_batch_default_tuning", None, [C.POINTER(BatchTuning)])
    _sig(lib, "nmfk_batch_create", vp, [u32, u32])
    _sig(lib, "nmfk_batch_destroy", None, [vp])
    _sig(lib, "nmfk_batch_set_tuning", i32, [vp, C.POINTER(BatchTuning)])
    _sig(lib, "nmfk_batch_get_tuning", i32, [vp, C.POINTER(BatchTuning)])
    _sig(lib, "nmfk_batch_add", i32, [vp, vp, C.POINTER(Options)])
    _sig(lib, "nmfk_batch_active", u32, [vp])
    _sig(lib, "nmfk_batch_step", i32, [vp, u32])
    _sig(lib, "nmfk_batch_poll", i32, [vp, C.POINTER(vp), C.POINTER(Result), u32, C.POINTER(u32)])
    _sig(lib, "nmfk_batch_autotune", i32, [vp, u32, C.POINTER(BatchTuning), C.POINTER(dbl)])
    _sig(lib, "nmfk_sweep_default_config", None, [C.POINTER(SweepConfig)])
    _sig(lib, "nmfk_sweep_seed", u64, [u64, u32, u32])
    _sig(lib, "nmfk_sweep_run", i32, [vp, C.POINTER(SweepConfig), C.POINTER(SweepOut)])
    _sig(lib, "nmfk_serve_default_config", None, [C.POINTER(ServeConfig)])
    _sig(lib, "nmfk_tags_free", None, [C.POINTER(Tags)])
    _sig(lib, "nmfk_matrix_load", i32, [C.c_char_p, i32, i32, i32, C.POINTER(C.POINTER(C.c_float)),
                                        C.POINTER(u32), C.POINTER(u32), C.POINTER(Tags)])
    _sig(lib, "nmfk_matrix_save", i32, [C.c_char_p, i32, C.POINTER(C.c_float), u32, u32, sz, i32,
                                        C.POINTER(Tags)])
    _sig(lib, "nmfk_request_bytes", sz, [u32, u32, u32, u32])
    _sig(lib, "nmfk_serve_run", i32, [C.POINTER(ServeConfig), C.POINTER(Request), u32,
                                      C.POINTER(RequestResult), C.POINTER(ServeStats)])
    _sig(lib, "nmfk_serve_run_detail", i32, [C.POINTER(ServeConfig), C.POINTER(Request), u32,
                                             C.POINTER(RequestResult), C.POINTER(ServeStats),
                                             C.POINTER(RestartResult)])
    _lib = lib
    return lib


def last_error() -> str:
    return load().nmfk_last_error().decode(errors="replace")


def _check(rc: int, what: str):
    if rc != 0:
        raise NmfkError(f"{what} failed ({rc}): {last_error()}")


def gpu_count() -> int:
    """Number of HIP devices, from the HIP runtime directly (no torch)."""
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        try:
            hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        except OSError:
            return 0
    n = C.c_int(0)
    rc = hip.hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


TUNE_WGS_VALU, TUNE_WGS_MFMA, TUNE_MIN_STEPS, TUNE_MAX_SPLITS, TUNE_MFMA_NA, TUNE_MFMA_NB, TUNE_WGS_BF16 = range(7)


def tune_set(key: int, value: int):
    """Planner tunables (nmfk_tune_set): apply to jobs created afterwards."""
    _check(load().nmfk_tune_set(key, value), "tune_set")


def tune_get(key: int) -> int:
    return load().nmfk_tune_get(key)


def job_kt(n: int, m: int, k: int) -> int:
    """Padded rank (row pitch of W and H) a job of n x m at rank k uses (nmfk_job_kt)."""
    return int(load().nmfk_job_kt(n, m, k))


def measure_stream(nbytes: int = 1 << 30, iters: int = 10, blocks: int = 4096):
    """Measured read and copy bandwidth (GB/s) of plain streaming kernels."""
    r, c = C.c_double(), C.c_double()
    _check(load().nmfk_measure_stream(nbytes, iters, blocks, C.byref(r), C.byref(c)), "measure_stream")
    return r.value, c.value


def default_options(k: int = 2, nIters: int = 2000, niter_test_conv: int = 10,
                    stop_threshold: int = 40) -> Options:
    return Options(k, nIters, niter_test_conv, stop_threshold)


class PinnedArray:
    """Pinned host memory (nmfk_host_alloc) seen as a C-contiguous float32 array."""

    def __init__(self, shape):
        self.lib = load()
        count = int(np.prod(shape))
        self.ptr = self.lib.nmfk_host_alloc(4 * count)
        if not self.ptr:
            raise NmfkError(f"nmfk_host_alloc failed: {last_error()}")
        self.array = np.ctypeslib.as_array((C.c_float * count).from_address(self.ptr)).reshape(shape)

    def close(self):
        if self.ptr:
            self.array = None
            self.lib.nmfk_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Job:
    """One factorization V (n x m) ~ W (n x k) . H^T (H is m x k), resident on the GPU.

    Job(n, m, k) owns its V; Job.shared(src, k) reads the V of <src> without a copy;
    Job.streamed(V, k) keeps V in host memory and streams it in row blocks."""

    def __init__(self, n: int, m: int, k: int, _src: "Job | None" = None):
        self.lib = load()
        self.n, self.m, self.k = int(n), int(m), int(k)
        self._src = _src
        self._sharers = weakref.WeakSet()   # open jobs reading this job's V
        self._batch = None                  # the Batch holding this job, until poll returns it
        if _src is None:
            self.h = self.lib.nmfk_job_create(self.n, self.m, self.k)
        else:
            self.h = self.lib.nmfk_job_create_shared(_src.h, self.k)
        if not self.h:
            raise NmfkError(f"nmfk_job_create failed: {last_error()}")
        if _src is not None:
            _src._sharers.add(self)

    @classmethod
    def shared(cls, src: "Job", k: int) -> "Job":
        return cls(src.n, src.m, k, _src=src)

    @classmethod
    def streamed(cls, V: np.ndarray, k: int, block_rows: int = 0, pinned: bool = False) -> "Job":
        """nmfk_job_create_streamed: V (n x m) stays in host memory, copied to the device in
        row blocks of <block_rows> rows (a multiple of 64; 0 = library default) once per
        iteration.  pinned=True first copies V into pinned memory that the job keeps."""
        self = cls.__new__(cls)
        self.lib = load()
        V = np.ascontiguousarray(V, dtype=np.float32)
        self.n, self.m = (int(x) for x in V.shape)
        self.k, self._src, self._pin = int(k), None, None
        self._sharers, self._batch = weakref.WeakSet(), None
        if pinned:
            self._pin = PinnedArray(V.shape)
            self._pin.array[...] = V
            V = self._pin.array
        self._hostV = V
        self.h = self.lib.nmfk_job_create_streamed(self.n, self.m, self.k, V, self.m, block_rows, int(pinned))
        if not self.h:
            self._hostV = None
            if self._pin is not None:
                self._pin.close()
            raise NmfkError(f"nmfk_job_create_streamed failed: {last_error()}")
        return self

    def streamed_info(self) -> dict:
        br, nb, by = C.c_uint32(), C.c_uint32(), C.c_double()
        _check(self.lib.nmfk_job_streamed_info(self.h, C.byref(br), C.byref(nb), C.byref(by)), "streamed_info")
        return {"block_rows": br.value, "nblocks": nb.value, "bytes_copied": by.value}

    def replan(self, target_wgs: int):
        _check(self.lib.nmfk_job_replan(self.h, target_wgs), "replan")

    def plan(self) -> dict:
        out = np.zeros(8, np.uint32)
        self.lib.nmfk_job_plan_get(self.h, out)
        keys = ("nab", "S", "bchunk", "path")
        return {"H": dict(zip(keys, map(int, out[:4]))), "W": dict(zip(keys, map(int, out[4:])))}

    def close(self):
        """Frees the job.  Refused while a batch still holds it (until Batch.poll returns it
        or the batch is closed) or while jobs made by Job.shared(self, k) are open: both
        would go on reading its device memory."""
        if self.h:
            batch = getattr(self, "_batch", None)
            if batch is not None and batch.h and self.h in batch._jobs:
                raise NmfkError("job is in an active batch: poll it out or close the batch first")
            open_sharers = [j for j in getattr(self, "_sharers", ()) if j.h]
            if open_sharers:
                raise NmfkError(f"{len(open_sharers)} job(s) made by Job.shared still read this job's V: "
                                "close them first")
            self.lib.nmfk_job_destroy(self.h)
            self.h = None
        pin = getattr(self, "_pin", None)
        if pin is not None:         # after the job: its streams no longer read the buffer
            self._hostV = None
            pin.close()
            self._pin = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # inputs -------------------------------------------------------------
    def set_V(self, V: np.ndarray):
        V = np.ascontiguousarray(V, dtype=np.float32)
        assert V.shape == (self.n, self.m), V.shape
        _check(self.lib.nmfk_job_set_V(self.h, V, self.m), "set_V")

    def gen_V_lognormal(self, seed: int, mu: float, sigma: float):
        _check(self.lib.nmfk_job_gen_V_lognormal(self.h, seed, mu, sigma), "gen_V_lognormal")

    def gen_V_planted(self, seed: int, kstar: int, sigma: float):
        _check(self.lib.nmfk_job_gen_V_planted(self.h, seed, kstar, sigma), "gen_V_planted")

    def get_V(self) -> np.ndarray:
        out = np.empty((self.n, self.m), np.float32)
        _check(self.lib.nmfk_job_get_V(self.h, out, self.m), "get_V")
        return out

    def init_random(self, seed: int):
        _check(self.lib.nmfk_job_init_random(self.h, seed), "init_random")

    def set_WH(self, W: np.ndarray, H: np.ndarray):
        W = np.ascontiguousarray(W, dtype=np.float32)
        H = np.ascontiguousarray(H, dtype=np.float32)
        assert W.shape == (self.n, self.k) and H.shape == (self.m, self.k)
        _check(self.lib.nmfk_job_set_WH(self.h, W, self.k, H, self.k), "set_WH")

    # outputs ------------------------------------------------------------
    def get_W(self) -> np.ndarray:
        out = np.empty((self.n, self.k), np.float32)
        _check(self.lib.nmfk_job_get_W(self.h, out, self.k), "get_W")
        return out

    def get_H(self) -> np.ndarray:
        out = np.empty((self.m, self.k), np.float32)
        _check(self.lib.nmfk_job_get_H(self.h, out, self.k), "get_H")
        return out

    def get_classification(self) -> np.ndarray:
        out = np.empty(self.m, np.uint32)
        _check(self.lib.nmfk_job_get_classification(self.h, out), "get_classification")
        return out

    def dot_kl(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(self.lib.nmfk_job_dot_kl(self.h, C.byref(a), C.byref(b), C.byref(c)), "dot_kl")
        return a.value, b.value, c.value

    # the loop -----------------------------------------------------------
    def run(self, nIters=2000, niter_test_conv=10, stop_threshold=40, trace=False) -> dict:
        """nmfk_job_run; with trace=True also 'diff_trace', the connectivity difference the
        device counted at every convergence test (nmfk_job_run_traced)."""
        opt = Options(self.k, nIters, niter_test_conv, stop_threshold)
        res = Result()
        if not trace:
            _check(self.lib.nmfk_job_run(self.h, C.byref(opt), C.byref(res)), "run")
            return res.as_dict()
        cap = nIters // max(niter_test_conv, 1) + 1
        buf = np.zeros(cap, np.uint64)
        nt = C.c_uint32()
        _check(self.lib.nmfk_job_run_traced(self.h, C.byref(opt), C.byref(res), buf.ctypes.data, cap,
                                            C.byref(nt)), "run")
        out = res.as_dict()
        out["diff_trace"] = buf[:min(nt.value, cap)].copy()
        return out

    def update_H(self):
        _check(self.lib.nmfk_job_update_H(self.h), "update_H")

    def update_W(self):
        _check(self.lib.nmfk_job_update_W(self.h), "update_W")

    def adjust(self):
        _check(self.lib.nmfk_job_adjust(self.h), "adjust")

    def sync(self):
        _check(self.lib.nmfk_job_sync(self.h), "sync")

    PATH_AUTO, PATH_VALU, PATH_MFMA32 = 0, 1, 2

    def set_kernel_path(self, path: int):
        _check(self.lib.nmfk_job_set_kernel_path(self.h, path), "set_kernel_path")

    def kernel_path(self) -> int:
        return self.lib.nmfk_job_kernel_path(self.h)

    SCHED_AUTO, SCHED_TWO_PASS, SCHED_ONE_PASS = 0, 1, 2

    def set_schedule(self, sched: int):
        _check(self.lib.nmfk_job_set_schedule(self.h, sched), "set_schedule")

    def schedule(self) -> int:
        return self.lib.nmfk_job_schedule(self.h)

    def time_one_pass(self, iters: int) -> dict:
        t = TimingOnePass()
        _check(self.lib.nmfk_job_time_one_pass(self.h, iters, C.byref(t)), "time_one_pass")
        return {f: getattr(t, f) for f, _ in TimingOnePass._fields_}

    def info(self) -> JobInfo:
        inf = JobInfo()
        _check(self.lib.nmfk_job_info_get(self.h, C.byref(inf)), "info")
        return inf

    def time_iterations(self, iters: int) -> dict:
        t = Timing()
        _check(self.lib.nmfk_job_time_iterations(self.h, iters, C.byref(t)), "time_iterations")
        return {f: getattr(t, f) for f, _ in Timing._fields_}


class Batch:
    """nmfk_batch: continuous batching of many jobs (include/nmfk.h)."""

    def __init__(self, niter_test_conv: int = 10, capacity: int = 256):
        self.lib = load()
        self.T = niter_test_conv
        self.cap = capacity
        self.h = self.lib.nmfk_batch_create(niter_test_conv, capacity)
        if not self.h:
            raise NmfkError(f"nmfk_batch_create failed: {last_error()}")
        self._jobs = {}

    def close(self):
        if self.h:
            self.lib.nmfk_batch_destroy(self.h)
            self.h = None
            for job in self._jobs.values():
                job._batch = None
            self._jobs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, job: Job, nIters=2000, stop_threshold=40) -> int:
        opt = Options(job.k, nIters, self.T, stop_threshold)
        slot = self.lib.nmfk_batch_add(self.h, job.h, C.byref(opt))
        if slot < 0:
            raise NmfkError(f"batch add failed: {last_error()}")
        self._jobs[job.h] = job
        job._batch = self
        return slot

    @property
    def active(self) -> int:
        return self.lib.nmfk_batch_active(self.h)

    def step(self, chunks: int = 1):
        _check(self.lib.nmfk_batch_step(self.h, chunks), "batch step")

    def poll(self):
        """[(job, result dict)] of the jobs that finished."""
        done = (C.c_void_p * self.cap)()
        res = (Result * self.cap)()
        nd = C.c_uint32(0)
        _check(self.lib.nmfk_batch_poll(self.h, done, res, self.cap, C.byref(nd)), "batch poll")
        out = [(self._jobs.pop(done[i]), res[i].as_dict()) for i in range(nd.value)]
        for job, _ in out:
            job._batch = None
        return out

    def tuning(self) -> dict:
        t = BatchTuning()
        self.lib.nmfk_batch_get_tuning(self.h, C.byref(t))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = BatchTuning()
        self.lib.nmfk_batch_get_tuning(self.h, C.byref(t))
        for k_, v in kw.items():
            setattr(t, k_, v)
        _check(self.lib.nmfk_batch_set_tuning(self.h, C.byref(t)), "batch tuning")

    def autotune(self, iters: int = 20):
        t = BatchTuning()
        ms = C.c_double(0.0)
        _check(self.lib.nmfk_batch_autotune(self.h, iters, C.byref(t), C.byref(ms)), "autotune")
        return t.as_dict(), ms.value

    def run(self, chunks_per_poll: int = 1):
        """Steps until every job has finished; returns [(job, result)] in finishing order."""
        out = []
        while self.active:
            self.step(chunks_per_poll)
            out += self.poll()
        return out


SWEEP_BATCHED, SWEEP_ONE_AT_A_TIME = 0, 1


def sweep_seed(base: int, k: int, restart: int) -> int:
    return load().nmfk_sweep_seed(base, k, restart)


def sweep(src: Job, ks, restarts: int, base_seed: int, nIters=2000, niter_test_conv=10,
          stop_threshold=40, mode=SWEEP_BATCHED, plan_target=0, autotune=False, autotune_iters=20,
          want_factors=False) -> dict:
    """nmfk_sweep_run: every (k, restart) job on src's V; returns results, final class vectors
    (nk x restarts x m), consensus matrices (nk x m x m int32) and timings."""
    lib = load()
    ks = np.ascontiguousarray(ks, dtype=np.uint32)
    nk, m = len(ks), src.m
    cfg = SweepConfig()
    lib.nmfk_sweep_default_config(C.byref(cfg))
    cfg.ks = ks.ctypes.data_as(C.POINTER(C.c_uint32))
    cfg.nk, cfg.restarts, cfg.base_seed = nk, restarts, base_seed
    cfg.nIters, cfg.niter_test_conv, cfg.stop_threshold = nIters, niter_test_conv, stop_threshold
    cfg.mode, cfg.plan_target = mode, plan_target
    cfg.autotune, cfg.autotune_iters = int(bool(autotune)), autotune_iters
    nj = nk * restarts
    res = (Result * nj)()
    cls = np.zeros((nk, restarts, m), np.uint32)
    cons = np.zeros((nk, m, m), np.int32)
    out = SweepOut()
    out.res = C.cast(res, C.POINTER(Result))
    out.cls = cls.ctypes.data_as(C.POINTER(C.c_uint32))
    out.consensus = cons.ctypes.data_as(C.POINTER(C.c_int32))
    fac = None
    if want_factors:
        fac = np.zeros(int(sum((src.n + m) * int(k) for k in ks) * restarts), np.float32)
        out.W_H_out = fac.ctypes.data_as(C.POINTER(C.c_float))
    _check(lib.nmfk_sweep_run(src.h, C.byref(cfg), C.byref(out)), "sweep")
    r = {"results": [res[i].as_dict() for i in range(nj)], "cls": cls, "consensus": cons,
         "seconds": out.seconds, "tune_seconds": out.tune_seconds, "tuning": out.tuning.as_dict(),
         "plan_target": out.plan_target}
    if want_factors:
        facs, o = [], 0
        for k in ks:
            for _ in range(restarts):
                W = fac[o:o + src.n * int(k)].reshape(src.n, int(k)); o += src.n * int(k)
                H = fac[o:o + m * int(k)].reshape(m, int(k)); o += m * int(k)
                facs.append((W, H))
        r["factors"] = facs
    return r


def serve(requests, hbm_budget=200 << 30, capacity=1024, nIters=2000, niter_test_conv=10,
          stop_threshold=40, mu=5.0, sigma=1.5, chunks_per_poll=1, plan_target=None, progress_every=0,
          detail=False, factors_of=()):
    """nmfk_serve_run over [(n, m, k, restarts, seed, arrival_s), ...] (arrivals sorted).
    Returns (per-request result dicts, stats dict).  With detail=True (nmfk_serve_run_detail)
    each request dict also has 'restarts': per-restart dicts with iterations, converged,
    ntests, last_diff, kl and the final class vector 'cls'; requests whose index is in
    factors_of also get the final 'W' and 'H' of every restart."""
    lib = load()
    cfg = ServeConfig()
    lib.nmfk_serve_default_config(C.byref(cfg))
    cfg.hbm_budget, cfg.capacity = hbm_budget, capacity
    cfg.nIters, cfg.niter_test_conv, cfg.stop_threshold = nIters, niter_test_conv, stop_threshold
    cfg.mu, cfg.sigma, cfg.chunks_per_poll = mu, sigma, chunks_per_poll
    if plan_target is not None:
        cfg.plan_target = plan_target
    cfg.progress_every = progress_every
    nr = len(requests)
    req = (Request * max(nr, 1))()
    for i, (n, m, k, r, seed, t) in enumerate(requests):
        req[i] = Request(n, m, k, r, seed, t)
    out = (RequestResult * max(nr, 1))()
    st = ServeStats()
    if not detail:
        _check(lib.nmfk_serve_run(C.byref(cfg), req, nr, out, C.byref(st)), "serve")
        return [out[i].as_dict() for i in range(nr)], st.as_dict()
    total = sum(q[3] for q in requests)
    recs = (RestartResult * max(total, 1))()
    bufs = []
    o = 0
    for i, (n, m, k, r, _, _) in enumerate(requests):
        for _ in range(r):
            cls = np.zeros(m, np.uint32)
            W = np.zeros((n, k), np.float32) if i in factors_of else None
            H = np.zeros((m, k), np.float32) if i in factors_of else None
            recs[o].cls = cls.ctypes.data
            recs[o].W = W.ctypes.data if W is not None else None
            recs[o].H = H.ctypes.data if H is not None else None
            bufs.append((cls, W, H))
            o += 1
    _check(lib.nmfk_serve_run_detail(C.byref(cfg), req, nr, out, C.byref(st), recs), "serve")
    res = [out[i].as_dict() for i in range(nr)]
    o = 0
    for i, q in enumerate(requests):
        res[i]["restarts"] = []
        for _ in range(q[3]):
            cls, W, H = bufs[o]
            d = {f: getattr(recs[o], f) for f in ("iterations", "converged", "ntests", "last_diff", "kl")}
            d["cls"] = cls
            if W is not None:
                d["W"], d["H"] = W, H
            res[i]["restarts"].append(d)
            o += 1
    return res, st.as_dict()


def request_stream(count, rate, seed, n_range=(1000, 60000), m_range=(20, 3000), k_range=(2, 32),
                   restarts_range=(1, 10)):
    """Synthetic request stream: Poisson arrivals at <rate> per second; n and m log-uniform
    in their ranges, k uniform in {k_lo .. min(k_hi, m)}, restarts uniform; per-request
    seeds.  Returns [(n, m, k, restarts, seed, arrival_s), ...]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count)) if rate > 0 else np.zeros(count)
    out = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1])))))
        m = int(round(np.exp(rng.uniform(np.log(m_range[0]), np.log(m_range[1])))))
        k = int(rng.integers(k_range[0], min(k_range[1], m) + 1))
        r = int(rng.integers(restarts_range[0], restarts_range[1] + 1))
        out.append((n, m, k, r, int(rng.integers(0, 2**63)), float(t[i])))
    return out


FMT_TEXT, FMT_BINARY, FMT_NATIVE = 0, 1, 2


def load_matrix(path, fmt=FMT_TEXT, numeric_hdrs=False, numeric_lbls=False):
    """nmfk_matrix_load: returns (matrix float32 rows x cols, name, headers, labels)."""
    lib = load()
    A = C.POINTER(C.c_float)()
    r, c_ = C.c_uint32(), C.c_uint32()
    t = Tags()
    _check(lib.nmfk_matrix_load(str(path).encode(), fmt, int(numeric_hdrs), int(numeric_lbls), C.byref(A),
                                C.byref(r), C.byref(c_), C.byref(t)), "load_matrix")
    M = np.ctypeslib.as_array(A, shape=(r.value, c_.value)).copy()
    C.CDLL(None).free(C.cast(A, C.c_void_p))
    name = t.name.decode() if t.name else None
    hd = [t.headers[i].decode() for i in range(t.nheaders)] if t.headers else None
    lb = [t.labels[i].decode() for i in range(t.nlabels)] if t.labels else None
    lib.nmfk_tags_free(C.byref(t))
    return M, name, hd, lb


def save_matrix(path, A, fmt=FMT_TEXT, transpose=False, name=None, headers=None, labels=None):
    """nmfk_matrix_save; tags describe the matrix as written (after the optional transpose)."""
    lib = load()
    A = np.ascontiguousarray(A, np.float32)
    keep = []

    def arr(lst):
        if lst is None:
            return None
        a = (C.c_char_p * len(lst))(*[x.encode() for x in lst])
        keep.append(a)
        return C.cast(a, C.POINTER(C.c_char_p))
    t = Tags(name.encode() if name is not None else None, arr(headers), arr(labels),
             len(headers) if headers else 0, len(labels) if labels else 0)
    r, c_ = A.shape
    _check(lib.nmfk_matrix_save(str(path).encode(), fmt, A.ctypes.data_as(C.POINTER(C.c_float)), r, c_, c_,
                                int(bool(transpose)), C.byref(t)), "save_matrix")
