"""The serving loop's device memory arena (csrc/nmfk_arena.c) on the host: the file is
compiled with gcc against a stub of the two HIP allocation calls it makes (host memory in
place of device memory), and a randomized workload checks its invariants -- 256-byte
alignment, no two live blocks overlapping, exact in-use accounting, coalescing of freed
neighbours, the byte limit, and segments returned only by trim."""
from __future__ import annotations

import ctypes as C
import subprocess

import numpy as np
import pytest

from conftest import ROOT

STUB = r"""
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <hip/hip_runtime_api.h>
int stub_mallocs, stub_frees;
hipError_t hipMalloc(void **p, size_t n) {
    *p = aligned_alloc(256, (n + 255) / 256 * 256);
    if (!*p) return hipErrorOutOfMemory;
    stub_mallocs++;
    return hipSuccess;
}
hipError_t hipFree(void *p) { if (p) { stub_frees++; free(p); } return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
char stub_err[512];
void nmfk_set_error(const char *fmt, ...) {
    va_list ap; va_start(ap, fmt); vsnprintf(stub_err, sizeof stub_err, fmt, ap); va_end(ap);
}
"""


@pytest.fixture(scope="module")
def arena(tmp_path_factory):
    d = tmp_path_factory.mktemp("arena")
    (d / "stub.c").write_text(STUB)
    so = d / "libarena.so"
    subprocess.run(["gcc", "-std=c11", "-O1", "-fPIC", "-shared", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", f"-I{ROOT / 'include'}", f"-I{ROOT / 'bionmf-gpu_amd' / 'csrc'}",
                    str(ROOT / "bionmf-gpu_amd" / "csrc" / "nmfk_arena.c"), str(d / "stub.c"), "-o", str(so)],
                   check=True)
    L = C.CDLL(str(so))
    vp, sz = C.c_void_p, C.c_size_t
    L.nmfk_arena_create.restype, L.nmfk_arena_create.argtypes = vp, [sz, sz]
    L.nmfk_arena_destroy.argtypes = [vp]
    L.nmfk_arena_alloc.restype, L.nmfk_arena_alloc.argtypes = vp, [vp, sz]
    L.nmfk_arena_free.restype, L.nmfk_arena_free.argtypes = C.c_int, [vp, vp]
    for f in ("nmfk_arena_round", "nmfk_arena_reserved", "nmfk_arena_in_use"):
        getattr(L, f).restype = sz
    L.nmfk_arena_round.argtypes = [sz]
    L.nmfk_arena_reserved.argtypes = L.nmfk_arena_in_use.argtypes = [vp]
    L.nmfk_arena_trim.argtypes = [vp]
    return L


def counters(L):
    return C.c_int.in_dll(L, "stub_mallocs").value, C.c_int.in_dll(L, "stub_frees").value


def test_randomized_alloc_free_invariants(arena):
    L = arena
    MiB = 1 << 20
    a = L.nmfk_arena_create(64 * MiB, 16 * MiB)
    rng = np.random.default_rng(7)
    live = {}   # address -> rounded size
    m0, f0 = counters(L)
    for step in range(4000):
        if live and rng.random() < 0.45:
            p = list(live)[rng.integers(len(live))]
            assert L.nmfk_arena_free(a, p) == 0
            del live[p]
        else:
            want = int(rng.choice([1, 4, 100, 256, 257, 4096, 65536, 1 << 20, 3 << 20]) * rng.integers(1, 4))
            p = L.nmfk_arena_alloc(a, want)
            rs = L.nmfk_arena_round(want)
            assert rs % 256 == 0 and rs >= want
            if p is None:
                # only when the live blocks leave no room within the limit
                assert sum(live.values()) + rs > 48 * MiB or L.nmfk_arena_reserved(a) == 64 * MiB
                continue
            assert p % 256 == 0
            for q, qs in live.items():
                assert p + rs <= q or q + qs <= p, "overlapping blocks"
            live[p] = rs
        assert L.nmfk_arena_in_use(a) == sum(live.values())
        assert L.nmfk_arena_reserved(a) <= 64 * MiB
    m1, f1 = counters(L)
    assert f1 == f0 and 1 <= m1 - m0 <= 4    # segments only, never released while in use
    for p in list(live):
        assert L.nmfk_arena_free(a, p) == 0
    assert L.nmfk_arena_in_use(a) == 0
    # everything coalesced: each segment takes a block of its own full size again
    p = L.nmfk_arena_alloc(a, 16 * MiB)
    assert p is not None
    assert L.nmfk_arena_free(a, p) == 0
    assert L.nmfk_arena_free(a, p) != 0       # double free is refused
    assert L.nmfk_arena_free(a, 12345) != 0   # so is a foreign pointer
    L.nmfk_arena_trim(a)
    assert L.nmfk_arena_reserved(a) == 0 and counters(L)[1] == f0 + (m1 - m0)
    L.nmfk_arena_destroy(a)


def test_limit_and_first_fit_reuse(arena):
    L = arena
    a = L.nmfk_arena_create(1 << 20, 1 << 20)
    x = L.nmfk_arena_alloc(a, 300 << 10)
    y = L.nmfk_arena_alloc(a, 300 << 10)
    z = L.nmfk_arena_alloc(a, 300 << 10)
    assert None not in (x, y, z) and y == x + (300 << 10) and z == y + (300 << 10)
    assert L.nmfk_arena_alloc(a, 300 << 10) is None          # beyond the 1 MiB limit
    assert L.nmfk_arena_free(a, y) == 0
    assert L.nmfk_arena_alloc(a, 200 << 10) == y              # first fit reuses the hole
    assert L.nmfk_arena_alloc(a, 200 << 10) is None           # 100 KiB + tail left, split
    assert L.nmfk_arena_free(a, x) == 0
    w = L.nmfk_arena_alloc(a, 300 << 10)
    assert w == x
    L.nmfk_arena_destroy(a)
