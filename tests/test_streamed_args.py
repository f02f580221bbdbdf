"""Argument checks of nmfk_job_create_streamed, which fail before any device call (so they
run without a GPU)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lib(pkg):
    return pkg.load()


@pytest.mark.parametrize("n,m,k,ldv,br,msg", [
    (100, 50, 3, 50, 100, "multiple of 64"),
    (100, 50, 3, 40, 64, "ldv >= m"),
    (100, 50, 0, 50, 64, "invalid shape"),
    (100, 50, 51, 50, 64, "invalid shape"),
    (100, 50, 65, 50, 64, "invalid shape"),
])
def test_create_streamed_rejects(pkg, lib, n, m, k, ldv, br, msg):
    V = np.ones((n, max(m, ldv)), np.float32)
    h = lib.nmfk_job_create_streamed(n, m, k, V, ldv, br, 0)
    assert not h
    assert msg in pkg.last_error()


def test_create_streamed_needs_V(pkg, lib):
    # a null V, through a binding of the entry point that takes a plain pointer
    raw = C.CDLL(str(pkg.LIB_PATH)).nmfk_job_create_streamed
    raw.restype = C.c_void_p
    raw.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_size_t, C.c_uint32, C.c_int]
    assert not raw(10, 10, 2, None, 10, 64, 0)
    assert "need V" in pkg.last_error()
