"""Serving loop (nmfk_serve_run) through the C-ABI: requests of uneven shape arriving over
time are admitted by HBM budget and advanced together; each request's result must be the
one it gets when served alone (bitwise: stop iterations, converged restarts, and the FP64
KL divergence of its best restart)."""
from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu

KEYS = ("iterations_max", "iterations_sum", "converged", "best_restart", "kl_best")


def small_stream(nmfk, count=8, rate=40.0, seed=5):
    return nmfk.request_stream(count, rate, seed, n_range=(300, 3000), m_range=(20, 300),
                               k_range=(2, 32), restarts_range=(1, 3))


def test_serve_same_alone_and_in_mix(nmfk):
    reqs = small_stream(nmfk)
    mix, st = nmfk.serve(reqs, nIters=300)
    assert st["requests"] == len(reqs) and st["p99"] >= st["p50"] > 0
    assert st["peak_jobs"] > max(r[3] for r in reqs)   # requests really overlapped
    for req, got in zip(reqs, mix):
        alone, _ = nmfk.serve([req[:5] + (0.0,)], nIters=300)
        for key in KEYS:
            assert alone[0][key] == got[key], (req, key)
        assert got["latency"] > 0 and got["finished"] >= got["admitted"] >= req[5] - 1e-9


def test_serve_admission_respects_budget(nmfk):
    reqs = small_stream(nmfk, count=6, rate=0.0, seed=9)   # all arrive at once
    sizes = [nmfk.load().nmfk_request_bytes(n, m, k, r) for n, m, k, r, _, _ in reqs]
    budget = 2 * max(sizes)
    out, st = nmfk.serve(reqs, hbm_budget=budget, nIters=200)
    assert st["requests"] == len(reqs)
    assert st["peak_bytes"] <= budget
    ref, _ = nmfk.serve(reqs, nIters=200)                  # unconstrained: same results
    for a, b in zip(out, ref):
        for key in KEYS:
            assert a[key] == b[key]
    with pytest.raises(nmfk.NmfkError):
        nmfk.serve(reqs, hbm_budget=min(sizes) - 1, nIters=200)
