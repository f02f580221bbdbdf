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


# ---------------------------------------------------------------- the served stream's distributions

C4 = dict(n_range=(1000, 60000), m_range=(20, 3000), k_range=(2, 32), restarts_range=(1, 10))
# The lower corner of the same ranges: requests the FP64 oracle finishes in about a second.
C4_LOW = dict(n_range=(1000, 2500), m_range=(20, 120), k_range=(2, 12), restarts_range=(1, 4))


def c4_mix(nmfk):
    """8 requests drawn from the served stream's distributions (n log-U[1000, 60000], m log-U[20,
    3000], k U{2..32}, restarts U{1..10}, V lognormal(5, 1.5^2)) and 6 from the lower corner of
    the same ranges, Poisson arrivals at 40/s, interleaved by arrival time."""
    big = nmfk.request_stream(8, 40.0, 2024, **C4)
    low = nmfk.request_stream(6, 40.0, 2025, **C4_LOW)
    reqs = sorted(big + low, key=lambda q: q[5])
    # the requests the FP64 oracle runs in a few seconds per restart (n m k <= 3.5e6)
    return reqs, {i for i, q in enumerate(reqs) if q[0] * q[1] * q[2] <= 3.5e6}


@pytest.fixture(scope="module")
def c4_served(nmfk):
    reqs, low = c4_mix(nmfk)
    out, st = nmfk.serve(reqs, hbm_budget=64 << 30, detail=True)
    return reqs, low, out, st


def test_served_restarts_match_fp64_oracle(nmfk, orc, c4_served):
    # Every restart of the requests the oracle can run: stop iteration, test count, last
    # difference and final class vector exact; KL divergence within 1e-3 of FP64.
    reqs, low, out, st = c4_served
    assert st["requests"] == len(reqs)
    checked = 0
    for i in sorted(low):
        n, m, k, r, seed, _ = reqs[i]
        V = orc.gen_lognormal(n, m, seed, 5.0, 1.5)
        for ri, got in enumerate(out[i]["restarts"]):
            W0, H0 = orc.init_WH(n, m, k, nmfk.sweep_seed(seed, k, ri))
            o = orc.nmf(V, W0, H0, 2000, 10, 40, fp64=True, max_trace=200)
            where = (i, n, m, k, ri)
            assert got["iterations"] == o["iterations"] and got["ntests"] == o["ntests"], where
            assert got["converged"] == int(o["converged"]), where
            assert got["last_diff"] == o["diff_trace"][-1], where
            assert (got["cls"] == o["cls"]).all(), where
            okl = orc.kl_div(V, o["W"], o["H"], fp64=True)
            assert abs(got["kl"] - okl) <= 1e-3 * okl, (where, got["kl"], okl)
            checked += 1
    assert checked >= 20


def test_served_requests_same_alone_and_in_mix(nmfk, c4_served):
    # The requests drawn from the full ranges (too large for the oracle): each one served
    # alone gives bitwise the same restarts as in the mix, and every restart lowered the KL
    # divergence of its starting point.
    reqs, low, out, st = c4_served
    assert st["peak_jobs"] > max(q[3] for q in reqs)   # requests really overlapped
    for i, q in enumerate(reqs):
        if i in low:
            continue
        n, m, k, r, seed, _ = q
        alone, _ = nmfk.serve([q[:5] + (0.0,)], hbm_budget=64 << 30, detail=True)
        for ri, (a, b) in enumerate(zip(alone[0]["restarts"], out[i]["restarts"])):
            for key in ("iterations", "converged", "ntests", "last_diff", "kl"):
                assert a[key] == b[key], (i, ri, key)
            assert (a["cls"] == b["cls"]).all()
            with nmfk.Job(n, m, k) as j:
                j.gen_V_lognormal(seed, 5.0, 1.5)
                j.init_random(nmfk.sweep_seed(seed, k, ri))
                kl0 = j.dot_kl()[2]
            assert b["kl"] < kl0, (i, ri)
        assert out[i]["kl_best"] == min(x["kl"] for x in out[i]["restarts"])
