"""One pass over V per iteration (k_fused_mfma + k_fused_finish, NMFK_SCHED_ONE_PASS) against
the FP64 oracle and against the two-pass schedule, through the C-ABI.

Bars as for the two-pass path: stop iterations, test counts, per-test connectivity
differences and class vectors bit-exact against the FP64 oracle; W and H within 2e-5
relative after one iteration, 1e-3 over whole runs."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

ONE, TWO = 2, 1

# n, m, k: ragged row blocks (n % 16), a single partial block, both workgroup widths
# (4 and 8 waves of 128 columns: m up to 512 and 1024), tiles and waves cut by m, and k 9..16
# (padded rank 16).
SHAPES = [(1000, 37, 12), (333, 1024, 16), (16, 100, 9), (15, 64, 10), (2049, 513, 13),
          (4097, 65, 16), (100, 1000, 16), (3000, 257, 11), (777, 700, 14)]


def _job(nmfk, orc, n, m, k, seed_v, seed_init, sched):
    V = orc.gen_lognormal(n, m, seed_v, 0.0, 1.0)
    j = nmfk.Job(n, m, k)
    j.set_V(V)
    j.init_random(seed_init)
    j.set_schedule(sched)
    return j, V


@pytest.mark.parametrize("n,m,k", SHAPES)
def test_one_pass_iterations_match_oracle(nmfk, orc, n, m, k):
    # 1, 2 and 7 iterations (no test: T larger than the run) against the FP64 update order.
    V = orc.gen_lognormal(n, m, n + m, 0.0, 1.0)
    W0, H0 = orc.init_WH(n, m, k, k + 7)
    with nmfk.Job(n, m, k) as j:
        j.set_V(V)
        j.set_schedule(ONE)
        assert j.schedule() == ONE
        for iters, tol in ((1, 2e-5), (2, 4e-5), (7, 2e-4)):
            j.init_random(k + 7)
            r = j.run(iters, 1000, 40)
            assert r["iterations"] == iters and r["ntests"] == 0
            o = orc.nmf(V, W0, H0, iters, 1000, 40, fp64=True)
            assert rel_err(j.get_H(), o["H"]) < tol, (iters, rel_err(j.get_H(), o["H"]))
            assert rel_err(j.get_W(), o["W"]) < tol, (iters, rel_err(j.get_W(), o["W"]))


@pytest.mark.parametrize("n,m,k,T,stop", [(1200, 300, 12, 5, 6), (2500, 128, 16, 3, 8),
                                          (700, 1000, 10, 10, 4)])
def test_one_pass_runs_match_oracle_test_by_test(nmfk, orc, n, m, k, T, stop):
    # Whole runs under the stop rule: every test's difference count, the stop iteration and
    # the class vector exactly as the FP64 oracle; factors within 1e-3.
    V = orc.gen_lognormal(n, m, 31, 0.0, 1.0)
    W0, H0 = orc.init_WH(n, m, k, 32)
    o = orc.nmf(V, W0, H0, 1500, T, stop, fp64=True, max_trace=1000)
    with nmfk.Job(n, m, k) as j:
        j.set_V(V)
        j.init_random(32)
        j.set_schedule(ONE)
        r = j.run(1500, T, stop, trace=True)
        assert r["iterations"] == o["iterations"] and r["ntests"] == o["ntests"]
        assert r["converged"] == o["converged"]
        assert np.array_equal(r["diff_trace"], o["diff_trace"])
        assert np.array_equal(j.get_classification(), o["cls"])
        assert rel_err(j.get_W(), o["W"]) < 1e-3 and rel_err(j.get_H(), o["H"]) < 1e-3


def test_one_pass_and_two_pass_agree(nmfk, orc):
    n, m, k = 5000, 800, 16
    res = []
    for sched in (TWO, ONE):
        j, V = _job(nmfk, orc, n, m, k, 41, 42, sched)
        with j:
            assert j.schedule() == sched
            r = j.run(400, 10, 20)
            res.append((r, j.get_classification(), j.get_W(), j.get_H(), j.dot_kl()[2]))
    (r2, c2, W2, H2, kl2), (r1, c1, W1, H1, kl1) = res
    assert r1["iterations"] == r2["iterations"] and r1["ntests"] == r2["ntests"]
    assert np.array_equal(c1, c2)
    assert rel_err(W1, W2) < 1e-3 and rel_err(H1, H2) < 1e-3
    assert abs(kl1 - kl2) <= 1e-5 * kl2


def test_one_pass_clamps_at_the_resident_points(nmfk, orc):
    # W0 and H0 with entries below R_MIN: the clamps at the end of each test interval (new W
    # rows in the pass, H in the test, the pass's H sums against clamped H) must fall on the
    # same iterations as the reference's adjust calls, so values ~1e-39 and R_MIN are told
    # apart entry by entry (no absolute floor).
    n, m, k, T = 1500, 100, 12, 5
    V = orc.gen_lognormal(n, m, 3, 0.0, 1.0)
    rng = np.random.default_rng(0)
    W0 = rng.uniform(0.1, 1.0, (n, k)).astype(np.float32)
    H0 = rng.uniform(0.1, 1.0, (m, k)).astype(np.float32)
    W0[rng.random((n, k)) < 0.05] = 1e-39
    H0[rng.random((m, k)) < 0.05] = 1e-39
    with nmfk.Job(n, m, k) as a, nmfk.Job(n, m, k) as b:
        for j, s in ((a, TWO), (b, ONE)):
            j.set_V(V)
            j.set_schedule(s)
        for iters in (T - 1, T, 3 * T + 2):
            for j in (a, b):
                j.set_WH(W0, H0)
            ra, rb = a.run(iters, T, 1000), b.run(iters, T, 1000)
            assert ra["iterations"] == rb["iterations"] == iters
            np.testing.assert_allclose(b.get_W(), a.get_W(), rtol=1e-3, atol=0)
            np.testing.assert_allclose(b.get_H(), a.get_H(), rtol=1e-3, atol=0)


def test_one_pass_is_deterministic(nmfk, orc):
    outs = []
    for _ in range(2):
        j, V = _job(nmfk, orc, 6000, 1000, 16, 5, 6, ONE)
        with j:
            j.run(30, 10, 1000)
            outs.append((j.get_W(), j.get_H()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_schedule_selection(nmfk, orc):
    # Automatic is two passes everywhere; one pass is available (and refused elsewhere) for
    # the FP32 matrix-core path at padded rank 16 with m <= 1024 (k <= 8 pads to 16 unless V
    # has fewer than 32 rows or columns).
    for (n, m, k), ok in [((500, 300, 16), True), ((500, 1024, 9), True), ((500, 1025, 16), False),
                          ((500, 300, 8), True), ((500, 24, 8), False), ((500, 300, 17), False),
                          ((500, 300, 40), False)]:
        with nmfk.Job(n, m, k) as j:
            assert j.schedule() == TWO, (n, m, k)
            if ok:
                j.set_schedule(ONE)
                assert j.schedule() == ONE
            else:
                with pytest.raises(nmfk.NmfkError, match="one pass"):
                    j.set_schedule(ONE)
            j.set_schedule(TWO)
            assert j.schedule() == TWO
    with nmfk.Job(500, 300, 16) as j:
        j.set_schedule(ONE)
        j.set_kernel_path(1)                 # vector ALUs: two passes
        assert j.schedule() == TWO
        j.set_kernel_path(2)
        assert j.schedule() == ONE
    with pytest.raises(nmfk.NmfkError):
        with nmfk.Job(500, 300, 16) as j:
            j.set_schedule(3)


def test_one_pass_timing(nmfk, orc):
    j, V = _job(nmfk, orc, 20000, 1000, 16, 1, 2, ONE)
    with j:
        t = j.time_one_pass(5)
        assert t["ms_pass"] > 0 and t["ms_finish"] > 0 and t["ms_per_iter"] > 0
        assert t["waves_per_workgroup"] == 8 and t["workgroups"] == 250
