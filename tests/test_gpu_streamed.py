"""Streamed (out-of-core) jobs: V stays in host memory and is copied to the device in row
blocks once per iteration (nmfk_job_create_streamed).

A streamed run keeps the reference's update order, test points and stop rule, so it must
stop at the same iteration with the same class vectors as the resident run of the same
job, and W, H and the KL divergence agree within FP32 summation-order differences (the
blocks split the H half-step's sums differently from the resident plan).  Against the FP64
oracle the bar is the resident path's: class vectors and stop iteration exact, KL within
1e-3.  The pinned and the staged copy paths move the same bytes, so they give the same bits."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import GOLDEN, rel_err

pytestmark = pytest.mark.gpu

# (n, m, k, block_rows): several blocks with a ragged last one, exactly divisible, one
# block larger than n, 64-row blocks; k over the matrix-core (kt 16, 32; ranks up to 8 pad to
# 16) and, on V of fewer than 32 columns, the vector-ALU (kt 4, 8) kernel paths.
CASES = [
    (3000, 200, 3, 512),
    (3000, 200, 16, 1024),
    (2500, 300, 24, 640),
    (777, 150, 5, 64),
    (1000, 90, 10, 4096),
    (4096, 128, 8, 1024),
    (3000, 24, 3, 512),
    (2000, 28, 7, 640),
]


def resident_and_streamed(nmfk, V, k, block_rows, seed, pinned=False):
    n, m = V.shape
    a = nmfk.Job(n, m, k)
    a.set_V(V)
    a.init_random(seed)
    b = nmfk.Job.streamed(V, k, block_rows, pinned=pinned)
    b.init_random(seed)
    return a, b


@pytest.mark.parametrize("n,m,k,br", CASES)
def test_streamed_run_matches_resident(nmfk, orc, n, m, k, br):
    V = orc.gen_planted(n, m, min(k, 6), 31 + k, 0.3)
    a, b = resident_and_streamed(nmfk, V, k, br, 100 + k)
    ra, rb = a.run(600, 10, 40), b.run(600, 10, 40)
    for key in ("iterations", "converged", "ntests", "last_diff"):
        assert ra[key] == rb[key], (key, ra, rb)
    assert np.array_equal(a.get_classification(), b.get_classification())
    assert rel_err(b.get_W(), a.get_W()) < 1e-3
    assert rel_err(b.get_H(), a.get_H()) < 1e-3
    kla, klb = a.dot_kl()[2], b.dot_kl()[2]
    assert abs(kla - klb) / kla < 1e-5
    info = b.streamed_info()
    assert info["nblocks"] == -(-n // info["block_rows"])
    # one copy of V per pass: the first H half-step, then one per iteration
    assert info["bytes_copied"] == (ra["iterations"] + 1) * n * m * 4
    a.close(), b.close()


def test_streamed_config0_against_fp64(nmfk, orc):
    # The golden config-0 run (V 5000x38 lognormal(5, 1.5^2), k=3, 2000 iterations, T=10,
    # stop_threshold=40) through 1024-row blocks: stop iteration and class vector of the
    # FP64 oracle, H and KL within 1e-3.
    g = np.load(GOLDEN / "config0.npz")
    n, m, k, sv, si = (int(x) for x in g["params"])
    V = orc.gen_lognormal(n, m, sv, 5.0, 1.5)
    with nmfk.Job.streamed(V, k, 1024) as job:
        job.init_random(si)
        r = job.run(2000, 10, 40)
        assert r["converged"] and r["iterations"] == int(g["f64_iterations"])
        assert np.array_equal(job.get_classification(), g["f64_cls"])
        assert rel_err(job.get_H(), g["f64_H"]) < 1e-3
        assert rel_err(job.get_W()[:64], g["f64_W_head"]) < 1e-3
        _, _, kl = job.dot_kl()
        assert abs(kl - float(g["f64_kl"])) / float(g["f64_kl"]) < 1e-3


def test_streamed_golden_tiny_runs(nmfk, orc):
    g = np.load(GOLDEN / "tiny.npz")
    for ci in range(int(g["ncases"])):
        p = f"c{ci}_"
        n, m, k, sv, si = (int(x) for x in g[p + "shape"])
        with nmfk.Job.streamed(g[p + "V"], k, 64) as job:
            job.init_random(si)
            r = job.run(200, 5, 3)
            assert r["iterations"] == int(g[p + "iterations"]), ci
            assert r["converged"] == bool(g[p + "converged"]) and r["ntests"] == int(g[p + "ntests"])
            assert np.array_equal(job.get_classification(), g[p + "cls"]), ci


@pytest.mark.parametrize("n,m,k,br,iters", [(2000, 170, 12, 384, 120), (20000, 1000, 8, 4096, 20)])
def test_streamed_pinned_and_staged_give_same_bits(nmfk, orc, n, m, k, br, iters):
    # The second case stages 16 MB blocks, copied by several host threads.
    V = orc.gen_lognormal(n, m, 77, 0.0, 1.0)
    a = nmfk.Job.streamed(V, k, br, pinned=True)
    b = nmfk.Job.streamed(V, k, br, pinned=False)
    for j in (a, b):
        j.init_random(5)
    ra, rb = a.run(iters, 10, 1000), b.run(iters, 10, 1000)
    assert ra["iterations"] == rb["iterations"] == iters
    assert np.array_equal(a.get_W(), b.get_W()) and np.array_equal(a.get_H(), b.get_H())
    a.close(), b.close()


def test_streamed_clamps_at_the_resident_points(nmfk, orc):
    # W0 and H0 with entries below R_MIN (= FLT_MIN, 1.18e-38): the R_MIN clamps at the end
    # of each test interval (W in the finish, H in the test) turn them into R_MIN at fixed
    # iterations.  A clamp one pass early or late would leave values ~1e-39 on one side and
    # R_MIN on the other, so the comparison is entry by entry with no absolute floor.
    n, m, k, T = 1500, 100, 4, 5
    V = orc.gen_lognormal(n, m, 3, 0.0, 1.0)
    rng = np.random.default_rng(0)
    W0 = rng.uniform(0.1, 1.0, (n, k)).astype(np.float32)
    H0 = rng.uniform(0.1, 1.0, (m, k)).astype(np.float32)
    W0[rng.random((n, k)) < 0.05] = 1e-39
    H0[rng.random((m, k)) < 0.05] = 1e-39
    a = nmfk.Job(n, m, k)
    a.set_V(V)
    b = nmfk.Job.streamed(V, k, 256)
    for iters in (T - 1, T, 3 * T + 2):
        for j in (a, b):
            j.set_WH(W0, H0)
        ra, rb = a.run(iters, T, 1000), b.run(iters, T, 1000)
        assert ra["iterations"] == rb["iterations"] == iters
        np.testing.assert_allclose(b.get_W(), a.get_W(), rtol=1e-3, atol=0)
        np.testing.assert_allclose(b.get_H(), a.get_H(), rtol=1e-3, atol=0)
    a.close(), b.close()


def test_streamed_job_refuses_device_V_operations(nmfk, orc):
    V = orc.gen_lognormal(300, 50, 1, 0.0, 1.0)
    with nmfk.Job.streamed(V, 3, 128) as j:
        with pytest.raises(nmfk.NmfkError, match="streamed"):
            j.set_V(V)
        with pytest.raises(nmfk.NmfkError, match="streamed"):
            j.get_V()
        with pytest.raises(nmfk.NmfkError, match="streamed"):
            j.time_iterations(2)
        with pytest.raises(nmfk.NmfkError, match="streamed"):
            nmfk.Job.shared(j, 2)
        batch = nmfk.Batch(10, 4)
        with pytest.raises(nmfk.NmfkError, match="streamed"):
            batch.add(j, 100, 40)
        batch.close()


def test_streamed_half_steps_match_resident(nmfk, orc):
    # nmfk_job_update_H / update_W on a streamed job: one pass over V per half-step, the
    # reference's blockwise structure, with the resident job's update order and clamps.
    n, m, k = 1700, 130, 6
    V = orc.gen_lognormal(n, m, 12, 0.0, 1.0)
    a, b = resident_and_streamed(nmfk, V, k, 320, 44)
    for step in range(12):
        for j in (a, b):
            j.update_H()
            j.update_W()
            if step % 4 == 3:
                j.adjust()
    assert rel_err(b.get_W(), a.get_W()) < 1e-5 and rel_err(b.get_H(), a.get_H()) < 1e-5
    W0, H0 = orc.init_WH(n, m, k, 44)
    W, H = W0.astype(np.float64), H0.astype(np.float64)
    for step in range(12):
        H = orc.update_H(V, W, H)
        W = orc.update_W(V, W, H)
        if step % 4 == 3:
            H, W = orc.adjust(H), orc.adjust(W)
    assert rel_err(b.get_W(), W) < 1e-4 and rel_err(b.get_H(), H) < 1e-4
    a.close(), b.close()
