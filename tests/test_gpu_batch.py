"""Batched execution (continuous batching) and the rank sweep, through the C-ABI.

A job advanced inside a batch must give the bits it gives alone (nmfk_job_run on a twin
job with the same V, seed and kernel plan): W, H, iterations, tests and class vectors.
The sweep's consensus matrices are checked exactly against the class vectors it returns,
and sweep jobs against the FP64 oracle (class vectors and stop iteration exact, KL within
1e-3)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

# (n, m, k, seed_v, seed_init, nIters, stop_threshold): k covers every batched width
# (kt 4 and 8 on V of fewer than 32 columns, where small ranks stay on the vector ALUs, with
# and without lane groups; kt 16 and 32), and the 4000 x 3000 (kt 16, both halves) and
# 4200 x 2800 (kt 32, W half) jobs plan 2 x 2 matrix-core waves (1000-4095 rows of U, at
# least 240 workgroups); the last two cannot converge in time and run the leftover
# nIters % T iterations after their test budget.
MIX = [
    (700, 150, 3, 1, 11, 400, 4),
    (1200, 90, 6, 2, 12, 400, 4),
    (3000, 24, 6, 11, 21, 400, 4),
    (333, 200, 10, 3, 13, 400, 4),
    (900, 64, 17, 4, 14, 400, 4),
    (2000, 300, 32, 5, 15, 400, 4),
    (4000, 3000, 16, 9, 19, 400, 4),
    (4200, 2800, 24, 10, 20, 400, 4),
    (64, 20, 2, 6, 16, 400, 4),
    (500, 77, 5, 7, 17, 123, 1000),
    (256, 300, 16, 8, 18, 57, 1000),
]
T = 5


def twins(nmfk, orc, spec):
    n, m, k, sv, si, _, _ = spec
    V = orc.gen_lognormal(n, m, sv, 0.0, 1.0)
    a, b = nmfk.Job(n, m, k), nmfk.Job(n, m, k)
    for j in (a, b):
        j.set_V(V)
        j.init_random(si)
    return a, b


def assert_same(a, b, ra, rb, what):
    for key in ("iterations", "converged", "ntests", "last_diff"):
        assert ra[key] == rb[key], (what, key, ra, rb)
    assert np.array_equal(a.get_W(), b.get_W()), what
    assert np.array_equal(a.get_H(), b.get_H()), what
    assert np.array_equal(a.get_classification(), b.get_classification()), what


def test_batch_equals_jobs_alone(nmfk, orc):
    pairs = [twins(nmfk, orc, s) for s in MIX]
    # 2 x 2 waves own 128 rows of the updated factor per workgroup (1 x 4: 64)
    plans = [pairs[i][0].plan() for i in (6, 7)]
    assert [plans[0]["H"]["nab"], plans[0]["W"]["nab"], plans[1]["W"]["nab"]] == [24, 32, 33]
    batch = nmfk.Batch(T, 16)
    for (a, _), s in zip(pairs, MIX):
        batch.add(a, s[5], s[6])
    done = dict((id(j), r) for j, r in batch.run())
    assert len(done) == len(MIX)
    for (a, b), s in zip(pairs, MIX):
        rb = b.run(s[5], T, s[6])
        assert_same(a, b, done[id(a)], rb, s)
    # the leftover-iteration jobs really ran past their test budget
    assert not done[id(pairs[-1][0])]["converged"] and done[id(pairs[-1][0])]["iterations"] == 57
    for a, b in pairs:
        a.close(), b.close()


def test_batch_several_chunks_per_poll(nmfk, orc):
    # Three chunks per step: the states reach the host once, after the last of them; jobs that
    # stop in an earlier chunk stay stopped (their tiles return at once), so every result is
    # still the job's result alone, leftover-iteration jobs included.
    pairs = [twins(nmfk, orc, s) for s in MIX]
    batch = nmfk.Batch(T, 16)
    for (a, _), s in zip(pairs, MIX):
        batch.add(a, s[5], s[6])
    done = dict((id(j), r) for j, r in batch.run(chunks_per_poll=3))
    assert len(done) == len(MIX)
    for (a, b), s in zip(pairs, MIX):
        assert_same(a, b, done[id(a)], b.run(s[5], T, s[6]), s)
    for a, b in pairs:
        a.close(), b.close()


def test_batch_admission_mid_run(nmfk, orc):
    # Jobs arrive while others are running (continuous batching): each result is still the
    # job's result alone.
    pairs = [twins(nmfk, orc, s) for s in MIX]
    batch = nmfk.Batch(T, 4)          # fewer slots than jobs: later jobs wait for a free slot
    pending = list(range(len(MIX)))
    results = {}
    while pending or batch.active:
        while pending and batch.active < 4:
            i = pending.pop(0)
            batch.add(pairs[i][0], MIX[i][5], MIX[i][6])
        batch.step(1)
        for j, r in batch.poll():
            results[id(j)] = r
    for (a, b), s in zip(pairs, MIX):
        rb = b.run(s[5], T, s[6])
        assert_same(a, b, results[id(a)], rb, s)
    for a, b in pairs:
        a.close(), b.close()


def test_batch_autotune_and_tilings_keep_bits(nmfk, orc):
    specs = MIX[:5]
    ref = None
    for tune in ("autotune", dict(grid_partial=7, grid_finish=3, order=1, compact_pct=0)):
        pairs = [twins(nmfk, orc, s) for s in specs]
        batch = nmfk.Batch(T, 8)
        for (a, _), s in zip(pairs, specs):
            batch.add(a, s[5], s[6])
        if tune == "autotune":
            best, ms = batch.autotune(5)
            assert ms > 0 and best["order"] in (0, 1)
        else:
            batch.set_tuning(**tune)
        done = dict((id(j), r) for j, r in batch.run())
        out = [(a.get_W(), a.get_H(), done[id(a)]["iterations"]) for a, _ in pairs]
        if ref is None:
            for (a, b), s in zip(pairs, specs):
                assert_same(a, b, done[id(a)], b.run(s[5], T, s[6]), s)
            ref = out
        else:
            for x, y in zip(out, ref):
                assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) and x[2] == y[2]
        for a, b in pairs:
            a.close(), b.close()


def test_batch_rejects_bad_jobs(nmfk, orc):
    batch = nmfk.Batch(T, 2)
    with nmfk.Job(300, 100, 40) as big:          # kt 64: not a batched width
        big.init_random(1)
        with pytest.raises(nmfk.NmfkError):
            batch.add(big, 100, 4)
    with nmfk.Job(300, 100, 4) as j:
        j.init_random(1)
        with pytest.raises(nmfk.NmfkError):
            batch.add(j, 3, 4)                   # nIters < niter_test_conv


def sweep_input(nmfk, orc, n=600, m=40, kstar=3, seed=21):
    V = orc.gen_planted(n, m, kstar, seed, 0.3)
    src = nmfk.Job(n, m, 1)
    src.set_V(V)
    return src, V


def test_sweep_batched_equals_one_at_a_time(nmfk, orc):
    src, V = sweep_input(nmfk, orc)
    ks, R = [2, 3, 4, 5, 9], 3
    a = nmfk.sweep(src, ks, R, 1234, nIters=600, plan_target=64, want_factors=True)
    b = nmfk.sweep(src, ks, R, 1234, nIters=600, plan_target=64, mode=nmfk.SWEEP_ONE_AT_A_TIME,
                   want_factors=True)
    for ra, rb in zip(a["results"], b["results"]):
        for key in ("iterations", "converged", "ntests", "last_diff"):
            assert ra[key] == rb[key]
    assert np.array_equal(a["cls"], b["cls"])
    assert np.array_equal(a["consensus"], b["consensus"])
    for (Wa, Ha), (Wb, Hb) in zip(a["factors"], b["factors"]):
        assert np.array_equal(Wa, Wb) and np.array_equal(Ha, Hb)
    # consensus: exact counts from the returned class vectors
    for q in range(len(ks)):
        c = a["cls"][q]
        ref = sum((c[r][:, None] == c[r][None, :]).astype(np.int32) for r in range(R))
        assert np.array_equal(a["consensus"][q], ref)
        assert np.all(np.diag(a["consensus"][q]) == R)
        assert np.array_equal(a["consensus"][q], a["consensus"][q].T)
    src.close()


def test_sweep_jobs_match_oracle(nmfk, orc):
    src, V = sweep_input(nmfk, orc)
    ks, R, base = [2, 3], 2, 99
    a = nmfk.sweep(src, ks, R, base, nIters=2000, autotune=True, autotune_iters=3, want_factors=True)
    i = 0
    for k in ks:
        for r in range(R):
            W0, H0 = orc.init_WH(V.shape[0], V.shape[1], k, nmfk.sweep_seed(base, k, r))
            o = orc.nmf(V, W0, H0, 2000, 10, 40, fp64=True)
            res = a["results"][i]
            assert res["iterations"] == o["iterations"] and res["converged"] == o["converged"], (k, r)
            assert np.array_equal(a["cls"][ks.index(k)][r], orc.idx_max(o["H"])), (k, r)
            W, H = a["factors"][i]
            assert rel_err(H, o["H"]) < 1e-3 and rel_err(W, o["W"]) < 1e-3, (k, r)
            kl_gpu, kl_ref = orc.kl_div(V, W, H), orc.kl_div(V, o["W"], o["H"])
            assert abs(kl_gpu - kl_ref) / kl_ref < 1e-3
            i += 1
    src.close()


def test_sweep_seed_is_splitmix(nmfk):
    def splitmix(base, k, r):
        z = (base ^ ((k << 32) | r)) & (2**64 - 1)
        z = (z + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        return z ^ (z >> 31)
    for base, k, r in [(0, 0, 0), (1234, 5, 7), (2**63 + 5, 10, 19)]:
        assert nmfk.sweep_seed(base, k, r) == splitmix(base, k, r)


def test_full_sweep_matches_fp64(nmfk, orc):
    # The benchmark's rank sweep at full shape (V 5000 x 200 planted rank 6, k = 2..10 x 20
    # restarts, batched with the autotuner); restart 0 of every rank against the FP64 oracle:
    # stop iteration and class vector exact, KL within 1e-3.
    n, m, ks, R, base = 5000, 200, list(range(2, 11)), 20, 4242
    V = orc.gen_planted(n, m, 6, 777, 0.3)
    src = nmfk.Job(n, m, 1)
    src.set_V(V)
    a = nmfk.sweep(src, ks, R, base, autotune=True, autotune_iters=5, want_factors=True)
    for q, k in enumerate(ks):
        i = q * R
        W0, H0 = orc.init_WH(n, m, k, nmfk.sweep_seed(base, k, 0))
        o = orc.nmf(V, W0, H0, 2000, 10, 40, fp64=True)
        assert a["results"][i]["iterations"] == o["iterations"], k
        assert np.array_equal(a["cls"][q][0], orc.idx_max(o["H"])), k
        W, H = a["factors"][i]
        kl, klo = orc.kl_div(V, W, H), orc.kl_div(V, o["W"], o["H"])
        assert abs(kl - klo) / klo < 1e-3, k
    src.close()


def test_jobs_are_not_freed_under_a_batch_or_a_sharer(nmfk, orc):
    # Closing a job that a batch still steps, or whose V a shared job still reads, would
    # free device memory in use: the bindings refuse it until the job is polled out (or the
    # batch closed) and the sharers are closed.
    V = orc.gen_lognormal(300, 40, 3, 0.0, 1.0)
    src = nmfk.Job(300, 40, 3)
    src.set_V(V)
    src.init_random(1)
    twin = nmfk.Job.shared(src, 4)
    twin.init_random(2)
    with pytest.raises(nmfk.NmfkError, match="Job.shared"):
        src.close()
    batch = nmfk.Batch(T, 4)
    batch.add(twin, 50, 4)
    with pytest.raises(nmfk.NmfkError, match="active batch"):
        twin.close()
    out = batch.run()
    assert [j for j, _ in out] == [twin]
    twin.close()
    src.close()
    assert src.h is None and twin.h is None
    j = nmfk.Job(300, 40, 3)
    j.set_V(V)
    j.init_random(3)
    other = nmfk.Batch(T, 2)
    other.add(j, 50, 4)
    other.close()                    # the batch goes first: the job may then be freed
    j.close()
    batch.close()
