"""Pins the CPU oracle (no GPU needed): known-answer vectors, the reference's own
common.c built in place, closed forms, brute force and invariants chosen so that a dropped
term, a wrong index or a transposed operand fails at least one of them."""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from conftest import GOLDEN, rel_err

FLT_MIN = np.float32(1.17549435e-38)


# ---------------------------------------------------------------- generator

def test_philox_known_answers(orc):
    # Random123 known-answer vectors for Philox4x32-10.
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(orc.philox(ctr, key)) == want


def test_uniform_fill_properties(orc):
    A = orc.uniform_fill(257, 5, seed=123, tag=0, pitch=8)
    vals = A[:, :5]
    assert np.all(A[:, 5:] == 0.0)
    assert vals.min() > 0.0 and vals.max() <= 1.0
    scaled = vals.astype(np.float64) * 2.0**24
    assert np.all(scaled == np.round(scaled))
    assert abs(vals.mean() - 0.5) < 0.02
    # element value depends on (seed, tag, logical index) only, not on the pitch
    B = orc.uniform_fill(257, 5, seed=123, tag=0, pitch=5)
    assert np.array_equal(B, vals)
    C = orc.uniform_fill(257, 5, seed=123, tag=1, pitch=5)
    assert not np.array_equal(B, C)


def test_lognormal_moments(orc):
    V = orc.gen_lognormal(2000, 50, seed=5, mu=5.0, sigma=1.5)
    L = np.log(V.astype(np.float64))
    assert abs(L.mean() - 5.0) < 0.02 and abs(L.std() - 1.5) < 0.02


def test_planted_structure(orc):
    n, m, r = 300, 40, 4
    V = orc.gen_planted(n, m, r, seed=9, sigma=0.0).astype(np.float64)
    # sigma = 0 gives an exact rank-r product: singular values beyond r vanish
    s = np.linalg.svd(V, compute_uv=False)
    assert s[r] / s[0] < 1e-6
    assert np.linalg.matrix_rank(V, tol=s[0] * 1e-6) == r


# ---------------------------------------------------------------- integer parts

def brute_difference(c, l):
    m = len(c)
    return sum((c[i] == c[j]) != (l[i] == l[j]) for i in range(m) for j in range(i + 1, m))


def test_get_difference_bruteforce_and_invariants(orc):
    rng = np.random.default_rng(0)
    for m in (1, 2, 3, 7, 38, 101):
        for _ in range(5):
            c = rng.integers(0, 4, m).astype(np.uint32)
            l = rng.integers(0, 4, m).astype(np.uint32)
            d = orc.get_difference(c, l)
            assert d == brute_difference(c, l)
            assert d == orc.get_difference(l, c)                       # symmetric
            perm = rng.permutation(4).astype(np.uint32)
            assert d == orc.get_difference(perm[c], l)                 # label names do not matter
            assert orc.get_difference(c, perm[c]) == 0                 # same partition -> 0
    # first test compares against all zeros: pairs in different classes
    c = np.array([0, 1, 1, 2], np.uint32)
    assert orc.get_difference(c, np.zeros(4, np.uint32)) == 5


def test_get_difference_matches_reference_common_c(orc):
    ref = orc.ref_common()
    if ref is None:
        pytest.skip("oracle/_ref/libref_common.so not built (reference tree absent)")
    rng = np.random.default_rng(1)
    for m in (2, 5, 38, 200, 1000):
        c = rng.integers(0, 9, m).astype(np.uint32)
        l = rng.integers(0, 9, m).astype(np.uint32)
        assert orc.get_difference(c, l) == ref.get_difference(c, l, m)


def test_stop_rule_sequences(orc):
    S = 40
    # all-zero differences: converges at test S+1 (src/NMF_GPU.c:695-708)
    assert orc.stop_rule([0] * 100, S) == S + 1
    # README sample: 430 iterations with -j 10 -t 40 = 43 tests -> two changed tests first
    assert orc.stop_rule([5, 3] + [0] * 60, S) * 10 == 430
    # a change resets the counter
    assert orc.stop_rule([0] * 20 + [1] + [0] * 100, S) == 21 + S + 1
    assert orc.stop_rule([1] * 50, S) == 0
    assert orc.stop_rule([0, 0, 0], 0) == 1


# ---------------------------------------------------------------- argmax / clamp

def test_idx_max_plain_and_ties(orc):
    rng = np.random.default_rng(2)
    A = rng.random((200, 13))
    assert np.array_equal(orc.idx_max(A), np.argmax(A, axis=1))
    # k <= 32: lowest index among tied maxima
    T = np.zeros((3, 10))
    T[0, [2, 7]] = 1.0
    T[1, :] = 0.5
    T[2, [9, 4, 6]] = 3.0
    assert list(orc.idx_max(T)) == [2, 0, 4]


def test_idx_max_reference_layout_k_over_32(orc):
    # 33 <= k <= 96: 32 reference threads scan columns t, t+32, t+64; ties go to the
    # smaller (col mod 32) first, then to the smaller column (derived from
    # src/GPU_kernels.cu:2603-2830 by hand).
    k = 40
    rows = []
    expect = []
    r = np.zeros(k); r[[1, 33]] = 2.0; rows.append(r); expect.append(1)     # same thread: first column
    r = np.zeros(k); r[[5, 34]] = 2.0; rows.append(r); expect.append(34)    # 34 % 32 = 2 < 5
    r = np.zeros(k); r[[31, 32]] = 2.0; rows.append(r); expect.append(32)   # 32 % 32 = 0
    r = np.zeros(k); r[[3, 39]] = 2.0; rows.append(r); expect.append(3)     # 39 % 32 = 7 > 3
    r = np.arange(k, dtype=float); rows.append(r); expect.append(k - 1)     # strict maximum
    assert list(orc.idx_max(np.array(rows))) == expect


def test_idx_max_nan_follows_reference_threads(orc):
    nan = np.nan
    # k=4: threads 0..3 start from their own column; NaN at thread 0 never loses
    # ('<' with NaN is false), so thread 0's NaN survives the tree.
    assert list(orc.idx_max(np.array([[nan, 1.0, 2.0, 3.0]]))) == [0]
    # NaN at thread 2: the tree compares thread 0 (value 1) with thread 2 (NaN): 1 < NaN is
    # false, thread 0 keeps; then thread 0 vs thread 1 (5) -> 1.
    assert list(orc.idx_max(np.array([[1.0, 5.0, nan, 0.5]]))) == [1]


def test_adjust_special_values(orc):
    x = np.array([[np.nan, -0.0, 0.0, 1e-39, -1.0, float(FLT_MIN), 1.0, np.inf, -np.inf, 3e-38]])
    want = [FLT_MIN, FLT_MIN, FLT_MIN, FLT_MIN, FLT_MIN, FLT_MIN, 1.0, np.inf, FLT_MIN, 3e-38]
    for fp64 in (True, False):
        got = orc.adjust(x, fp64=fp64)[0]
        assert np.array_equal(got.astype(np.float64), np.array(want, np.float64).astype(got.dtype).astype(np.float64))


# ---------------------------------------------------------------- elementwise ops

def _mul_div_inputs(dt):
    # Column-dependent divisors on a non-square shape, so a row/column mix-up in acc[] fails.
    rng = np.random.default_rng(31)
    A = (rng.random((37, 5)) * 3 + 0.01).astype(dt)
    Aux = (rng.random((37, 5)) * 7 + 0.01).astype(dt)
    acc = np.array([0.3, 1.7, 3.1, 11.0, 0.07], dt)
    return A, Aux, acc


@pytest.mark.parametrize("fp64", [True, False])
def test_mul_div_multiplies_then_divides(orc, fp64):
    # src/GPU_kernels.cu:2029-2033: A = (A * Aux) / acc[col], one rounding per operation.
    dt = np.float64 if fp64 else np.float32
    A, Aux, acc = _mul_div_inputs(dt)
    want = np.divide(np.multiply(A, Aux, dtype=dt), acc[None, :], dtype=dt)
    got = orc.mul_div(A, Aux, acc, fp64=fp64)
    assert got.dtype == dt and np.array_equal(got, want)
    # The inputs tell the orders apart: dividing first rounds differently on some entries.
    other1 = np.multiply(A, np.divide(Aux, acc[None, :], dtype=dt), dtype=dt)
    other2 = np.multiply(np.divide(A, acc[None, :], dtype=dt), Aux, dtype=dt)
    assert np.any(other1 != want) and np.any(other2 != want)
    # and a row-indexed divisor would differ too
    assert not np.array_equal(got, np.divide(np.multiply(A, Aux, dtype=dt)[:5, :].T, acc[None, :], dtype=dt).T)


def test_mul_div_overflow_reveals_order(orc):
    # In FP32 (A * Aux) overflows before the division; A * (Aux / acc) would not.
    A = np.array([[1e30, 2.0]], np.float32)
    Aux = np.array([[1e30, 3.0]], np.float32)
    acc = np.array([1e30, 4.0], np.float32)
    got = orc.mul_div(A, Aux, acc, fp64=False)
    assert np.isposinf(got[0, 0]) and got[0, 1] == np.float32(1.5)
    # FP64 has the range for the product: (a * a) / a with a = float(1e30f)
    a = float(A[0, 0])
    got64 = orc.mul_div(A.astype(np.float64), Aux.astype(np.float64), acc.astype(np.float64))
    assert got64[0, 0] == (a * a) / a and got64[0, 1] == 1.5


@pytest.mark.parametrize("fp64", [True, False])
def test_div_sub_both_branches(orc, fp64):
    # src/GPU_kernels.cu:1676-1700: div_operator != 0 -> A = B ./ A, else A = B - A.
    dt = np.float64 if fp64 else np.float32
    rng = np.random.default_rng(32)
    A = (rng.random((23, 9)) * 5 + 0.1).astype(dt)
    B = (rng.random((23, 9)) * 5 + 0.1).astype(dt)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = orc.div_sub(A, B, True, fp64=fp64)
        assert np.array_equal(q, np.divide(B, A, dtype=dt))
        assert not np.allclose(q, np.divide(A, B, dtype=dt))       # operands not swapped
        d = orc.div_sub(A, B, False, fp64=fp64)
        assert np.array_equal(d, np.subtract(B, A, dtype=dt))
        assert np.any(np.sign(d) != np.sign(np.subtract(A, B, dtype=dt)))  # B - A, not A - B
        # special values of the IEEE division and subtraction
        Z = np.array([[0.0, 0.0, 2.0, np.inf]], dt)
        Y = np.array([[3.0, 0.0, 2.0, 1.0]], dt)
        q = orc.div_sub(Z, Y, True, fp64=fp64)[0]
        assert np.isposinf(q[0]) and np.isnan(q[1]) and q[2] == 1.0 and q[3] == 0.0
        d = orc.div_sub(Z, Y, False, fp64=fp64)[0]
        assert d[0] == 3.0 and d[1] == 0.0 and d[2] == 0.0 and np.isneginf(d[3])


# ---------------------------------------------------------------- updates

def numpy_update_H(V, W, H):
    WH = W @ H.T
    return H * ((V / WH).T @ W) / W.sum(axis=0)


def numpy_update_W(V, W, H):
    WH = W @ H.T
    return W * ((V / WH) @ H) / H.sum(axis=0)


@pytest.mark.parametrize("n,m,k", [(9, 5, 1), (16, 8, 2), (30, 11, 3), (50, 31, 7), (40, 70, 16)])
def test_updates_match_matrix_form(orc, n, m, k):
    rng = np.random.default_rng(n * 100 + k)
    V = rng.random((n, m)) + 0.1
    W = rng.random((n, k)) + 0.05
    H = rng.random((m, k)) + 0.05
    Hn = orc.update_H(V, W, H)
    assert rel_err(Hn, numpy_update_H(V, W, H)) < 1e-12
    Wn = orc.update_W(V, W, Hn)
    assert rel_err(Wn, numpy_update_W(V, W, Hn)) < 1e-12
    # the FP32 form follows the same algorithm within single-precision rounding
    Hf = orc.update_H(V, W, H, fp64=False)
    assert rel_err(Hf, Hn) < 1e-5


def test_exact_factorization_is_a_fixed_point(orc):
    rng = np.random.default_rng(3)
    W = rng.random((40, 4)) + 0.1
    H = rng.random((25, 4)) + 0.1
    V = W @ H.T
    assert rel_err(orc.update_H(V, W, H), H) < 1e-13
    assert rel_err(orc.update_W(V, W, H), W) < 1e-13


def test_kl_is_non_increasing(orc):
    # Lee & Seung (2001): the multiplicative KL updates never increase D(V||WH).
    V = orc.gen_lognormal(120, 30, seed=4, mu=0.0, sigma=1.0).astype(np.float64)
    W, H = (x.astype(np.float64) for x in orc.init_WH(120, 30, 5, seed=5))
    prev = orc.kl_div(V, W, H)
    for _ in range(60):
        H = orc.update_H(V, W, H)
        cur = orc.kl_div(V, W, H)
        assert cur <= prev * (1 + 1e-12)
        W = orc.update_W(V, W, H)
        nxt = orc.kl_div(V, W, H)
        assert nxt <= cur * (1 + 1e-12)
        prev = nxt


@pytest.mark.parametrize("fp64", [True, False])
def test_scale_equivariance_is_exact(orc, fp64):
    # V -> 4V (a power of two) scales H by 4 and leaves W unchanged, bit for bit.
    V = orc.gen_lognormal(60, 20, seed=6, mu=0.0, sigma=1.0)
    W0, H0 = orc.init_WH(60, 20, 3, seed=7)
    a = orc.nmf(V, W0, H0, 57, 5, 2, fp64=fp64)
    b = orc.nmf(4 * V, W0, H0, 57, 5, 2, fp64=fp64)
    assert np.array_equal(a["W"], b["W"])
    assert np.array_equal(4 * a["H"], b["H"])
    assert np.array_equal(a["cls"], b["cls"]) and a["iterations"] == b["iterations"]


def test_dot_and_kl_closed_forms(orc):
    V = np.array([[1.0, 2.0], [3.0, 0.0]])
    W = np.array([[1.0], [2.0]])
    H = np.array([[1.0], [1.0]])   # WH = [[1,1],[2,2]]
    dv, dvwh = orc.dot_VWH(V, W, H)
    assert dv == 14.0 and dvwh == 0.0 + 1.0 + 1.0 + 4.0
    kl = orc.kl_div(V, W, H)
    want = (0 - 1 + 1) + (2 * np.log(2) - 2 + 1) + (3 * np.log(1.5) - 3 + 2) + (0 - 0 + 2)
    assert abs(kl - want) < 1e-14
    assert orc.kl_div(W @ H.T, W, H) == 0.0


# ---------------------------------------------------------------- control flow

def test_nmf_control_flow(orc):
    V = orc.gen_lognormal(30, 12, seed=8, mu=0.0, sigma=1.0)
    W0, H0 = orc.init_WH(30, 12, 2, seed=9)
    r = orc.nmf(V, W0, H0, nIters=7, niter_test_conv=10, stop_threshold=3)
    assert r["ntests"] == 0 and r["iterations"] == 7 and not r["converged"]
    r = orc.nmf(V, W0, H0, nIters=25, niter_test_conv=10, stop_threshold=30)
    assert r["ntests"] == 2 and r["iterations"] == 25
    r = orc.nmf(V, W0, H0, nIters=2000, niter_test_conv=4, stop_threshold=3, max_trace=600)
    assert r["converged"] and r["iterations"] == r["ntests"] * 4
    assert orc.stop_rule(r["diff_trace"], 3) == r["ntests"]
    # the class vector output is the argmax of the final (clamped) H
    assert np.array_equal(r["cls"], orc.idx_max(r["H"]))
    # the trace's consecutive vectors reproduce the recorded differences
    tr = r["cls_trace"]
    assert r["diff_trace"][0] == orc.get_difference(tr[0], np.zeros_like(tr[0]))
    for t in range(1, len(tr)):
        assert r["diff_trace"][t] == orc.get_difference(tr[t], tr[t - 1])


def test_manual_loop_equals_nmf(orc):
    # the driver is exactly T x (update_H, update_W), adjust(H), adjust(W), argmax
    V = orc.gen_lognormal(25, 9, seed=10, mu=0.0, sigma=1.0)
    W0, H0 = orc.init_WH(25, 9, 3, seed=11)
    W, H = W0.astype(np.float64), H0.astype(np.float64)
    for _ in range(6):
        H = orc.update_H(V, W, H)
        W = orc.update_W(V, W, H)
    H, W = orc.adjust(H), orc.adjust(W)
    for _ in range(2):
        H = orc.update_H(V, W, H)
        W = orc.update_W(V, W, H)
    r = orc.nmf(V, W0, H0, nIters=8, niter_test_conv=6, stop_threshold=5)
    assert np.array_equal(r["W"], W) and np.array_equal(r["H"], H)


# ---------------------------------------------------------------- fixtures

def test_golden_tiny_regression(orc):
    g = np.load(GOLDEN / "tiny.npz")
    for ci in range(int(g["ncases"])):
        p = f"c{ci}_"
        n, m, k, sv, si = (int(x) for x in g[p + "shape"])
        V = orc.gen_lognormal(n, m, sv, 0.0, 1.0)
        W0, H0 = orc.init_WH(n, m, k, si)
        assert np.array_equal(V, g[p + "V"]) and np.array_equal(W0, g[p + "W0"])
        assert np.array_equal(H0, g[p + "H0"])
        H1 = orc.update_H(V, W0, H0)
        W1 = orc.update_W(V, W0, H1)
        assert np.array_equal(H1, g[p + "H1"]) and np.array_equal(W1, g[p + "W1"])
        r = orc.nmf(V, W0, H0, 200, 5, 3, max_trace=64)
        assert np.array_equal(r["cls"], g[p + "cls"]) and r["iterations"] == int(g[p + "iterations"])


def test_golden_config0_f32_matches_f64(orc):
    g = np.load(GOLDEN / "config0.npz")
    assert np.array_equal(g["f64_cls"], g["f32_cls"])
    assert int(g["f64_iterations"]) == int(g["f32_iterations"])
    assert abs(float(g["f32_kl"]) - float(g["f64_kl"])) / float(g["f64_kl"]) < 1e-6


def test_all_aml_real_data_sample(orc):
    # README.md:218-249: ALL_AML, k=2, -j 10 -t 40 -i 2000 converged (430 iterations there,
    # from an unknown random seed) with distance 0.566049; a loose real-data check.
    V = np.load(GOLDEN / "all_aml.npz")["V"].astype(np.float32)
    W0, H0 = orc.init_WH(5000, 38, 2, seed=3)
    r = orc.nmf(V, W0, H0, 2000, 10, 40, fp64=False)
    assert r["converged"] and r["iterations"] % 10 == 0
    dv, dvwh = orc.dot_VWH(V, r["W"], r["H"], fp64=False)
    dist = np.sqrt(dvwh) / np.sqrt(dv)
    assert abs(dist - 0.566049) < 0.02
