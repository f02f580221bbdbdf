"""Writes the golden fixtures under tests/golden/ from the CPU oracle (and converts the
reference's ALL_AML test matrix into a compact array).

    python tests/golden/make_golden.py            # rewrite the small fixtures
    python tests/golden/make_golden.py --ref DIR  # reference tree (default /root/reference)
    python tests/golden/make_golden.py --large    # also the two FP64 runs at full size
                                                  # (about 2 and 12 minutes on 8 cores)

Fixtures:
  tiny.npz       small shapes: V, W0, H0 (Philox seeds), W/H after 1 and 10 iterations of
                 the FP64 update order, and a whole FP64 nmf() run (class vector, iteration
                 count, per-test differences).
  config0.npz    5000 x 38 lognormal(5, 1.5^2) V, k=3: class vector, stop iteration and KL
                 of the FP64 and FP32 oracle runs (T=10, stop_threshold=40, 2000 iterations).
  all_aml.npz    test/ALL_AML_data_noLabels.txt of the reference (5000 x 38 integers) as uint16.
  config1_fp64.npz  the single large job at full size: V 20000 x 1000 planted rank 16
                 (seed 20250101, sigma 0.3), k = 16, W0/H0 from seed 7, 500 iterations with a
                 test every 10 and no stop: FP64 oracle's KL, Frobenius sums, class vector,
                 per-test differences, H and every 20th row of W.
  config3_fp64.npz  the dense-rank job: V 50000 x 2000 planted rank 48 (seed 4848), k = 64,
                 W0/H0 from seed 9, 300 iterations: the same quantities, every 25th row of W.
"""
from __future__ import annotations

import argparse
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1] / "oracle"))
import oracle as O  # noqa: E402

TINY_CASES = [  # n, m, k, seed_v, seed_init
    (16, 8, 2, 11, 12),
    (12, 7, 3, 13, 14),
    (33, 17, 4, 15, 16),
    (64, 40, 5, 17, 18),
    (70, 45, 12, 19, 20),
    (90, 50, 33, 21, 22),
]

CONFIG0 = dict(n=5000, m=38, k=3, seed_v=1, seed_init=2, mu=5.0, sigma=1.5)


def tiny():
    out = {}
    for ci, (n, m, k, sv, si) in enumerate(TINY_CASES):
        V = O.gen_lognormal(n, m, sv, 0.0, 1.0)
        W0, H0 = O.init_WH(n, m, k, si)
        W, H = W0.astype(np.float64), H0.astype(np.float64)
        H = O.update_H(V, W, H)
        W = O.update_W(V, W, H)
        W1, H1 = W.copy(), H.copy()
        for _ in range(9):
            H = O.update_H(V, W, H)
            W = O.update_W(V, W, H)
        r = O.nmf(V, W0, H0, nIters=200, niter_test_conv=5, stop_threshold=3, fp64=True, max_trace=64)
        p = f"c{ci}_"
        out.update({p + "shape": np.array([n, m, k, sv, si]), p + "V": V, p + "W0": W0, p + "H0": H0,
                    p + "W1": W1, p + "H1": H1, p + "W10": W, p + "H10": H,
                    p + "cls": r["cls"], p + "iterations": np.array(r["iterations"]),
                    p + "converged": np.array(int(r["converged"])), p + "ntests": np.array(r["ntests"]),
                    p + "diff_trace": r["diff_trace"]})
    out["ncases"] = np.array(len(TINY_CASES))
    return out


def config0():
    c = CONFIG0
    V = O.gen_lognormal(c["n"], c["m"], c["seed_v"], c["mu"], c["sigma"])
    W0, H0 = O.init_WH(c["n"], c["m"], c["k"], c["seed_init"])
    out = {"params": np.array([c["n"], c["m"], c["k"], c["seed_v"], c["seed_init"]]),
           "V_checksum": np.array(float(V.astype(np.float64).sum()))}
    for fp64 in (True, False):
        r = O.nmf(V, W0, H0, 2000, 10, 40, fp64=fp64, max_trace=256)
        tag = "f64" if fp64 else "f32"
        out[f"{tag}_cls"] = r["cls"]
        out[f"{tag}_iterations"] = np.array(r["iterations"])
        out[f"{tag}_converged"] = np.array(int(r["converged"]))
        out[f"{tag}_diff_trace"] = r["diff_trace"]
        out[f"{tag}_kl"] = np.array(O.kl_div(V, r["W"], r["H"], fp64=True))
        out[f"{tag}_W_head"] = r["W"][:64].astype(np.float64)
        out[f"{tag}_H"] = r["H"].astype(np.float64)
    return out


LARGE = {  # name: n, m, k, kstar, seed_v, seed_init, iterations, W row step
    "config1_fp64": (20000, 1000, 16, 16, 20250101, 7, 500, 20),
    "config3_fp64": (50000, 2000, 64, 48, 4848, 9, 300, 25),
}


def large(name):
    """FP64 oracle run of a full-size workload (V planted, sigma 0.3, test every 10
    iterations, never stopping early), reduced to what a device run is compared with."""
    n, m, k, kstar, sv, si, iters, step = LARGE[name]
    V = O.gen_planted(n, m, kstar, sv, 0.3)
    W0, H0 = O.init_WH(n, m, k, si)
    r = O.nmf(V, W0, H0, iters, 10, 10**9, fp64=True, max_trace=iters // 10 + 1)
    dv, dvwh = O.dot_VWH(V, r["W"], r["H"])
    rows = np.arange(0, n, step)
    probe = [(i * 7919) % n for i in range(64)]
    return {"params": np.array([n, m, k, kstar, sv, si, iters, step]),
            "V_checksum": np.array(float(V.astype(np.float64).sum())),
            "V_probe_idx": np.array(probe), "V_probe": V[probe, [(i * 104729) % m for i in range(64)]],
            "kl": np.array(O.kl_div(V, r["W"], r["H"], fp64=True)), "dot_V": np.array(dv),
            "dot_VWH": np.array(dvwh), "cls": r["cls"], "ntests": np.array(r["ntests"]),
            "iterations": np.array(r["iterations"]), "diff_trace": r["diff_trace"],
            "H": r["H"].astype(np.float32), "W_rows": r["W"][rows].astype(np.float32)}


def all_aml(ref: pathlib.Path):
    src = ref / "test" / "ALL_AML_data_noLabels.txt"
    A = np.loadtxt(src, dtype=np.float64)
    assert A.shape == (5000, 38) and A.min() >= 0 and A.max() < 65536 and np.all(A == np.round(A))
    return {"V": A.astype(np.uint16)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--large", action="store_true", help="also the full-size FP64 runs")
    ap.add_argument("--only-large", action="store_true", help="only the full-size FP64 runs")
    args = ap.parse_args()
    if args.large or args.only_large:
        for name in LARGE:
            np.savez_compressed(HERE / f"{name}.npz", **large(name))
            print("written", name, flush=True)
    if args.only_large:
        return
    np.savez_compressed(HERE / "tiny.npz", **tiny())
    np.savez_compressed(HERE / "config0.npz", **config0())
    ref = pathlib.Path(args.ref)
    if (ref / "test" / "ALL_AML_data_noLabels.txt").exists():
        np.savez_compressed(HERE / "all_aml.npz", **all_aml(ref))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
