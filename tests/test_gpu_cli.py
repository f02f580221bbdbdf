"""The command-line driver bionmf-gpu_amd/NMF_GPU (the reference's bin/NMF_GPU): options,
messages and output files, and its results equal the job API's on the same seed."""
from __future__ import annotations

import os
import pathlib
import subprocess

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

DRIVER = ROOT / "bionmf-gpu_amd" / "NMF_GPU"


def run_driver(args, seed=5, **env_extra):
    if not DRIVER.exists():
        pytest.fail("bionmf-gpu_amd/NMF_GPU is not built")
    env = dict(os.environ, NMFGPU_SEED=str(seed), **env_extra)
    return subprocess.run([str(DRIVER)] + [str(a) for a in args], capture_output=True, text=True, env=env,
                          timeout=120)


def test_driver_text_io_matches_job_api(nmfk, orc, tmp_path):
    n, m, k = 300, 40, 3
    V = orc.gen_lognormal(n, m, 17, 3.0, 1.0)
    hd = [f"sample_{j}" for j in range(m)]
    lb = [f"gene_{i}" for i in range(n)]
    f = tmp_path / "input.txt"
    nmfk.save_matrix(f, V, nmfk.FMT_TEXT, name="Expression", headers=hd, labels=lb)
    Vf = nmfk.load_matrix(f)[0]                 # the values the driver reads (%g text)
    r = run_driver([f, "-k", k, "-i", 500, "-j", 10, "-t", 20])
    assert r.returncode == 0, r.stderr
    assert "Starting NMF( K=3 )..." in r.stdout and "Distance between V and W*H:" in r.stdout
    W, nm, whd, wlb = nmfk.load_matrix(tmp_path / "input.txt_W.txt")
    H, hnm, hhd, hlb = nmfk.load_matrix(tmp_path / "input.txt_H.txt")
    assert W.shape == (n, k) and H.shape == (k, m)
    assert nm == hnm == "Expression" and wlb == lb and whd == [f"Factor_{i}" for i in range(k)]
    assert hhd == hd and hlb == [f"Factor_{i}" for i in range(k)]
    cls_lines = (tmp_path / "input.txt_classes.txt").read_text().splitlines()
    cls = np.array([int(x.split("\t")[1]) for x in cls_lines], np.uint32)
    with nmfk.Job(n, m, k) as j:
        j.set_V(Vf)
        j.init_random(5)
        res = j.run(500, 10, 20)
        W_ref, H_ref = j.get_W(), j.get_H()
        assert np.array_equal(cls, j.get_classification())
    # text output carries 6 significant digits (%g)
    assert np.allclose(W, W_ref, rtol=1e-5, atol=0) and np.allclose(H, H_ref.T, rtol=1e-5, atol=0)
    word = "converged in" if res["converged"] else "iterations performed"
    assert word in r.stdout and str(res["iterations"]) in r.stdout


def test_driver_binary_io_and_errors(nmfk, orc, tmp_path):
    n, m, k = 200, 30, 4
    V = orc.gen_lognormal(n, m, 3, 2.0, 1.0)
    f = tmp_path / "in.dat"
    nmfk.save_matrix(f, V, nmfk.FMT_BINARY)
    r = run_driver([f, "-b", 0, "-e", 0, "-k", k, "-i", 200])
    assert r.returncode == 0, r.stderr
    W = nmfk.load_matrix(tmp_path / "in.dat_W.dat", nmfk.FMT_BINARY)[0]
    H = nmfk.load_matrix(tmp_path / "in.dat_H.dat", nmfk.FMT_BINARY)[0]
    with nmfk.Job(n, m, k) as j:
        j.set_V(V)
        j.init_random(5)
        j.run(200, 10, 40)
        assert np.array_equal(W, j.get_W()) and np.array_equal(H, j.get_H().T)   # exact: FP32 -> FP64
    # native output keeps H as stored (M x K)
    r = run_driver([f, "-b", 0, "-e", 1, "-k", k, "-i", 200])
    assert r.returncode == 0, r.stderr
    assert nmfk.load_matrix(tmp_path / "in.dat_H.native.dat", nmfk.FMT_NATIVE)[0].shape == (m, k)
    # errors: rank above min(N, M), missing file
    r = run_driver([f, "-b", 0, "-k", m + 1])
    assert r.returncode != 0 and "invalid factorization rank" in r.stderr
    r = run_driver([tmp_path / "missing.txt"])
    assert r.returncode != 0


def test_driver_streams_V_from_host_memory(nmfk, orc, tmp_path):
    # NMFGPU_STREAM_ROWS forces the out-of-core mode the driver picks when V does not fit the
    # device: same stop iteration and classes as the resident run, factors within 1e-3.
    n, m, k = 1000, 60, 4
    V = orc.gen_planted(n, m, 4, 23, 0.3)
    outs = {}
    for mode, env in (("resident", {}), ("streamed", {"NMFGPU_STREAM_ROWS": "192"})):
        d = tmp_path / mode
        d.mkdir()
        f = d / "v.txt"
        nmfk.save_matrix(f, V, nmfk.FMT_TEXT)
        r = run_driver([f, "-k", k, "-i", 800, "-j", 10, "-t", 30], **env)
        assert r.returncode == 0, r.stderr
        assert ("streamed from host memory in 6 blocks of 192 rows" in r.stdout) == (mode == "streamed")
        it_line = [ln for ln in r.stdout.splitlines() if ln.startswith("NMF:")]
        outs[mode] = (it_line, nmfk.load_matrix(d / "v.txt_W.txt")[0], nmfk.load_matrix(d / "v.txt_H.txt")[0],
                      (d / "v.txt_classes.txt").read_text())
    a, b = outs["resident"], outs["streamed"]
    assert a[0] == b[0] and a[3] == b[3]
    assert np.allclose(b[1], a[1], rtol=1e-3, atol=1e-6 * a[1].max())
    assert np.allclose(b[2], a[2], rtol=1e-3, atol=1e-6 * a[2].max())


REF_DRIVER = ROOT / "oracle" / "_ref" / "NMF_GPU_ref"


def test_reference_driver_on_library(nmfk, orc, tmp_path):
    # The reference's own src/NMF_GPU.c + matrix I/O, built against libnmfgpu_amd.so by the
    # integration recipe (make -C oracle refdriver): its main() and nmf() loop drive the
    # drop-in API.  Its outputs equal the job API's run from the same seed.
    if not REF_DRIVER.exists():
        pytest.skip("oracle/_ref/NMF_GPU_ref was not built (reference tree absent at build time)")
    n, m, k = 400, 50, 3
    V = orc.gen_lognormal(n, m, 29, 3.0, 1.0)
    f = tmp_path / "data.txt"
    nmfk.save_matrix(f, V, nmfk.FMT_TEXT)
    Vf = nmfk.load_matrix(f)[0]
    env = dict(os.environ, NMFGPU_SEED="7")
    r = subprocess.run([str(REF_DRIVER), str(f), "-k", str(k), "-i", "600", "-j", "10", "-t", "20"],
                       capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    W = nmfk.load_matrix(tmp_path / "data.txt_W.txt")[0]
    H = nmfk.load_matrix(tmp_path / "data.txt_H.txt")[0]
    assert W.shape == (n, k) and H.shape == (k, m)
    with nmfk.Job(n, m, k) as j:
        j.set_V(Vf)
        j.init_random(7)
        res = j.run(600, 10, 20)
        W_ref, H_ref = j.get_W(), j.get_H()
        dv, dvwh, _ = j.dot_kl()
    # text output carries 6 significant digits (%g)
    assert np.allclose(W, W_ref, rtol=1e-5, atol=0) and np.allclose(H, H_ref.T, rtol=1e-5, atol=0)
    out = r.stdout + r.stderr
    if res["converged"]:
        assert f"converged in {res['iterations']} iterations" in out, out[-2000:]
    dist = float(out.split("Distance between V and W*H:")[1].split()[0])
    assert abs(dist - np.sqrt(dvwh / dv)) <= 1e-4 * dist
