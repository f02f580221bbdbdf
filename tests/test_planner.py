"""The launch planner of the half-step kernels (host code, no GPU): which kernel path a rank
gets, how the contraction range is split, and the process-wide tunables it reads."""
from __future__ import annotations

import ctypes as C

import pytest

PATH_AUTO, PATH_VALU, PATH_MFMA32, PATH_BF16 = 0, 1, 2, 3
DEFAULTS = {0: 512, 1: 1024, 2: 8, 3: 64, 4: 1, 5: 4, 6: 256, 7: 2048}


class HPlan(C.Structure):
    _fields_ = [(f, C.c_uint32) for f in ("nab", "S", "bchunk", "ap", "nw", "path", "lb", "ke")]


@pytest.fixture
def lib(pkg):
    lib = pkg.load()
    lib.nmfk_plan_half_path.restype = None
    lib.nmfk_plan_half_path.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(HPlan)]
    lib.nmfk_plan_half_target.restype = None
    lib.nmfk_plan_half_target.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32,
                                          C.POINTER(HPlan)]
    saved = {key: lib.nmfk_tune_get(key) for key in DEFAULTS}
    yield lib
    for key, value in saved.items():
        lib.nmfk_tune_set(key, value)


def plan(lib, p, q, kt, path=PATH_AUTO):
    pl = HPlan()
    lib.nmfk_plan_half_path(p, q, kt, path, C.byref(pl))
    return pl


def test_tunable_defaults_and_bounds(lib):
    assert {key: lib.nmfk_tune_get(key) for key in DEFAULTS} == DEFAULTS
    assert lib.nmfk_tune_get(8) == -1 and lib.nmfk_tune_get(-1) == -1
    assert lib.nmfk_tune_set(8, 1) != 0 and lib.nmfk_tune_set(1, 0) != 0
    assert lib.nmfk_tune_get(1) == 1024


@pytest.mark.parametrize("kt,path", [(4, PATH_VALU), (8, PATH_VALU), (16, PATH_MFMA32), (32, PATH_MFMA32),
                                     (64, PATH_BF16)])
def test_automatic_path_by_rank(lib, kt, path):
    assert plan(lib, 1000, 20000, kt).path == path


def test_requested_paths_fall_back_where_the_kernel_cannot_run(lib):
    assert plan(lib, 1000, 20000, 16, PATH_BF16).path == PATH_MFMA32   # BF16 needs kt % 32 == 0
    assert plan(lib, 1000, 20000, 8, PATH_MFMA32).path == PATH_VALU    # matrix cores need kt % 16 == 0
    assert plan(lib, 1000, 20000, 64, PATH_MFMA32).path == PATH_MFMA32
    assert plan(lib, 1000, 20000, 32, PATH_VALU).path == PATH_VALU


SHAPES = [(1, 1), (3, 7), (38, 5000), (5000, 38), (200, 5000), (1000, 20000), (20000, 1000), (2000, 50000),
          (50000, 2000), (60000, 20), (20, 60000), (3000, 4_300_000)]


@pytest.mark.parametrize("p,q", SHAPES)
@pytest.mark.parametrize("kt,path", [(4, PATH_VALU), (8, PATH_VALU), (16, PATH_VALU), (16, PATH_MFMA32),
                                     (32, PATH_MFMA32), (64, PATH_MFMA32), (64, PATH_BF16)])
def test_plan_covers_both_ranges(lib, p, q, kt, path):
    pl = plan(lib, p, q, kt, path)
    assert pl.path == path
    if path == PATH_VALU:
        step, rows_a = pl.lb, 64 // pl.lb * pl.ap
        assert pl.nw == 4 and pl.lb in (1, 2, 4, 8)
        # lane groups: the largest count that still gives every lane a row of the factor
        # (none for a factor of more than 32 ap rows)
        assert pl.lb == 1 or p <= 64 * pl.ap // pl.lb
        assert pl.lb == 8 or p > 64 * pl.ap // (2 * pl.lb)
    else:
        step, rows_a = (32 if path == PATH_BF16 else 16), 64 * pl.ap
        # matrix cores: 1 x 4 waves, or 2 x 2 for BF16 over fewer than 4096 rows of U and for
        # FP32 at kt 16/32 over 1000-4095 rows where that gives at least 240 workgroups
        assert pl.nw == 4 and pl.lb == 1
        if path == PATH_BF16 and q < 4096:
            assert pl.ap == 2
        elif path == PATH_MFMA32 and kt <= 32 and 1000 <= q < 4096:
            assert pl.ap == 1 or (pl.nab + 1) * pl.S >= 240
        else:
            assert pl.ap == 1
    assert pl.nab == -(-p // rows_a)
    assert pl.bchunk % step == 0 and pl.S >= 1 and pl.S <= DEFAULTS[3]
    assert (pl.S - 1) * pl.bchunk < q <= pl.S * pl.bchunk   # every split holds rows, together all q


def test_bf16_plans_of_the_dense_rank_workload(lib):
    # V 50000 x 2000, k = 64: the H half (rows of H against the 50000 rows of V) in 8 splits
    # of 6272 rows, 33 x 8 = 264 workgroups (with the denominator column), 1 x 4 waves; the W
    # half in one split over the 2000 columns, 2 x 2 waves (128 rows of W per workgroup).
    h, w = plan(lib, 2000, 50000, 64), plan(lib, 50000, 2000, 64)
    assert (h.path, h.nab, h.S, h.bchunk, h.ap, h.nw) == (PATH_BF16, 32, 8, 6272, 1, 4)
    assert (w.path, w.nab, w.S, w.bchunk, w.ap, w.nw) == (PATH_BF16, 391, 1, 2016, 2, 4)
    # the FP32 matrix-core kernel keeps its own target of 1024 workgroups
    assert plan(lib, 2000, 50000, 64, PATH_MFMA32).S == 32


def test_tunables_move_the_plan(lib):
    # c1's H half, 16 workgroup columns.  The minimum of 8 steps per wave caps the splits at
    # ceil(20000 / (4 * 16 * 8)) = 40 under the 1024-workgroup target; balanced over the CUs
    # that becomes 30 (672 rows each).  With any tunable changed the planner does what the
    # tunables say: 512 workgroups ask for 32 splits: 625 rows, rounded up to 640 (16-row
    # steps), 32 splits.  2048 workgroups with 1 step per wave ask for 128, capped at 64: 313
    # rows -> 320, 63 splits.
    base = plan(lib, 1000, 20000, 16)
    assert (base.nab, base.S, base.bchunk) == (16, 30, 672)
    assert lib.nmfk_tune_set(1, DEFAULTS[1] // 2) == 0
    assert (plan(lib, 1000, 20000, 16).S, plan(lib, 1000, 20000, 16).bchunk) == (32, 640)
    assert lib.nmfk_tune_set(2, 1) == 0 and lib.nmfk_tune_set(1, 2 * DEFAULTS[1]) == 0
    assert (plan(lib, 1000, 20000, 16).S, plan(lib, 1000, 20000, 16).bchunk) == (63, 320)
    assert lib.nmfk_tune_set(6, 512) == 0
    assert plan(lib, 2000, 50000, 64).S == 16
    assert lib.nmfk_tune_set(4, 2) == 0 and lib.nmfk_tune_set(5, 2) == 0
    pl = plan(lib, 1000, 20000, 16)
    assert (pl.ap, pl.nw, pl.nab) == (2, 4, 8)
    assert lib.nmfk_tune_set(4, 3) == 0                     # 3 x 2 is no kernel shape: default 1 x 4
    pl = plan(lib, 1000, 20000, 16)
    assert (pl.ap, pl.nw) == (1, 4)


def busiest_cu_load(nab, S, kt):
    # workgroups the busiest CU of an XCD runs (8 XCDs of 32 CUs, 4 resident four-wave
    # workgroups per CU at kt 16 and 3 at kt 32, queued ones evening out) per split
    per = -(-(nab + 1) * S // 8)
    resident = 4 if kt <= 16 else 3
    return (-(-per // 32) if per <= 32 * resident else per / 32) / S


@pytest.mark.parametrize("p,q,kt,S", [(1000, 20000, 16, 30), (1000, 20000, 32, 30), (1000, 10000, 16, 15),
                                      (1000, 30000, 16, 30), (500, 20000, 16, 28), (500, 60000, 16, 56),
                                      (1000, 40000, 16, 63), (2000, 20000, 16, 32), (3000, 8000, 32, 16),
                                      (30000, 1000, 16, 2), (8000, 3000, 32, 4), (14000, 600, 16, 2)])
def test_balanced_splits(lib, p, q, kt, S):
    # A lone FP32 job's split count: the least busiest-CU load among half the default split
    # count up to the default, the fewest splits on a tie, at least 240 workgroups.
    pl = plan(lib, p, q, kt)
    assert pl.S == S
    t = HPlan()
    lib.nmfk_plan_half_target(p, q, kt, PATH_AUTO, 1024, C.byref(t))     # the same job, unbalanced
    assert (t.nab, t.ap) == (pl.nab, pl.ap) or pl.ap == 2
    if pl.ap == 1:
        assert S <= t.S <= 2 * S + 1
        assert busiest_cu_load(pl.nab, S, kt) <= busiest_cu_load(t.nab, t.S, kt) + 1e-12
    assert (pl.nab + 1) * S >= 240 or S == t.S


def test_small_lone_job_plan(lib):
    # A lone FP32 job below 240 workgroups at 8 steps per wave group takes 4.  V 5000 x 38,
    # k = 3 (padded to 16), H half: 1 workgroup column; 8 steps of 4 x 16 rows cap the 5000
    # rows at 10 splits (20 workgroups), 4 steps at 20 splits of 256 rows (40 workgroups).
    h = plan(lib, 38, 5000, 16)
    assert (h.path, h.nab, h.S, h.bchunk, h.ap) == (PATH_MFMA32, 1, 20, 256, 1)
    # V 2500 x 1000, k = 16: 85 -> 170 workgroups (H half), 82 -> 164 (W half, 1 x 4)
    h, w = plan(lib, 1000, 2500, 16), plan(lib, 2500, 1000, 16)
    assert ((h.nab + 1) * h.S, (w.nab + 1) * w.S, w.ap) == (170, 164, 1)
    # at 240 workgroups or more the 8 steps stay: V 14000 x 600, k = 12 (H half balanced to
    # 11 x 23 = 253, W half 220 x 2 = 440 in 304-row splits) and the H half of V 10000 x 1000
    # (balanced to 17 x 15 = 255)
    h, w = plan(lib, 600, 14000, 16), plan(lib, 14000, 600, 16)
    assert ((h.nab + 1) * h.S, (w.nab + 1) * w.S, w.bchunk) == (253, 440, 304)
    assert plan(lib, 1000, 10000, 16).S == 15
    # a workgroup target (batched and served plans) or a changed tunable keeps 8 steps:
    # 10 splits of 512 rows
    t = HPlan()
    lib.nmfk_plan_half_target(38, 5000, 16, PATH_AUTO, 16, C.byref(t))
    assert (t.S, t.bchunk) == (10, 512)
    assert lib.nmfk_tune_set(1, DEFAULTS[1] // 2) == 0
    assert plan(lib, 38, 5000, 16).S == 10


def test_fp32_short_range_plan(lib):
    # c1's W half (20000 rows of W against the 1000 rows of Vt): 2 x 2 waves, 128 rows of W per
    # workgroup (157 columns), at least 12 steps of 16 rows per wave group, so at most
    # ceil(1000 / 384) = 3 splits of 336 rows.
    w = plan(lib, 20000, 1000, 16)
    assert (w.path, w.ap, w.nw, w.nab, w.S, w.bchunk) == (PATH_MFMA32, 2, 4, 157, 3, 336)
    assert (plan(lib, 20000, 1000, 32).ap, plan(lib, 20000, 999, 16).ap, plan(lib, 20000, 4096, 16).ap) == (2, 1, 1)
    # too few workgroups in 2 x 2 (V 2500 x 1000: 9 x 7 = 63 tiles for the H half): 1 x 4
    assert (plan(lib, 1000, 2500, 16).ap, plan(lib, 2500, 1000, 16).ap) == (1, 1)
    assert plan(lib, 20000, 2000, 64, PATH_MFMA32).ap == 1          # kt 64 keeps 1 x 4 in FP32
    # an explicit workgroup target (batched plans) keeps 1 x 4
    t = HPlan()
    lib.nmfk_plan_half_target(20000, 1000, 16, PATH_AUTO, 64, C.byref(t))
    assert (t.ap, t.nw) == (1, 4)
    # so do changed step or shape tunables
    assert lib.nmfk_tune_set(2, 6) == 0
    assert plan(lib, 20000, 1000, 16).ap == 1
    assert lib.nmfk_tune_set(2, 8) == 0 and lib.nmfk_tune_set(4, 4) == 0 and lib.nmfk_tune_set(5, 1) == 0
    assert (plan(lib, 20000, 1000, 16).ap, plan(lib, 20000, 1000, 16).nw) == (4, 4)


def test_plan_rank_width(lib):
    # FP32 matrix cores at kt 16 sum only a small rank's columns in GEMM1; every other plan
    # keeps the full padded width.
    lib.nmfk_plan_set_rank.restype = None
    lib.nmfk_plan_set_rank.argtypes = [C.POINTER(HPlan), C.c_uint32, C.c_uint32]
    for k, ke in [(1, 4), (3, 4), (4, 4), (5, 8), (8, 8), (9, 16), (16, 16)]:
        pl = plan(lib, 1000, 20000, 16)
        assert pl.ke == 16
        lib.nmfk_plan_set_rank(C.byref(pl), 16, k)
        assert pl.ke == ke, k
    for kt, path, k in [(16, PATH_VALU, 3), (8, PATH_AUTO, 5), (32, PATH_AUTO, 20), (64, PATH_BF16, 40)]:
        pl = plan(lib, 1000, 20000, kt, path)
        lib.nmfk_plan_set_rank(C.byref(pl), kt, k)
        assert pl.ke == kt, (kt, path)


@pytest.mark.parametrize("n,m,k,kt", [(20000, 1000, 3, 16), (5000, 38, 3, 16), (5000, 200, 8, 16), (32, 32, 1, 16),
                                     (60000, 20, 4, 4), (31, 5000, 8, 8), (3000, 24, 6, 8), (64, 20, 2, 4),
                                     (100, 100, 9, 16), (100, 100, 16, 16), (100, 100, 17, 32), (100, 100, 33, 64),
                                     (5, 3, 2, 4)])
def test_job_padded_rank(pkg, n, m, k, kt):
    # Ranks up to 8 pad to 16 (FP32 matrix cores) unless V has fewer than 32 rows or columns,
    # where the vector ALUs keep 4 or 8 columns; larger ranks pad to 16, 32 or 64.
    assert pkg.job_kt(n, m, k) == kt


@pytest.mark.parametrize("n,m,kt,ok", [(20000, 1000, 16, True), (20000, 1024, 16, True), (20000, 1025, 16, False),
                                       (20000, 1000, 32, False), (20000, 1000, 8, False), (15, 64, 16, True),
                                       (1, 1, 16, True), (50000, 2000, 64, False), (0, 10, 16, False)])
def test_one_pass_plan(lib, n, m, kt, ok):
    # k_fused_mfma takes padded rank 16 and m <= 1024 (up to 8 waves of 128 columns); its
    # workgroups split the ceil(n / 16) row blocks evenly (row blocks per workgroup differ by
    # at most one) and aim for 2048 waves in all.
    G, nwv = C.c_uint32(), C.c_uint32()
    rc = lib.nmfk_fused_plan(n, m, kt, C.byref(G), C.byref(nwv))
    assert (rc == 0) == ok
    if not ok:
        return
    nb = (n + 15) // 16
    assert nwv.value == (4 if m <= 512 else 8)
    per_wg = -(-nb // (2048 // nwv.value))          # row blocks per workgroup at the wave target
    assert G.value == -(-nb // per_wg) and 1 <= G.value <= nb
    per = [((g + 1) * nb) // G.value - (g * nb) // G.value for g in range(G.value)]
    assert min(per) >= 1 and max(per) - min(per) <= 1
    if (n, m) == (20000, 1000):
        assert (G.value, nwv.value) == (250, 8)
