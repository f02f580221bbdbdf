"""Matrix files (text, non-native binary, native binary) of the command-line driver.

The build's reader and writer (csrc/nmfk_matrix_io.c) are checked against the
reference's own src/matrix_io/*.c, compiled unmodified into oracle/_ref by
oracle/Makefile: the same files must load to the same matrix and tags, and the same
matrix and tags must be written to byte-identical files.  Host code only (no GPU)."""
from __future__ import annotations

import pathlib

import numpy as np
import pytest

REF_TEST = pathlib.Path("/root/reference/test")


@pytest.fixture(scope="module")
def io(pkg):
    pkg.load()
    return pkg


@pytest.fixture(scope="module")
def ref(orc):
    r = orc.ref_matrix_io()
    if r is None:
        pytest.skip("oracle/_ref/libref_matrix_io.so not built (reference tree absent)")
    return r


def rand_matrix(rows, cols, seed, integer=False):
    rng = np.random.default_rng(seed)
    A = rng.lognormal(2.0, 1.0, (rows, cols)).astype(np.float32)
    if integer:
        A = np.round(A) + 1
    return A.astype(np.float32)


def write_text(path, A, name=None, headers=None, labels=None, delim="\t", crlf=False):
    eol = "\r\n" if crlf else "\n"
    lines = []
    if headers is not None or name is not None:
        first = ([name] if name is not None else []) + (headers or [])
        lines.append(delim.join(first))
    for i, row in enumerate(A):
        vals = [f"{v:g}" for v in row]
        lines.append(delim.join(([labels[i]] if labels is not None else []) + vals))
    path.write_text(eol.join(lines) + eol)


LAYOUTS = [
    dict(),
    dict(headers=True),
    dict(labels=True),
    dict(headers=True, labels=True),
    dict(name=True, headers=True, labels=True),
    dict(name=True, headers=True, labels=True, crlf=True),
    dict(headers=True, labels=True, delim=" "),
]


@pytest.mark.parametrize("layout", range(len(LAYOUTS)))
def test_text_load_matches_reference(io, ref, tmp_path, layout):
    lay = LAYOUTS[layout]
    A = rand_matrix(17, 9, layout, integer=(layout % 2 == 0))
    hd = [f"col_{j}" for j in range(9)] if lay.get("headers") else None
    lb = [f"gene_{i}" for i in range(17)] if lay.get("labels") else None
    nm = "Description" if lay.get("name") else None
    p = tmp_path / "m.txt"
    write_text(p, A, nm, hd, lb, lay.get("delim", "\t"), lay.get("crlf", False))
    got = io.load_matrix(p)
    exp = ref[0](p, 0)
    assert exp is not None
    assert np.array_equal(got[0], exp[0])
    assert same_tags(got[1:], exp[1:])
    assert got[2] == hd and got[3] == lb and got[1] == nm


def test_reference_first_header_defect():
    """The reference reader corrupts the first column header when a description string
    precedes the headers, depending on the two strings' lengths ("Name" then "col_0"
    loads as "", "Name" then "HHHHH0" as "0", while "Description" then "col_0" is read
    correctly; seen on the reference's own test/ALL_AML_data.txt, whose first header
    ALL_19769_B-cell loads as 9769_B-cell).  The build reads the documented format
    (doc/user_guide.txt.md:186-230): the first header is the whole first field after the
    name.  Comparisons with the reference below therefore skip that one field."""


def same_tags(got, exp):
    name_g, hd_g, lb_g = got
    name_e, hd_e, lb_e = exp
    if name_g is not None and hd_g is not None:
        hd_g, hd_e = hd_g[1:], hd_e[1:]
    return (name_g, hd_g, lb_g) == (name_e, hd_e, lb_e)


def test_allaml_file_matches_reference(io, ref):
    p = REF_TEST / "ALL_AML_data.txt"
    if not p.exists():
        pytest.skip("reference test data absent")
    got, exp = io.load_matrix(p), ref[0](p, 0)
    assert got[0].shape == (5000, 38)
    assert np.array_equal(got[0], exp[0]) and same_tags(got[1:], exp[1:])
    assert got[1] == "Name" and got[2][0] == "ALL_19769_B-cell" and got[3][0] == "M12759_at"
    q = REF_TEST / "ALL_AML_data_noLabels.txt"
    if q.exists():
        g2, e2 = io.load_matrix(q), ref[0](q, 0)
        assert np.array_equal(g2[0], e2[0]) and same_tags(g2[1:], e2[1:])
        assert np.array_equal(g2[0], got[0])


@pytest.mark.parametrize("fmt", [0, 1, 2])
@pytest.mark.parametrize("tagged", [False, True])
@pytest.mark.parametrize("transpose", [False, True])
def test_writer_is_byte_identical_to_reference(io, ref, tmp_path, fmt, tagged, transpose):
    A = rand_matrix(6, 4, 3 + fmt)
    R, Cn = (4, 6) if transpose else (6, 4)
    kw = {}
    if tagged:
        kw = dict(name="matrix name", headers=[f"h{j}" for j in range(Cn)], labels=[f"l{i}" for i in range(R)])
    a, b = tmp_path / "ours", tmp_path / "ref"
    # the reference takes tags in the stored orientation and swaps them when transposing;
    # native files are written as stored (no transposition of data or tags)
    rkw = dict(kw)
    if tagged and transpose:
        rkw = dict(name=kw["name"], headers=kw["labels"], labels=kw["headers"])
    io.save_matrix(a, A, fmt, transpose, **(rkw if fmt == 2 else kw))
    assert ref[1](b, fmt, A, transpose, **rkw) == 0
    assert a.read_bytes() == b.read_bytes()


@pytest.mark.parametrize("fmt", [0, 1, 2])
def test_round_trip(io, ref, tmp_path, fmt):
    A = rand_matrix(12, 7, 9 + fmt, integer=(fmt == 0))   # %g text: exact for small integers
    kw = dict(name="n", headers=[f"c{j}" for j in range(7)], labels=[f"r{i}" for i in range(12)])
    p = tmp_path / "rt"
    io.save_matrix(p, A, fmt, **kw)
    M, name, hd, lb = io.load_matrix(p, fmt)
    assert np.array_equal(M, A) and (name, hd, lb) == (kw["name"], kw["headers"], kw["labels"])
    exp = ref[0](p, fmt)
    assert exp is not None and np.array_equal(exp[0], A)


def test_invalid_files_are_rejected(io, ref, tmp_path):
    cases = {
        "negative": "1\t2\t3\n4\t-5\t6\n",
        "zero_row": "1\t2\t3\n0\t0\t0\n4\t5\t6\n",
        "zero_col": "1\t0\t3\n4\t0\t6\n",
        "ragged": "1\t2\t3\n4\t5\n",
        "bad_char": "1\t2\t3\n4\t5,5\t6\n",
    }
    for name, text in cases.items():
        p = tmp_path / f"{name}.txt"
        p.write_text(text)
        with pytest.raises(io.NmfkError):
            io.load_matrix(p)
        # The reference reader frees a buffer twice on an all-zero column (glibc aborts
        # the process), so that case is only checked on the build's reader.
        if name != "zero_col":
            assert ref[0](p, 0) is None, name
    p = tmp_path / "badsig.dat"
    p.write_bytes(b"\x00\x01\x02\x03" + b"\x00" * 40)
    with pytest.raises(io.NmfkError):
        io.load_matrix(p, 1)
    assert ref[0](p, 1) is None
