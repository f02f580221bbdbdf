"""The C-ABI boundary, without a GPU: the library loads, exports every function that
include/*.h declares, carries no oracle code, and its host-side drop-in functions
(get_difference, check_arguments, get_padding) behave like the reference's own src/common.c
built unmodified under oracle/_ref."""
from __future__ import annotations

import ctypes as C
import pathlib
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT

INCLUDE = ROOT / "include"


def declared_functions():
    names = set()
    for h in sorted(INCLUDE.glob("*.h")):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(([^;{}]*)\)\s*;", text):
            name = m.group(1)
            if name in {"sizeof", "if", "while", "for", "return"}:
                continue
            names.add(name)
    return names


def exported_symbols(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib)], capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.split()[-2:-1] in (["T"], ["W"])}


def test_every_declared_function_is_exported(pkg):
    pkg.load()
    decl = declared_functions()
    assert len(decl) > 60, sorted(decl)
    missing = decl - exported_symbols(pkg.LIB_PATH)
    assert not missing, sorted(missing)


def test_library_has_no_oracle_code(pkg):
    syms = exported_symbols(pkg.LIB_PATH)
    assert not any(s.startswith("orc_") for s in syms)
    ldd = subprocess.run(["ldd", str(pkg.LIB_PATH)], capture_output=True, text=True).stdout
    assert "oracle" not in ldd and "ref_common" not in ldd


def test_headers_compile_as_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text("".join(f'#include "{h.name}"\n' for h in sorted(INCLUDE.glob("*.h"))) + "int main(void){return 0;}\n")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    f"-I{INCLUDE}", "-c", str(src), "-o", str(tmp_path / "t.o")], check=True)


class InputArgs(C.Structure):
    # struct input_arguments (include/common.h of the reference and of this library)
    _fields_ = [("filename", C.c_char_p), ("numeric_hdrs", C.c_bool), ("numeric_lbls", C.c_bool),
                ("input_file_fmt", C.c_int), ("output_file_fmt", C.c_int), ("k", C.c_uint32),
                ("kp", C.c_uint32), ("nIters", C.c_uint32), ("niter_test_conv", C.c_uint32),
                ("stop_threshold", C.c_uint32), ("gpu_device", C.c_uint32), ("idx_other_args", C.c_uint32)]


def _bind_common(lib):
    lib.get_difference.restype = C.c_size_t
    lib.get_difference.argtypes = [np.ctypeslib.ndpointer(np.uint32), np.ctypeslib.ndpointer(np.uint32), C.c_uint32]
    lib.check_arguments.restype = C.c_int
    lib.check_arguments.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_bool), C.POINTER(InputArgs)]
    lib.get_padding.restype = C.c_uint32
    lib.get_padding.argtypes = [C.c_uint32]
    lib.set_default_matrix_limits.restype = C.c_int
    return lib


@pytest.fixture(scope="module")
def libs(pkg, orc):
    prod = _bind_common(pkg.load())
    ref = orc.ref_common()
    if ref is None:
        pytest.skip("oracle/_ref/libref_common.so not built (reference tree absent)")
    return prod, _bind_common(ref)


def test_get_difference_matches_reference(libs, orc):
    prod, ref = libs
    rng = np.random.default_rng(7)
    for m in (1, 2, 3, 38, 500, 3000):
        for labels in (2, 9, 64, 300):
            c = rng.integers(0, labels, m).astype(np.uint32)
            l = rng.integers(0, labels, m).astype(np.uint32)
            assert prod.get_difference(c, l, m) == ref.get_difference(c, l, m) == orc.get_difference(c, l)


ARGV_CASES = [
    ["NMF_GPU", "data.txt"],
    ["NMF_GPU", "data.txt", "-k", "7", "-i", "500", "-j", "5", "-t", "12", "-z", "1"],
    ["NMF_GPU", "-b", "1", "-e", "0", "-c", "-r", "data.dat", "extra"],
    ["NMF_GPU", "-K", "33", "-I", "9", "-J", "3", "-T", "4", "f"],
    ["NMF_GPU", "-k", "1", "f"],        # rank below 2: error
    ["NMF_GPU", "-i", "0", "f"],        # zero iterations: error
    ["NMF_GPU", "-j", "x", "f"],        # not a number: error
    ["NMF_GPU", "-k"],                  # missing argument: error
    ["NMF_GPU", "-k", "4"],             # no file name: error
    ["NMF_GPU", "-h", "f"],             # help
    ["NMF_GPU", "-q", "f"],             # unknown option ignored
]


@pytest.mark.parametrize("argv", ARGV_CASES)
def test_check_arguments_matches_reference(libs, argv):
    outs = []
    for lib in libs:
        lib.set_default_matrix_limits()
        arr = (C.c_char_p * (len(argv) + 1))(*[a.encode() for a in argv], None)
        hv = C.c_bool(False)
        a = InputArgs()
        rc = lib.check_arguments(len(argv), arr, C.byref(hv), C.byref(a))
        fields = None
        if rc == 0 and not hv.value:
            fields = tuple((f, getattr(a, f)) for f, _ in InputArgs._fields_)
        outs.append((rc, hv.value, fields))
    assert outs[0] == outs[1]


def test_get_padding_matches_reference(libs):
    prod, ref = libs
    for lib in libs:
        lib.set_default_matrix_limits()
    for d in (1, 2, 15, 16, 17, 38, 1000, 4097):
        assert prod.get_padding(d) == ref.get_padding(d)


def test_exports_cover_the_reference_driver(pkg):
    """Every function and global the reference's src/NMF_GPU.c uses from the headers this
    library replaces (NMF_routines.h, matrix_operations.h, GPU_setup.h, common.h) is
    exported, so that driver links against libnmfgpu_amd.so (its matrix I/O comes from the
    reference's own ISO-C sources)."""
    import re
    import subprocess
    ref = pathlib.Path("/root/reference")
    if not (ref / "src" / "NMF_GPU.c").exists():
        pytest.skip("reference tree absent")

    def strip(t):
        return re.sub(r"//[^\n]*", "", re.sub(r"/\*.*?\*/", "", t, flags=re.S))
    src = strip((ref / "src" / "NMF_GPU.c").read_text(errors="replace"))
    hdrs = strip("".join((ref / "include" / h).read_text(errors="replace")
                         for h in ("NMF_routines.h", "matrix_operations.h", "GPU_setup.h", "common.h")))
    keywords = {"if", "for", "while", "switch", "return", "sizeof"}
    declared = set(re.findall(r"\b([A-Za-z_]\w*)\s*\(", hdrs)) - keywords
    used = {c for c in re.findall(r"\b([A-Za-z_]\w*)\s*\(", src) if c in declared and not c.isupper()}
    glob = set(re.findall(r"\bextern\s+[^;(]*?\b([A-Za-z_]\w*)\s*(?:\[[^\]]*\])?\s*;", hdrs))
    used |= {g for g in glob if re.search(r"\b" + g + r"\b", src)}
    out = subprocess.run(["nm", "-D", "--defined-only", str(pkg.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert len(used) > 20
    assert sorted(used - syms) == []


def _integration_blocks():
    text = (ROOT / "INTEGRATION.md").read_text()
    sec = text[text.index("## 1. Rebuilding"):text.index("## 1a.")]
    diff = re.search(r"```diff\n(.*?)```", sec, re.S).group(1)
    make = re.search(r"```make\n(.*?)```", sec, re.S).group(1)
    return diff, make


def test_integration_recipe_builds_reference_driver(pkg, tmp_path):
    """The integration guide's rebuild recipe, word for word: its one-line diff applied to the
    reference's src/NMF_GPU.c, and its make rule run from a tree laid out like the reference's (the
    matrix I/O sources and headers are the reference's own, reached through symlinks; only
    the patched driver file is a temporary copy).  It must compile without implicit
    declarations and link against libnmfgpu_amd.so; the built driver's -h runs on the host."""
    ref = pathlib.Path("/root/reference")
    if not (ref / "src" / "NMF_GPU.c").exists():
        pytest.skip("reference tree absent")
    pkg.load()
    diff, make = _integration_blocks()
    hunk = re.search(r"@@ -(\d+) \+(\d+) @@\n-(.*)\n\+(.*)\n", diff)
    line_no, old, new = int(hunk.group(1)), hunk.group(3), hunk.group(4)
    lines = (ref / "src" / "NMF_GPU.c").read_text(errors="replace").split("\n")
    assert lines[line_no - 1] == old, (line_no, lines[line_no - 1])
    lines[line_no - 1] = new
    (tmp_path / "src").mkdir()
    (tmp_path / "src" / "NMF_GPU.c").write_text("\n".join(lines))
    (tmp_path / "src" / "matrix_io").symlink_to(ref / "src" / "matrix_io")
    (tmp_path / "include").symlink_to(ref / "include")
    (tmp_path / "Makefile").write_text(make)
    r = subprocess.run(["make", f"NMFGPU_AMD={pkg.LIB_PATH.parent}", "bin/NMF_GPU"], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert "implicit declaration" not in r.stderr
    drv = tmp_path / "bin" / "NMF_GPU"
    ldd = subprocess.run(["ldd", str(drv)], capture_output=True, text=True).stdout
    assert str(pkg.LIB_PATH) in ldd
    h = subprocess.run([str(drv), "-h"], capture_output=True, text=True, timeout=60)
    assert h.returncode == 0 and "Usage:" in h.stdout + h.stderr


def test_drop_in_headers_define_reference_macros(tmp_path):
    # The macros the reference's host sources take from real_type.h / index_type.h, with the
    # reference's release-build meanings (include/real_type.h:71,133-135,169-174,
    # include/index_type.h:121-123,160-163).
    src = tmp_path / "m.c"
    src.write_text(r'''
#include <stdio.h>
#include "real_type.h"
#include "index_type.h"
int main(void) {
    real x = 0; index_t i = 0; char buf[8] = "2.5 7";
    if (sscanf(buf, "%" SCNgREAL " %" SCN_IDX, &x, &i) != 2) return 1;
    char *end = NULL;
    real y = STRTOREAL("0.25", &end);
    printf("%g %" PRI_IDX " %g %g %d %d %d %d\n", (double)x, i, (double)y, (double)SQRTR(REAL_C(16.0)),
           MAX(3, 4), MIN(3, 4), MAX3(1, 5, 2), MIN3(4, 2, 3));
    return (sizeof(real) == 4 && sizeof(index_t) == 4 && isnan(NAN) && R_MIN == FLT_MIN) ? 0 : 2;
}
''')
    exe = tmp_path / "m"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{INCLUDE}", str(src), "-o", str(exe), "-lm"],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and out.stdout.split() == ["2.5", "7", "0.25", "4", "4", "3", "5", "2"]
