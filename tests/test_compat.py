"""libExtreMz: the reference library's API (include/api/ExtreM.h) over the GPU
engine.

CPU: the library loads and defines every function and global the header
declares, the header declares what the reference's header declares (when the
reference tree is at hand), and the test programs are built.

GPU: the correction driven through the API in the order of the reference's
main.cpp (tests/compat/dropin_main.cpp), and the reference's main.cpp itself
compiled against this header and library (oracle/_ref/main_dropin), write a
fixed.bin byte-identical to extremz_correct's on the same field, with the
same number of outer iterations; main.cpp's split-tree report agrees with the
test program's.
"""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT, make_field

HEADER = os.path.join(ROOT, "include", "api", "ExtreM.h")
LIB = os.path.join(ROOT, "extrem_z_amd", "_build", "libExtreMz.so")
DRIVER = os.path.join(ROOT, "tests", "compat", "_build", "dropin_main")
REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "main_dropin")
REF_HEADER = "/root/reference/include/api/ExtreM.h"


def declarations(path):
    """(functions, globals) declared in an ExtreM.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    body = text[text.index('extern "C"'):]
    funcs = set(re.findall(r"\b([A-Za-z_]\w*)\s*\(", body)) - {"if", "for", "while"}
    globs = set()
    for decl in re.findall(r"extern\s+([^;]*);", body):
        decl = re.sub(r"<[^;]*>", "", decl)  # template arguments
        for part in decl.split(","):
            m = re.search(r"([A-Za-z_]\w*)\s*(\[[^\]]*\])*\s*$", part.strip())
            if m:
                globs.add(m.group(1))
    types = {"int", "double", "atomic_int", "vector", "unordered_map"}
    return funcs, globs - types


def test_header_matches_the_reference_header():
    if not os.path.exists(REF_HEADER):
        pytest.skip("reference tree not present")
    ours, theirs = declarations(HEADER), declarations(REF_HEADER)
    assert ours[0] == theirs[0]
    assert ours[1] == theirs[1]
    assert re.search(r"extern int directions\[42\]", open(HEADER).read())


def test_library_defines_every_declared_symbol():
    assert os.path.exists(LIB), "libExtreMz.so not built (python __graft_entry__.py build)"
    funcs, globs = declarations(HEADER)
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    defined = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = sorted(n for n in funcs | globs if n not in defined)
    assert not missing, missing
    # the static archive of the reference's CMake target carries the same object
    assert os.path.exists(os.path.join(ROOT, "extrem_z_amd", "_build", "libExtreMz.a"))
    assert os.path.exists(DRIVER)


def test_cmake_package_links_the_reference_target(tmp_path):
    """find_package(ExtreMz) gives the reference's exported target
    ExtreMz::ExtreMz (a static library); a program against the API links."""
    import shutil
    if not shutil.which("cmake"):
        pytest.skip("cmake not available")
    (tmp_path / "CMakeLists.txt").write_text(
        "cmake_minimum_required(VERSION 3.20)\nproject(app CXX)\n"
        f"find_package(ExtreMz 1.0 REQUIRED PATHS {ROOT}/extrem_z_amd/cmake NO_DEFAULT_PATH)\n"
        f"add_executable(app {ROOT}/tests/compat/dropin_main.cpp)\n"
        "target_link_libraries(app PRIVATE ExtreMz::ExtreMz)\n")
    b = tmp_path / "b"
    for cmd in (["cmake", "-S", str(tmp_path), "-B", str(b)], ["cmake", "--build", str(b)]):
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert (b / "app").exists()


def write_case(tmp_path, shape, seed, kind, er, n_gauss=4):
    o, d, xi = make_field(shape, seed, kind, er, n_gauss)
    f = tmp_path / "field.f64"
    dp = tmp_path / "decp.f64"
    o.astype("<f8").tofile(f)
    d.astype("<f8").tofile(dp)
    return o, d, xi, str(f), str(dp)


def run(cmd, decp_path, cwd):
    env = dict(os.environ, EXTREMZ_DECP=decp_path, OMP_NUM_THREADS="4")
    print("run:", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, cwd=cwd, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


CASES = [
    # shape, seed, kind, er, maxNeighbors
    ((24, 20, 16), 7, "gauss", 1e-3, 14),
    ((24, 20, 16), 7, "gauss", 1e-3, 6),   # main.cpp's maxNeighbors on a 3-D field
    ((16, 16, 16), 11, "noise", 1e-2, 14),
    ((40, 30, 1), 5, "gauss", 1e-3, 6),    # the CPU path's plane adjacency
    ((33, 17, 9), 21, "grf", 1e-3, 14),
]


@pytest.mark.gpu
@pytest.mark.parametrize("shape,seed,kind,er,M", CASES)
def test_dropin_sequence_equals_extremz_correct(tmp_path, shape, seed, kind, er, M):
    import extrem_z_amd as ez
    o, d, xi, f, dp = write_case(tmp_path, shape, seed, kind, er)
    W, H, D = shape
    out = tmp_path / "fixed.bin"
    text = run([DRIVER, f, str(W), str(H), str(D), repr(er), str(M), str(out)], dp, tmp_path)
    m = re.search(r"outer (\d+) changed (\d+) branches (\d+)/(\d+) unmatched (\d+)", text)
    assert m, text
    adj = ez.ADJ_PLANE4 if (M == 6 and D == 1) else ez.ADJ_STENCIL14
    fixed, counts, st = ez.correct(o, d, xi, adjacency=adj)
    got = np.fromfile(out, dtype="<f8").reshape(o.shape)
    assert got.tobytes() == fixed.tobytes()
    assert int(m.group(1)) == st["outer_iters"]
    assert int(m.group(2)) == int(np.count_nonzero(fixed != d))
    # every maximum of the original keeps its branch (main.cpp:480-504 then
    # compares the branches' vertex lists position by position)
    assert int(m.group(3)) == int(m.group(4)) > 0


@pytest.mark.gpu
@pytest.mark.parametrize("shape,seed,kind,er", [((24, 20, 16), 7, "gauss", 1e-3),
                                                ((40, 30, 1), 5, "gauss", 1e-3)])
def test_reference_main_runs_on_libextremz(tmp_path, shape, seed, kind, er):
    """The reference's own main.cpp, compiled unchanged against
    include/api/ExtreM.h and linked with libExtreMz (main.cpp sets
    maxNeighbors = 6: the plane adjacency on the depth-1 field, the stencil
    on the 3-D one)."""
    if not os.path.exists(REF_MAIN):
        pytest.skip("oracle/_ref/main_dropin not built (needs the reference tree at build time)")
    import extrem_z_amd as ez
    o, d, xi, f, dp = write_case(tmp_path, shape, seed, kind, er)
    W, H, D = shape
    text = run([REF_MAIN, f"{f},{W},{H},{D}", repr(er), "sz3"], dp, tmp_path)
    adj = ez.ADJ_PLANE4 if D == 1 else ez.ADJ_STENCIL14
    fixed, counts, st = ez.correct(o, d, xi, adjacency=adj)
    got = np.fromfile(tmp_path / "fixed.bin", dtype="<f8").reshape(o.shape)
    assert got.tobytes() == fixed.tobytes()
    assert text.count("whole loops:") == st["outer_iters"]
    # main.cpp:499 prints one line per unmatched branch position: as many as
    # the test program counts on the same field
    out = tmp_path / "fixed2.bin"
    mine = run([DRIVER, f, str(W), str(H), str(D), repr(er), "6", str(out)], dp, tmp_path)
    m = re.search(r"unmatched (\d+)", mine)
    assert text.count("unmatched at vertex") == int(m.group(1))
