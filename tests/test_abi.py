"""The C-ABI library loads and exports every symbol include/extremz.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "extremz.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(extremz_[a-z_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("extremz_correct", "extremz_correct_device", "extremz_classify", "extremz_labels",
              "extremz_lut", "extremz_encode", "extremz_ctx_create", "extremz_ctx_destroy",
              "extremz_last_error", "extremz_pool_create", "extremz_pool_destroy",
              "extremz_batch_correct", "extremz_serve_submit", "extremz_serve_wait",
              "extremz_serve_wait_any", "extremz_pool_info_get"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import extrem_z_amd as ez
    L = ctypes.CDLL(ez.LIB_PATH)
    for n in declared():
        assert hasattr(L, n), n


def test_invalid_arguments_are_rejected_without_gpu():
    import numpy as np
    import extrem_z_amd as ez
    o = np.zeros((2, 2, 2))
    with pytest.raises(ez.ExtremzError) as e:
        ez.correct(o, o, 0.0)      # xi must be > 0
    assert e.value.code in (-22, -19)
    with pytest.raises(ValueError):
        ez.correct(o, np.zeros((2, 2, 3)), 0.1)


def test_lut_needs_no_device(golden):
    import numpy as np
    import extrem_z_amd as ez
    ref = np.fromfile(os.path.join(golden, "LUT.bin"), dtype="<i4")
    assert np.array_equal(ez.lut(), ref)


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of extremz_job / extremz_stats / extremz_pool_info
    have the header's sizes and field offsets (compiled with the host C
    compiler)."""
    import subprocess
    import extrem_z_amd as ez
    fields = {"extremz_job": (ez.Job, [f for f, _ in ez.Job._fields_]),
              "extremz_stats": (ez.Stats, [f for f, _ in ez.Stats._fields_]),
              "extremz_pool_info": (ez.PoolInfo, [f for f, _ in ez.PoolInfo._fields_]),
              "extremz_opts": (ez.Opts, [f for f, _ in ez.Opts._fields_])}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "extremz.h"',
             "int main(void) {"]
    for name, (_, fl) in fields.items():
        lines.append(f'printf("%zu\\n", sizeof({name}));')
        for f in fl:
            lines.append(f'printf("%zu\\n", offsetof({name}, {f}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = []
    for name, (cls, fl) in fields.items():
        want.append(ctypes.sizeof(cls))
        want += [getattr(cls, f).offset for f in fl]
    assert got == want


def test_pool_without_device_reports_enodev():
    import extrem_z_amd as ez
    with pytest.raises(ez.ExtremzError) as e:
        ez.Pool(0, lanes=0)
    assert e.value.code == -22
    try:
        p = ez.Pool(0, lanes=2)
    except ez.ExtremzError as e:
        assert e.code == -19
        return
    p.close()
    pytest.skip("a gfx950 device is present")
