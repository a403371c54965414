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
              "extremz_last_error"):
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
