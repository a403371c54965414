"""cost() container and statistics (src/ExtreM.cpp:252-389): zstd binding
against the reference's own edit-stream fixtures, the oracle's PSNR against
its closed form, known answers (CPU), and the GPU path against the oracle."""
import math
import os

import numpy as np
import pytest

from oracle import oracle
from tests.conftest import GOLDEN, make_field


def ez():
    import extrem_z_amd as m
    return m


def test_reference_edit_streams_decode():
    """src/mantissa.bin / src/exponent.bin: zstd frames of one edit stream split
    as in src/ExtreM_cuda.cu:4466-4476 (52-bit mantissa, unbiased exponent)."""
    m = np.frombuffer(ez().zstd_decompress(open(os.path.join(GOLDEN, "mantissa.zst"), "rb").read()),
                      "<u8")
    e = np.frombuffer(ez().zstd_decompress(open(os.path.join(GOLDEN, "exponent.zst"), "rb").read()),
                      "<i4")
    assert m.size == e.size == 9521
    assert int(m.max()) < 2 ** 52
    assert e.min() == -20 and e.max() == -5
    mag = (1.0 + m.astype(np.float64) / 2.0 ** 52) * np.exp2(e.astype(np.float64))
    assert 1e-6 < mag.min() and mag.max() < 0.06
    # the split is invertible: rebuild the bit patterns, split them again
    bits = (((e.astype(np.int64) + 1023) << 52) | m.astype(np.int64)).astype(np.uint64)
    assert np.array_equal(bits & np.uint64((1 << 52) - 1), m)
    assert np.array_equal(((bits >> np.uint64(52)) & np.uint64(0x7FF)).astype(np.int32) - 1023, e)


@pytest.mark.parametrize("n", [0, 1, 1000, 123457])
def test_zstd_round_trip(n):
    data = np.random.default_rng(n).integers(0, 7, n).astype(np.uint8).tobytes()
    frame = ez().zstd_compress(data, 3)
    assert ez().zstd_decompress(frame) == data


def test_oracle_psnr_closed_form():
    """A constant offset d gives mse = d^2: psnr = -20 log10(d / range)."""
    orig = np.arange(4096, dtype=np.float64) / 1024.0          # range 4095/1024
    d = 2.0 ** -10
    before = orig + d
    after = orig.copy()
    after[7] -= 2 * d
    counts = np.zeros(orig.size, np.uint8)
    counts[7] = 6
    st, bm, edits = oracle.cost(orig, before, after, counts, 0.5)
    rng = 4095.0 / 1024.0
    assert st["psnr"] == pytest.approx(-20 * math.log10(d / rng), rel=1e-14)
    want_fixed = -20 * math.log10(math.sqrt((2 * d) ** 2 / orig.size) / rng)
    assert st["fixed_psnr"] == pytest.approx(want_fixed, rel=1e-14)
    assert st["changed"] == orig.size and st["snapped"] == 1
    # the one snapped edit: -(decp0 - (orig - xi))
    assert edits.tolist() == [-(before[7] - (orig[7] - 0.5))]
    # indexs = counts where changed: one value 6 at position 7
    vals = np.zeros(orig.size, np.uint8)
    vals[7] = 6
    assert np.array_equal(bm, oracle.bitmap3(vals))
    assert math.isinf(oracle.cost(orig, orig, orig, counts * 0, 0.5)[0]["psnr"])


@pytest.mark.gpu
@pytest.mark.parametrize("shape,seed,kind,er", [((24, 20, 16), 51, "noise", 2e-2),
                                                ((40, 40, 40), 52, "gauss", 1e-3)])
def test_cost_parity(shape, seed, kind, er):
    orig, decp, xi = make_field(shape, seed, kind, er)
    fixed, counts, _ = oracle.correct(orig, decp, xi)
    want, bm, edits = oracle.cost(orig, decp, fixed, counts, xi)
    orig_b, cmp_b = orig.size * 8, 12345
    st, iz, ez_ = ez().cost(orig, decp, fixed, counts, xi, orig_b, cmp_b)
    for k in ("changed", "snapped"):
        assert st[k] == want[k], k
    assert st["edit_ratio"] == want["edit_ratio"]
    for k in ("psnr", "fixed_psnr"):
        assert st[k] == pytest.approx(want[k], rel=1e-12), k
    assert ez().zstd_decompress(iz) == bm.tobytes()
    assert ez().zstd_decompress(ez_) == edits.tobytes()
    assert st["index_bytes"] == bm.size and st["edits_bytes"] == 8 * edits.size
    ocr = orig_b / (cmp_b + len(iz) + len(ez_))
    assert st["ocr"] == pytest.approx(ocr, rel=1e-15)
    assert st["obr"] == pytest.approx(64 / ocr, rel=1e-15)
    assert st["cr"] == pytest.approx(orig_b / cmp_b, rel=1e-15)
    assert st["br"] == pytest.approx(64 / (bm.size / cmp_b), rel=1e-15)


@pytest.mark.gpu
def test_cli_writes_fixed_field_and_container(tmp_path):
    """extremz_cli with the reference's argument form (main.cpp:13-37) writes
    fixed.bin and, with --cost, the two streams of cost() and their frames."""
    import subprocess
    orig, decp, xi0 = make_field((20, 18, 12), 53, "noise", 1e-2)
    src = tmp_path / "field.bin"
    orig.astype("<f8").tofile(src)
    er = 1e-2
    xi = (orig.max() - orig.min()) * er          # getdata's bound, src/ExtreM.cpp:136-138
    from extrem_z_amd import synth
    decp = synth.quantise(orig, xi)
    dp = tmp_path / "decp.bin"
    decp.astype("<f8").tofile(dp)
    cli = os.path.join(os.path.dirname(ez().LIB_PATH), "extremz_cli")
    out = tmp_path / "fixed.bin"
    pre = str(tmp_path / "c_")
    r = subprocess.run([cli, f"{src},20,18,12", str(er), "sz3", "--decp", str(dp), "--out", str(out),
                        "--cost", pre, "--cmp-bytes", "1000"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    fixed, counts, _ = oracle.correct(orig, decp, xi)
    assert np.fromfile(out, "<f8").tobytes() == fixed.ravel().tobytes()
    _, bm, edits = oracle.cost(orig, decp, fixed, counts, xi)
    assert open(pre + "index.bin", "rb").read() == bm.tobytes()
    assert open(pre + "edits.bin", "rb").read() == edits.tobytes()
    assert ez().zstd_decompress(open(pre + "index.bin.zst", "rb").read()) == bm.tobytes()
    assert ez().zstd_decompress(open(pre + "edits.bin.zst", "rb").read()) == edits.tobytes()
    assert "OCR: " in open(pre + "stats.txt").read()


@pytest.mark.gpu
def test_cli_plane_adjacency(tmp_path):
    """--neighbors 6: the CPU path's maxNeighbors = 6 adjacency on a plane
    field, against the oracle in the same mode."""
    import subprocess
    orig, _, _ = make_field((30, 22, 1), 54, "gauss", 1e-3)
    src = tmp_path / "field.bin"
    orig.astype("<f8").tofile(src)
    er = 1e-3
    xi = (orig.max() - orig.min()) * er
    from extrem_z_amd import synth
    decp = synth.quantise(orig, xi)
    dp = tmp_path / "decp.bin"
    decp.astype("<f8").tofile(dp)
    cli = os.path.join(os.path.dirname(ez().LIB_PATH), "extremz_cli")
    out = tmp_path / "fixed.bin"
    r = subprocess.run([cli, f"{src},30,22,1", str(er), "sz3", "--decp", str(dp), "--out", str(out),
                        "--neighbors", "6"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    with oracle.plane_adjacency():
        fixed, _, _ = oracle.correct(orig, decp, xi)
    assert np.fromfile(out, "<f8").tobytes() == fixed.ravel().tobytes()
