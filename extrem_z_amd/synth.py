"""Synthetic correction workloads: original field, decompressed field, bound.

Fields are float32 on disk and promoted exactly to float64 (the reference
computes in float64, src/ExtreM.cpp:33).  The error bound follows getdata,
src/ExtreM.cpp:136-138: xi = er * (max - min).  The decompressed field is a
uniform quantiser, decp = 2*xi*rint(orig / (2*xi)), which stands in for the
reference's SZ3 round trip (src/ExtreM.cpp:121-131) and creates plateaus and
false critical points.  Layout is x fastest: idx = x + W*(y + H*z).
"""
from __future__ import annotations

import numpy as np


def gaussians(shape, n_gauss, seed, noise=1e-3):
    """Sum of anisotropic Gaussians plus uniform noise of noise * range.

    shape is (W, H, D).  Amplitudes ~ U[-1, 1], centres ~ U[0, n) per axis,
    sigmas ~ U[4, 16] voxels per axis.  Returns float32 with shape (D, H, W).
    """
    W, H, D = shape
    rng = np.random.default_rng(seed)
    f = np.zeros((D, H, W), np.float64)
    axes = [np.arange(n, dtype=np.float64) for n in (W, H, D)]
    for _ in range(n_gauss):
        a = rng.uniform(-1.0, 1.0)
        c = [rng.uniform(0, n) for n in (W, H, D)]
        s = [rng.uniform(4.0, 16.0) for _ in range(3)]
        gx, gy, gz = (np.exp(-0.5 * ((axes[k] - c[k]) / s[k]) ** 2) for k in range(3))
        f += a * (gz[:, None, None] * gy[None, :, None] * gx[None, None, :])
    rng_span = float(f.max() - f.min()) or 1.0
    f += rng.uniform(-1.0, 1.0, size=f.shape) * (noise * rng_span)
    return f.astype(np.float32)


def grf(shape, seed, slope=-3.0):
    """Gaussian random field with power spectrum |k|^slope, normalised to [0, 1].

    White N(0,1) -> rFFT -> multiply by |k|^(slope/2) with DC = 0 -> irFFT.
    Returns float32 with shape (D, H, W).
    """
    import scipy.fft as sfft

    W, H, D = shape
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((D, H, W), dtype=np.float32)
    spec = sfft.rfftn(white, workers=16)
    del white
    kz = np.fft.fftfreq(D).astype(np.float32)[:, None, None]
    ky = np.fft.fftfreq(H).astype(np.float32)[None, :, None]
    kx = np.fft.rfftfreq(W).astype(np.float32)[None, None, :]
    k2 = kz * kz + ky * ky + kx * kx
    k2[0, 0, 0] = 1.0
    amp = k2 ** np.float32(slope / 4.0)
    amp[0, 0, 0] = 0.0
    spec *= amp
    del amp, k2
    f = sfft.irfftn(spec, s=(D, H, W), workers=16).astype(np.float32)
    del spec
    lo, hi = float(f.min()), float(f.max())
    f -= lo
    f /= np.float32(hi - lo if hi > lo else 1.0)
    return f


def bound_from_er(orig, er):
    """xi = er * (max - min) over the float64 promotion, src/ExtreM.cpp:136-138."""
    o = np.asarray(orig, dtype=np.float64)
    return float(er * (o.max() - o.min()))


def quantise(orig, xi):
    """decp = 2*xi*rint(orig/(2*xi)) in float64, stepped so |decp-orig| <= xi."""
    o = np.asarray(orig, dtype=np.float64)
    q = 2.0 * xi
    decp = q * np.rint(o / q)
    over = np.abs(decp - o) > xi
    if over.any():
        decp[over] -= np.sign(decp[over] - o[over]) * q
    # a value half-way between two grid points can miss the bound by one
    # rounding on both sides: such voxels keep their original value
    still = np.abs(decp - o) > xi
    if still.any():
        decp[still] = o[still]
    assert np.all(np.abs(decp - o) <= xi)
    return decp


def make_case(kind, shape, seed, er=None, xi=None, n_gauss=8):
    """Returns (orig float64 (D,H,W), decp float64, xi)."""
    if kind == "gauss":
        f32 = gaussians(shape, n_gauss, seed)
    elif kind == "grf":
        f32 = grf(shape, seed)
    else:
        raise ValueError(kind)
    orig = f32.astype(np.float64)
    if xi is None:
        xi = bound_from_er(orig, er)
    return orig, quantise(orig, xi), float(xi)


def config(n):
    """The five workload definitions (index 1-based as in their descriptions)."""
    if n == 1:
        return make_case("gauss", (64, 64, 64), seed=1, er=1e-3, n_gauss=8)
    if n == 2:
        return make_case("gauss", (256, 256, 256), seed=2, er=1e-4, n_gauss=32)
    if n == 3:
        return make_case("grf", (512, 512, 512), seed=3, xi=1e-3)
    raise ValueError(n)


def mixed_batch(count=64, seed=4, lo=(32, 32, 32), hi=(384, 384, 256)):
    """Shapes log-uniform per axis; er from {1e-2, 1e-3, 1e-4}; built as config 2."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        shp = tuple(int(round(np.exp(rng.uniform(np.log(a), np.log(b))))) for a, b in zip(lo, hi))
        er = float(rng.choice([1e-2, 1e-3, 1e-4]))
        ng = max(1, int(round(32 * np.prod(shp) / 256 ** 3))) if np.prod(shp) > 64 ** 3 else 8
        out.append(make_case("gauss", shp, seed=4000 + i, er=er, n_gauss=ng))
    return out
