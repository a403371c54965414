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


def _zchunks(D, parts=16):
    step = max(1, -(-D // parts))
    return [(z, min(D, z + step)) for z in range(0, D, step)]


def _threads():
    import os
    return max(1, min(16, os.cpu_count() or 1))


def gaussians(shape, n_gauss, seed, noise=1e-3):
    """Sum of anisotropic Gaussians plus uniform noise of noise * range.

    shape is (W, H, D).  Amplitudes ~ U[-1, 1], centres ~ U[0, n) per axis,
    sigmas ~ U[4, 16] voxels per axis.  Returns float32 with shape (D, H, W).
    The sum is accumulated Gaussian by Gaussian in float64 over z-slabs in
    parallel threads; every voxel sees the same operations in the same order
    as a whole-array loop, so the result does not depend on the thread count.
    """
    from concurrent.futures import ThreadPoolExecutor

    W, H, D = shape
    rng = np.random.default_rng(seed)
    axes = [np.arange(n, dtype=np.float64) for n in (W, H, D)]
    terms = []
    for _ in range(n_gauss):
        a = rng.uniform(-1.0, 1.0)
        c = [rng.uniform(0, n) for n in (W, H, D)]
        s = [rng.uniform(4.0, 16.0) for _ in range(3)]
        gx, gy, gz = (np.exp(-0.5 * ((axes[k] - c[k]) / s[k]) ** 2) for k in range(3))
        terms.append((a, gx, gy, gz))
    f = np.empty((D, H, W), np.float64)
    chunks = _zchunks(D)

    def slab(z0, z1):
        fz = f[z0:z1]
        fz[...] = 0.0
        for a, gx, gy, gz in terms:
            fz += a * (gz[z0:z1, None, None] * gy[None, :, None] * gx[None, None, :])
        return fz.min(), fz.max()

    with ThreadPoolExecutor(_threads()) as ex:
        ext = list(ex.map(lambda c: slab(*c), chunks))
        rng_span = float(max(e[1] for e in ext) - min(e[0] for e in ext)) or 1.0
        u = rng.uniform(-1.0, 1.0, size=f.shape)
        out = np.empty((D, H, W), np.float32)
        amp = noise * rng_span

        def add(z0, z1):
            fz = f[z0:z1]
            fz += u[z0:z1] * amp
            out[z0:z1] = fz

        list(ex.map(lambda c: add(*c), chunks))
    return out


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
    """decp = 2*xi*rint(orig/(2*xi)) in float64, stepped so |decp-orig| <= xi.

    Elementwise, evaluated over z-slabs in parallel threads."""
    from concurrent.futures import ThreadPoolExecutor

    o = np.asarray(orig, dtype=np.float64)
    q = 2.0 * xi
    decp = np.empty_like(o)
    o3 = o.reshape(-1, *o.shape[-2:]) if o.ndim >= 2 else o.reshape(1, 1, -1)
    d3 = decp.reshape(o3.shape)

    def slab(z0, z1):
        ov, dv = o3[z0:z1], d3[z0:z1]
        dv[...] = q * np.rint(ov / q)
        over = np.abs(dv - ov) > xi
        if over.any():
            dv[over] -= np.sign(dv[over] - ov[over]) * q
        # a value half-way between two grid points can miss the bound by one
        # rounding on both sides: such voxels keep their original value
        still = np.abs(dv - ov) > xi
        if still.any():
            dv[still] = ov[still]
        return bool(np.all(np.abs(dv - ov) <= xi))

    with ThreadPoolExecutor(_threads()) as ex:
        ok = list(ex.map(lambda c: slab(*c), _zchunks(o3.shape[0])))
    assert all(ok)
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


def serve_set(count=16, seed=5, lo=16, hi=512):
    """Serving jobs: cubic edges log-uniform in [lo, hi]; half sums of
    Gaussians (Gaussian density of the 256^3 case), half k^-3 Gaussian random
    fields; er from {1e-2, 1e-3, 1e-4}.  Returns [(orig, decp, xi, kind)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        er = float(rng.choice([1e-2, 1e-3, 1e-4]))
        kind = "grf" if i % 2 else "gauss"
        shp = (n, n, n)
        if kind == "gauss":
            ng = max(1, int(round(32 * n ** 3 / 256 ** 3))) if n ** 3 > 64 ** 3 else 8
            f32 = gaussians(shp, ng, seed=5000 + i)
        else:
            f32 = grf(shp, seed=5000 + i)
        orig = f32.astype(np.float64)
        xi = bound_from_er(orig, er)
        out.append((orig, quantise(orig, xi), xi, kind))
    return out
