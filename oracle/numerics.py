"""Oracle numerics: staggered stencils, CPML profiles, CFL, wavelets, sampling.

Test infrastructure only (see the package docstring).  Every function states
the reference lines whose arithmetic it reproduces.  Scalars are cast to the
array dtype before use, which is what NumPy 2's weak-scalar rule does for the
reference's Python-float constants, so float32 results match bit-for-bit.
"""

from __future__ import annotations

import numpy as np

# Levander / Fornberg staggered first-derivative weights, ertm/numerics.py:35-40.
# sum_k c_k (2k-1) == 1 is the consistency condition (checked in the tests).
COEFFS = {
    2: (1.0,),
    4: (9.0 / 8.0, -1.0 / 24.0),
    6: (75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0),
    8: (1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0),
}


def ghost(order: int) -> int:
    """Padding width, ertm/numerics.py:60-62: half-stencil plus one."""
    return len(COEFFS[order]) + 1


def coef_l1(order: int) -> float:
    """Sum of |c_k|, ertm/numerics.py:54-57."""
    return float(sum(abs(c) for c in COEFFS[order]))


def derivative(f: np.ndarray, axis: int, order: int, h: float, plus: bool) -> np.ndarray:
    """Staggered derivative, ertm/numerics.py:68-117.

    ``plus``: value at i+1/2, sum_k c_k (f[i+k] - f[i-k+1]) / h.
    minus:    value at i-1/2, sum_k c_k (f[i+k-1] - f[i-k]) / h.
    Only indices [K, n-K) along ``axis`` are written (K = half stencil); the
    other axis is computed in full; everything else is exactly zero.
    Accumulation starts from zero and adds term k=1..K in order, then divides.
    """
    f = np.asarray(f)
    dt = np.result_type(f.dtype, np.float32)
    out = np.zeros(f.shape, dtype=dt)
    kk = len(COEFFS[order])
    n = f.shape[axis]
    lo, hi = kk, n - kk
    if lo >= hi:
        raise ValueError("axis too short for the stencil")
    g = np.moveaxis(f, axis, 0)
    o = np.moveaxis(out, axis, 0)
    tgt = o[lo:hi]
    for k, c in enumerate(COEFFS[order], start=1):
        cc = dt.type(c)
        if plus:
            a, b = g[lo + k : hi + k], g[lo - k + 1 : hi - k + 1]
        else:
            a, b = g[lo + k - 1 : hi + k - 1], g[lo - k : hi - k]
        tgt += cc * (a - b)
    out /= dt.type(h)
    return out


def cpml_profile(n: int, h: float, npml: int, dt: float, vmax: float, f0: float,
                 lo_side: bool = True, hi_side: bool = True, enabled: bool = True):
    """CFS-PML recursion coefficients on the padded node indices.

    Follows ertm/numerics.py:158-219 with kappa_max = 1, profile order 2,
    reflection 1e-5 and alpha_max = 2*pi*max(f0, 1) (ertm/numerics.py:252).
    Returns (a, b, npml_eff) with a, b float32 of length n.  Outside the
    strips a = 0 and b = 1, so the recursion is an exact no-op there.
    """
    w = int(max(1, min(npml, n // 3)))
    sigma = np.zeros(n)
    alpha = np.zeros(n)
    depth = max(w * h, h)
    smax = -3.0 * vmax * np.log(1e-5) / (2.0 * depth)
    amax = 2.0 * np.pi * float(max(f0, 1.0))
    d = (w - np.arange(w, dtype=np.float64)) / w
    if enabled and lo_side:
        sigma[:w] = smax * d ** 2
        alpha[:w] = amax * (1.0 - d)
    if enabled and hi_side:
        sigma[n - w :] = smax * (d ** 2)[::-1]
        alpha[n - w :] = amax * (1.0 - d[::-1])
    den = sigma + alpha  # kappa == 1
    b = np.exp(-(sigma + alpha) * dt)
    a = np.zeros(n)
    ok = den > 0
    a[ok] = sigma[ok] / den[ok] * (b[ok] - 1.0)
    return a.astype(np.float32), b.astype(np.float32), w


def cfl_dt(vmax: float, dx: float, dz: float, order: int, safety: float = 0.5) -> float:
    """ertm/numerics.py:314-326: safety / (vmax * hypot(C/dx, C/dz))."""
    c = coef_l1(order)
    return float(safety / (vmax * np.hypot(c / dx, c / dz)))


def ricker(t: np.ndarray, f0: float, t0: float) -> np.ndarray:
    """ertm/numerics.py:466-469, float64 math then float32."""
    u = np.pi * float(f0) * (np.asarray(t) - float(t0))
    return ((1.0 - 2.0 * u ** 2) * np.exp(-(u ** 2))).astype(np.float32)


def gaussian(t: np.ndarray, f0: float, t0: float) -> np.ndarray:
    """ertm/numerics.py:472-475."""
    u = np.pi * float(f0) * (np.asarray(t) - float(t0))
    return np.exp(-0.5 * u ** 2).astype(np.float32)


def resample_sinc(data: np.ndarray, dt_in: float, dt_out: float, half_length: int = 8,
                  beta: float = 8.6) -> np.ndarray:
    """ertm/numerics.py:337-392 along axis 0 (1-D or 2-D): Kaiser-windowed
    sinc taps j = floor(p) - L + 1 .. floor(p) + L around each output
    position p = k * dt_out / dt_in, weights normalised to unit sum, indices
    clamped to the record, float64.  The taps are summed one at a time in
    order, starting from zero."""
    x = np.asarray(data, dtype=np.float64)
    one = x.ndim == 1
    if one:
        x = x[:, None]
    n_in = x.shape[0]
    if n_in < 3:
        raise ValueError("too few samples")
    n_out = int(np.floor((n_in - 1) * dt_in / dt_out + 1e-9)) + 1
    p = np.arange(n_out, dtype=np.float64) * (dt_out / dt_in)
    j0 = np.floor(p).astype(np.int64)
    taps = 2 * half_length
    window = np.kaiser(taps, beta)
    j = j0[:, None] + (np.arange(taps) - half_length + 1)[None, :]
    w = np.sinc(p[:, None] - j) * window[None, :]
    # NumPy's row sum (pairwise over the taps), as the reference normalises
    total = w.sum(axis=1)
    total[total == 0] = 1.0
    w = w / total[:, None]
    out = np.zeros((n_out, x.shape[1]))
    for k in range(taps):
        out = out + w[:, k, None] * x[np.clip(j[:, k], 0, n_in - 1)]
    return out[:, 0] if one else out


def bilinear(field: np.ndarray, zi: np.ndarray, xi: np.ndarray) -> np.ndarray:
    """ertm/numerics.py:478-511: float64 bilinear gather, 0 where out of range.

    Term order (v00*wz0)*wx0 + (v01*wz0)*wx1 + (v10*wz1)*wx0 + (v11*wz1)*wx1.
    """
    zi = np.asarray(zi, np.float64)
    xi = np.asarray(xi, np.float64)
    z0 = np.floor(zi).astype(np.int64)
    x0 = np.floor(xi).astype(np.int64)
    tz = zi - z0
    tx = xi - x0
    nz, nx = field.shape
    ok = (z0 >= 0) & (z0 < nz - 1) & (x0 >= 0) & (x0 < nx - 1)
    zc = np.clip(z0, 0, nz - 2)
    xc = np.clip(x0, 0, nx - 2)
    v = (
        field[zc, xc] * (1 - tz) * (1 - tx)
        + field[zc, xc + 1] * (1 - tz) * tx
        + field[zc + 1, xc] * tz * (1 - tx)
        + field[zc + 1, xc + 1] * tz * tx
    )
    return np.where(ok, v, 0.0)


def laplacian(img: np.ndarray, dx: float, dz: float) -> np.ndarray:
    """ertm/numerics.py:398-404: 5-point Laplacian with edge padding, float64."""
    a = np.asarray(img, np.float64)
    q = np.pad(a, 1, mode="edge")
    c = q[1:-1, 1:-1]
    return (q[1:-1, 2:] - 2.0 * c + q[1:-1, :-2]) / (dx * dx) + (
        q[2:, 1:-1] - 2.0 * c + q[:-2, 1:-1]
    ) / (dz * dz)


def lowpass(img: np.ndarray, dx: float, dz: float, frac: float) -> np.ndarray:
    """ertm/numerics.py:424-442: radial wavenumber ramp of width 0.15, float64."""
    a = np.asarray(img, np.float64)
    nz, nx = a.shape
    kz = np.fft.fftfreq(nz, d=dz) / max(0.5 / dz, 1e-12)
    kx = np.fft.fftfreq(nx, d=dx) / max(0.5 / dx, 1e-12)
    r = np.sqrt(kz[:, None] ** 2 + kx[None, :] ** 2)
    resp = np.clip(1.0 - (r - frac) / 0.15, 0.0, 1.0)
    return np.real(np.fft.ifft2(np.fft.fft2(a) * resp))
