"""Numerical building blocks (ertm/numerics.py).

Host-side set-up math lives here: stencil weights, CFS-PML profiles, the CFL
bound, wavelets, Kaiser-sinc resampling and the per-shot image filters
(Laplacian, wavenumber low-pass).  The staggered derivative itself --
``dplus`` / ``dminus`` -- runs on the GPU through the same stencil code the
propagators use; there is no host implementation of it.
"""

from __future__ import annotations

from typing import Iterable

import numpy as np

# Staggered first-derivative weights (Levander 1988), ertm/numerics.py:35-40.
STAGGERED_COEFFICIENTS: dict[int, tuple[float, ...]] = {
    2: (1.0,),
    4: (9.0 / 8.0, -1.0 / 24.0),
    6: (75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0),
    8: (1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0),
}
SUPPORTED_ORDERS = tuple(sorted(STAGGERED_COEFFICIENTS))


def staggered_coefficients(order: int) -> tuple[float, ...]:
    try:
        return STAGGERED_COEFFICIENTS[int(order)]
    except (KeyError, TypeError) as exc:
        raise ValueError(f"unsupported stencil order {order!r}; choose from {SUPPORTED_ORDERS}") from exc


def coefficient_l1(order: int) -> float:
    """sum |c_k| (CFL constant), ertm/numerics.py:54-57."""
    return float(sum(abs(c) for c in staggered_coefficients(order)))


def ghost_width(order: int) -> int:
    """Array padding: half-stencil + 1, ertm/numerics.py:60-62."""
    return len(staggered_coefficients(order)) + 1


def coef_f32(order: int) -> tuple[float, float, float, float]:
    """Stencil weights as the float32 values the kernels use, zero-filled to 4."""
    cs = [float(np.float32(c)) for c in staggered_coefficients(order)]
    return tuple(cs + [0.0] * (4 - len(cs)))  # type: ignore[return-value]


# ---------------------------------------------------------------------------
# derivative on the GPU
# ---------------------------------------------------------------------------
def _derivative(field: np.ndarray, axis: int, coeffs, spacing: float, plus: bool) -> np.ndarray:
    from . import device

    array = np.asarray(field)
    if array.ndim != 2:
        raise ValueError(f"derivative needs a 2-D array, got {array.ndim}-D")
    if axis not in (0, 1):
        raise ValueError(f"axis must be 0 or 1, got {axis}")
    if spacing <= 0:
        raise ValueError(f"grid spacing must be positive, got {spacing}")
    order = 2 * len(tuple(coeffs))
    if order not in STAGGERED_COEFFICIENTS or tuple(coeffs) != STAGGERED_COEFFICIENTS[order]:
        raise ValueError("only the standard staggered weights are supported")
    if array.shape[axis] <= order:
        raise ValueError(f"axis {axis} of length {array.shape[axis]} is too short for order {order}")
    return device.derivative(array, axis, order, float(spacing), plus)


def dplus(field: np.ndarray, axis: int, coeffs: Iterable[float], spacing: float) -> np.ndarray:
    """d/dx at i+1/2 (ertm/numerics.py:120-127), computed by the HIP kernel."""
    return _derivative(field, axis, tuple(coeffs), spacing, True)


def dminus(field: np.ndarray, axis: int, coeffs: Iterable[float], spacing: float) -> np.ndarray:
    """d/dx at i-1/2 (ertm/numerics.py:130-137), computed by the HIP kernel."""
    return _derivative(field, axis, tuple(coeffs), spacing, False)


# ---------------------------------------------------------------------------
# CFS-PML
# ---------------------------------------------------------------------------
def cpml_profile(length: int, spacing: float, npml: int, dt: float, vmax: float, f0: float,
                 lo_side: bool = True, hi_side: bool = True, enabled: bool = True):
    """Recursion coefficients (a, b) per padded node and the strip width.

    CFS-PML of Komatitsch & Martin (2007) with kappa = 1, quadratic damping
    profile, R = 1e-5 and alpha_max = 2*pi*max(f0, 1) -- the parameters of
    ertm/numerics.py:158-219 (f0 is fixed at 25 Hz by the propagators).  The
    profile is evaluated in float64 and stored in float32.  Outside the strips
    a = 0 and b = 1, which makes the memory recursion an exact no-op; the
    engine therefore stores memories only inside the strips.
    """
    n = int(length)
    width = int(max(1, min(int(npml), n // 3)))
    depth = max(width * spacing, spacing)
    sig_max = -3.0 * vmax * np.log(1e-5) / (2.0 * depth)
    a_max = 2.0 * np.pi * float(max(f0, 1.0))
    ramp = (width - np.arange(width, dtype=np.float64)) / width
    sigma = np.zeros(n)
    alpha = np.zeros(n)
    if enabled and lo_side:
        sigma[:width] = sig_max * ramp ** 2
        alpha[:width] = a_max * (1.0 - ramp)
    if enabled and hi_side:
        sigma[n - width :] = sig_max * (ramp ** 2)[::-1]
        alpha[n - width :] = a_max * (1.0 - ramp[::-1])
    den = sigma + alpha
    b = np.exp(-(sigma + alpha) * dt)
    a = np.zeros(n)
    nz = den > 0
    a[nz] = sigma[nz] / den[nz] * (b[nz] - 1.0)
    return a.astype(np.float32), b.astype(np.float32), width


# ---------------------------------------------------------------------------
# stability / sampling
# ---------------------------------------------------------------------------
def cfl_dt(vmax: float, dx: float, dz: float, order: int, safety: float = 0.5) -> float:
    """Largest stable dt: safety / (vmax * hypot(C/dx, C/dz)), ertm/numerics.py:314-326."""
    if vmax <= 0:
        raise ValueError(f"vmax must be positive, got {vmax}")
    c = coefficient_l1(order)
    return float(safety / (vmax * np.hypot(c / dx, c / dz)))


def points_per_wavelength(vmin: float, f0: float, spacing: float) -> float:
    return float(vmin / (f0 * spacing)) if vmin > 0 and f0 > 0 else 0.0


def ricker(t: np.ndarray, f0: float, t0: float = 0.0) -> np.ndarray:
    """Ricker wavelet, peak 1 at t0 (ertm/numerics.py:466-469)."""
    u = np.pi * float(f0) * (np.asarray(t) - float(t0))
    return ((1.0 - 2.0 * u ** 2) * np.exp(-(u ** 2))).astype(np.float32)


def gaussian_pulse(t: np.ndarray, f0: float, t0: float = 0.0) -> np.ndarray:
    """Gaussian pulse (ertm/numerics.py:472-475)."""
    u = np.pi * float(f0) * (np.asarray(t) - float(t0))
    return np.exp(-0.5 * u ** 2).astype(np.float32)


def sinc_taps(n_in: int, dt_in: float, dt_out: float, *, half_length: int = 8,
              beta: float = 8.6) -> tuple[np.ndarray, np.ndarray]:
    """Weights (n_out, 2*half_length) float64 and tap rows (unclipped) of
    ``resample_sinc`` (ertm/numerics.py:370-383): output sample k sits at
    p = k*dt_out/dt_in, taps floor(p)-half_length+1 .. floor(p)+half_length,
    Kaiser-windowed sinc weights normalised to unit sum (NumPy's row sum)."""
    if dt_in <= 0 or dt_out <= 0:
        raise ValueError("dt_in and dt_out must be positive")
    if n_in < 3:
        raise ValueError(f"too few samples ({n_in}) to interpolate")
    n_out = int(np.floor((n_in - 1) * dt_in / dt_out + 1e-9)) + 1
    pos = np.arange(n_out, dtype=np.float64) * (dt_out / dt_in)
    taps = np.arange(-half_length + 1, half_length + 1)
    idx = np.floor(pos).astype(np.int64)[:, None] + taps[None, :]
    w = np.sinc(pos[:, None] - idx) * np.kaiser(2 * half_length, beta)[None, :]
    norm = w.sum(axis=1, keepdims=True)
    norm[norm == 0] = 1.0
    w /= norm
    return w, idx


def resample_sinc(data: np.ndarray, dt_in: float, dt_out: float, *, axis: int = 0,
                  half_length: int = 8, beta: float = 8.6) -> np.ndarray:
    """Kaiser-windowed sinc interpolation along time (ertm/numerics.py:329-392).

    Output length floor((n_in-1)*dt_in/dt_out) + 1; 2*half_length taps per
    output sample, normalised to unit sum; edge samples clamped.
    """
    x = np.asarray(data, dtype=np.float64)
    if dt_in <= 0 or dt_out <= 0:
        raise ValueError("dt_in and dt_out must be positive")
    squeeze = x.ndim == 1
    if squeeze:
        x = x[:, None]
    elif x.ndim != 2:
        raise ValueError(f"resample_sinc takes 1-D or 2-D data, got {x.ndim}-D")
    if axis == 1:
        x = x.T
    n_in = x.shape[0]
    w, idx = sinc_taps(n_in, dt_in, dt_out, half_length=half_length, beta=beta)
    out = np.einsum("ot,ots->os", w, x[np.clip(idx, 0, n_in - 1)])
    out = np.ascontiguousarray(out)
    if squeeze:
        out = out[:, 0]
    if axis == 1:
        out = out.T
    return out


# ---------------------------------------------------------------------------
# image filters (once per shot, float64)
# ---------------------------------------------------------------------------
def laplacian(field: np.ndarray, dx: float, dz: float) -> np.ndarray:
    """5-point Laplacian with edge padding (ertm/numerics.py:398-404)."""
    a = np.asarray(field, dtype=np.float64)
    q = np.pad(a, 1, mode="edge")
    mid = q[1:-1, 1:-1]
    return (q[1:-1, 2:] - 2.0 * mid + q[1:-1, :-2]) / (dx * dx) + (
        q[2:, 1:-1] - 2.0 * mid + q[:-2, 1:-1]) / (dz * dz)


def lowpass_wavenumber(field: np.ndarray, dx: float, dz: float,
                       cutoff_fraction: float = 0.85) -> np.ndarray:
    """Radial wavenumber low-pass with a 0.15-wide linear ramp (ertm/numerics.py:424-442)."""
    if not 0.0 < cutoff_fraction <= 1.0:
        raise ValueError(f"cutoff_fraction must lie in (0, 1], got {cutoff_fraction}")
    a = np.asarray(field, dtype=np.float64)
    nz, nx = a.shape
    kz = np.fft.fftfreq(nz, d=dz) / max(0.5 / dz, 1e-12)
    kx = np.fft.fftfreq(nx, d=dx) / max(0.5 / dx, 1e-12)
    radial = np.sqrt(kz[:, None] ** 2 + kx[None, :] ** 2)
    resp = np.clip(1.0 - (radial - cutoff_fraction) / 0.15, 0.0, 1.0)
    return np.real(np.fft.ifft2(np.fft.fft2(a) * resp))


def cosine_taper(field: np.ndarray, width: int, axes: tuple[int, ...] = (0, 1)) -> np.ndarray:
    """Cosine edge taper (ertm/numerics.py:445-463)."""
    out = np.array(field, dtype=np.float64, copy=True)
    width = int(max(0, width))
    if width == 0:
        return out
    for ax in axes:
        n = out.shape[ax]
        w = min(width, n // 2)
        if w <= 0:
            continue
        ramp = 0.5 - 0.5 * np.cos(np.pi * (np.arange(w) + 0.5) / w)
        view = np.moveaxis(out, ax, 0)
        shape = (w,) + (1,) * (view.ndim - 1)
        view[:w] *= ramp.reshape(shape)
        view[-w:] *= ramp[::-1].reshape(shape)
    return out
