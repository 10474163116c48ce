"""Generalized standard linear solid (GSLS) constant-Q fit (ertm/viscoelastic.py).

Host-side set-up for the viscoacoustic engine: for each distinct Q of the
model, L relaxation mechanisms (tau_l, delta_l) are fitted so that
Q(f) = Re K / Im K stays constant over [f_ref/r, f_ref*r].  The kernels then
evaluate, per cell, a_l = exp(-dt/tau_l) and
b_l = K_R delta_l (1 - a_l) tau_l / dt.

The fit is a weighted least squares in (log tau, log delta) over 24
log-spaced frequencies, solved with SciPy's bounded trust-region solver
(``least_squares``) with the reference's start point, bounds and tolerances
(ertm/viscoelastic.py:157-192), so the fitted mechanisms are the reference's.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass
from typing import Iterable

import numpy as np

from .exceptions import ERTMError

logger = logging.getLogger("ertm.viscoelastic")


def normalized_modulus(tau_sigma: Iterable[float], delta: Iterable[float],
                       freqs: Iterable[float]) -> np.ndarray:
    """K(w)/K_R = 1 - sum_l delta_l / (1 + i w tau_l)  (ertm/viscoelastic.py:47-57)."""
    w = 2.0 * np.pi * np.asarray(freqs, dtype=np.float64)
    k = np.ones_like(w, dtype=np.complex128)
    for tau, dl in zip(tau_sigma, delta):
        k -= dl / (1.0 + 1j * w * tau)
    return k


@dataclass(frozen=True)
class GSLSParameters:
    """Mechanisms fitted for one Q value (ertm/viscoelastic.py:60-101)."""

    q: float
    f_ref: float
    tau_sigma: tuple[float, ...]
    delta: tuple[float, ...]
    k_relaxed: float
    k_unrelaxed: float
    f_band: tuple[float, float] = (0.0, 0.0)
    fit_error: float = 0.0

    def modulus(self, freqs: Iterable[float]) -> np.ndarray:
        return normalized_modulus(self.tau_sigma, self.delta, freqs)

    def quality(self, freqs: Iterable[float]) -> np.ndarray:
        k = self.modulus(freqs)
        re, im = np.real(k), np.imag(k)
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(np.abs(im) > 1e-12, re / im, np.inf)

    def phase_velocity(self, freqs: Iterable[float], vp: float) -> np.ndarray:
        re = np.real(self.modulus(freqs))
        ref = float(np.real(self.modulus([self.f_ref])[0]))
        return vp * np.sqrt(np.maximum(re, 1e-9) / max(ref, 1e-12))

    def summary(self) -> dict[str, float]:
        return {"q": float(self.q), "f_ref": float(self.f_ref), "n_mechanisms": len(self.tau_sigma),
                "k_relaxed": float(self.k_relaxed), "k_unrelaxed": float(self.k_unrelaxed),
                "fit_error": float(self.fit_error)}


def _params(q, f_ref, tau, delta, band, err) -> GSLSParameters:
    return GSLSParameters(q=float(q), f_ref=float(f_ref), tau_sigma=tuple(float(t) for t in tau),
                          delta=tuple(float(d) for d in delta), k_relaxed=1.0,
                          k_unrelaxed=float(1.0 + np.sum(delta)),
                          f_band=(float(band[0]), float(band[1])), fit_error=float(err))


def single_mechanism(q: float, f_ref: float) -> tuple[float, float]:
    """One mechanism with its Q minimum (= q) at f_ref.

    Q_min = 2 sqrt(1-delta)/delta gives sqrt(1-delta) = (sqrt(1+q^2)-1)/q and
    the peak sits at 1/sqrt(tau_eps tau_sigma) = 2 pi f_ref.
    """
    u = (-1.0 + np.sqrt(1.0 + q * q)) / max(q, 1e-12)
    return float(u / (2.0 * np.pi * f_ref)), float(1.0 - u * u)


# Least-squares fits by (q, f_ref, n_mechanisms, band_ratio).  The fit is a
# deterministic function of these, so a cached result equals a new one; every
# propagator of a batch (and the backward propagators) fits the same few Q
# values of one model, at tens of milliseconds per fit.  Fallbacks to the
# closed form are not cached, so their warnings are logged on every call.
_FIT_CACHE: dict = {}
_FIT_CACHE_MAX = 4096


def fit_gsls(q: float, f_ref: float, *, n_mechanisms: int = 3, band_ratio: float = 1.5,
             min_q: float = 2.0) -> GSLSParameters:
    """Constant-Q fit over [f_ref/band_ratio, f_ref*band_ratio] (ertm/viscoelastic.py:132-205)."""
    if q <= min_q:
        raise ERTMError(f"GSLS fit needs Q > {min_q}, got {q}")
    if f_ref <= 0:
        raise ERTMError(f"f_ref must be positive, got {f_ref}")
    lo, hi = f_ref / band_ratio, f_ref * band_ratio
    freqs = np.geomspace(lo, hi, 24)
    if q > 1.0e4 or int(n_mechanisms) <= 1:
        tau, dl = single_mechanism(q, f_ref)
        return _params(q, f_ref, [tau], [dl], (lo, hi), 0.0)
    try:
        from scipy.optimize import least_squares
    except Exception:  # noqa: BLE001 - scipy is a dependency; the closed form as last resort
        logger.warning("scipy unavailable: GSLS falls back to the single-mechanism closed form")
        tau, dl = single_mechanism(q, f_ref)
        return _params(q, f_ref, [tau], [dl], (lo, hi), 0.0)
    n = int(n_mechanisms)
    key = (float(q), float(f_ref), n, float(band_ratio))
    hit = _FIT_CACHE.get(key)
    if hit is not None:
        return hit
    weights = 1.0 + 0.5 * np.exp(-0.5 * (np.log(freqs / f_ref) / 0.25) ** 2)
    x0 = np.concatenate([np.log(1.0 / (2.0 * np.pi * np.geomspace(lo, hi, n))),
                         np.log(np.full(n, 2.0 / q))])

    def q_of(x):
        k = normalized_modulus(np.exp(x[:n]), np.exp(x[n:]), freqs)
        re, im = np.real(k), np.imag(k)
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(np.abs(im) > 1e-12, re / im, np.inf)

    def residual(x):
        return (np.log(np.maximum(q_of(x), 1e-6)) - np.log(q)) * weights

    try:
        sol = least_squares(residual, x0, bounds=(x0 - 3.0, x0 + 3.0), max_nfev=400,
                            xtol=1e-10, ftol=1e-10)
    except Exception as exc:  # noqa: BLE001 - degenerate fit: closed form
        logger.warning("GSLS fit failed (%s): single-mechanism closed form", exc)
        tau, dl = single_mechanism(q, f_ref)
        return _params(q, f_ref, [tau], [dl], (lo, hi), 0.0)
    tau, dl = np.exp(sol.x[:n]), np.exp(sol.x[n:])
    err = float(np.max(np.abs(np.log(np.maximum(q_of(sol.x), 1e-6)) - np.log(q))))
    params = _params(q, f_ref, tau, dl, (lo, hi), err)
    if len(_FIT_CACHE) >= _FIT_CACHE_MAX:
        _FIT_CACHE.clear()
    _FIT_CACHE[key] = params
    return params


def gsls_table(q_values: Iterable[float], f_ref: float, *, n_mechanisms: int = 3,
               band_ratio: float = 1.5) -> dict[float, GSLSParameters]:
    """Fit every distinct Q once (ertm/viscoelastic.py:208-222)."""
    table: dict[float, GSLSParameters] = {}
    for v in sorted({float(q) for q in np.asarray(list(q_values), dtype=float)}):
        if not np.isfinite(v) or v <= 2.0:
            v = 1.0e4
        table[v] = fit_gsls(v, f_ref, n_mechanisms=n_mechanisms, band_ratio=band_ratio)
    return table


def estimate_q_spectral_ratio(reference: np.ndarray, attenuated: np.ndarray, freqs: np.ndarray,
                              travel_time: float, *, min_relative: float = 0.05) -> float:
    """Spectral-ratio Q: ln(A_ref/A_att) = pi f t / Q (ertm/viscoelastic.py:229-260)."""
    ref = np.asarray(reference, dtype=np.float64).ravel()
    att = np.asarray(attenuated, dtype=np.float64).ravel()
    f = np.asarray(freqs, dtype=np.float64).ravel()
    if ref.size != att.size or f.size != ref.size:
        raise ERTMError("spectral ratio needs equal-length reference, attenuated and frequency axes")
    ratio = np.log(np.maximum(np.abs(ref), 1e-30) / np.maximum(np.abs(att), 1e-30))
    rmax = float(np.max(np.abs(ref))) if ref.size else 0.0
    amax = float(np.max(np.abs(att))) if att.size else 0.0
    band = ((f > 0) & (np.abs(ref) > min_relative * rmax) & (np.abs(att) > min_relative * amax)
            & (ratio > 0))
    if np.count_nonzero(band) < 3 or travel_time <= 0:
        return float("inf")
    slope = float(np.polyfit(f[band], ratio[band], 1)[0])
    return float(np.pi * travel_time / max(slope, 1e-12))


def relaxation_tables(q_padded: np.ndarray, f_ref: float, n: int, band_ratio: float, dt: float,
                      nt: int, *, compensation: bool = False, alpha_c: float = 1.0,
                      f_max: float | None = None, gain_budget_db: float = 40.0) -> dict:
    """Per-unique-Q mechanism tables for the kernel (ertm/propagation.py:563-639).

    Returns unique Q values (at most 32, else their median -- the reference's
    rule), tau and delta (U, n), modulus_ref (U,) = Re K_norm(f_ref) from the
    fitted (attenuating) delta, per-cell class index, fit error, and, when
    compensating, the band-limited, budget-capped alpha_l with delta -> -alpha delta.
    """
    vals = np.clip(np.asarray(q_padded, dtype=np.float64).ravel(), 2.5, 1.0e5)
    uniq = np.unique(vals)
    if uniq.size > 32:
        uniq = np.array([float(np.median(vals))])
    tau = np.zeros((uniq.size, n))
    delta = np.zeros((uniq.size, n))
    mref = np.zeros(uniq.size)
    err = 0.0
    for i, q in enumerate(uniq):
        if q > 1.0e4:
            tau[i, :] = 1.0e-3
            mref[i] = 1.0
            continue
        p = fit_gsls(float(q), f_ref, n_mechanisms=n, band_ratio=band_ratio)
        err = max(err, p.fit_error)
        t, d = np.asarray(p.tau_sigma), np.asarray(p.delta)
        tau[i, : t.size] = t
        if t.size < n:
            tau[i, t.size :] = t[-1]
        delta[i, : d.size] = d
        mref[i] = float(np.real(normalized_modulus(t, d, [f_ref])[0]))
    index = np.searchsorted(uniq, vals)
    if index.max(initial=0) >= uniq.size:
        raise ERTMError("more than 32 distinct Q values: the per-Q tables cannot cover the model")
    alpha = np.zeros((uniq.size, n))
    fpk = np.sqrt(np.maximum(1.0 - delta, 1e-6)) / (2.0 * np.pi * tau)
    if compensation:
        total = max(nt * dt, dt)
        budget = 10.0 ** (float(gain_budget_db) / 20.0)
        for i, q in enumerate(uniq):
            keep = np.ones(n, bool) if f_max is None else fpk[i] <= float(f_max)
            need = np.log(budget) * max(float(q), 1.0) / (np.pi * np.maximum(fpk[i], 1e-6) * total)
            alpha[i] = np.where(keep, np.minimum(float(alpha_c), need), 0.0)
            delta[i] = -alpha[i] * delta[i]
    return dict(unique=uniq, tau=tau, delta=delta, modulus_ref=mref, index=index.reshape(
        np.shape(q_padded)), alpha=alpha, f_peak=fpk, fit_error=err)
