"""Oracle GSLS constant-Q fit and per-Q relaxation tables.

Test infrastructure only.  ``fit`` restates ertm/viscoelastic.py:132-205
(log-parameterised weighted least squares over 24 log-spaced frequencies,
solved with SciPy's trust-region-reflective ``least_squares`` exactly as the
reference configures it); ``relaxation`` restates the per-unique-Q table
assembly of ertm/propagation.py:563-639, including the compensation rule
(band limit + gain budget, then delta -> -alpha * delta).
"""

from __future__ import annotations

import numpy as np


def modulus(tau, delta, freqs):
    """K(w)/K_R = 1 - sum_l delta_l / (1 + i w tau_l), ertm/viscoelastic.py:47-57."""
    w = 2.0 * np.pi * np.asarray(freqs, np.float64)
    out = np.ones_like(w, dtype=np.complex128)
    for t, d in zip(tau, delta):
        out -= d / (1.0 + 1j * w * t)
    return out


def quality(tau, delta, freqs):
    """Q = Re K / Im K, ertm/viscoelastic.py:81-88."""
    m = modulus(tau, delta, freqs)
    re, im = np.real(m), np.imag(m)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(np.abs(im) > 1e-12, re / im, np.inf)


def single(q, f_ref):
    """Closed-form single mechanism peaking at f_ref, ertm/viscoelastic.py:120-129."""
    u = (-1.0 + np.sqrt(1.0 + q * q)) / max(q, 1e-12)
    return float(u / (2.0 * np.pi * f_ref)), float(1.0 - u * u)


def fit(q, f_ref, n=3, band_ratio=1.5):
    """Return (tau, delta, fit_error), ertm/viscoelastic.py:132-205."""
    lo, hi = f_ref / band_ratio, f_ref * band_ratio
    freqs = np.geomspace(lo, hi, 24)
    if q > 1.0e4 or int(n) <= 1:
        t, d = single(q, f_ref)
        return np.array([t]), np.array([d]), 0.0
    from scipy.optimize import least_squares

    n = int(n)
    tau0 = 1.0 / (2.0 * np.pi * np.geomspace(lo, hi, n))
    d0 = np.full(n, 2.0 / q)
    wts = 1.0 + 0.5 * np.exp(-0.5 * (np.log(freqs / f_ref) / 0.25) ** 2)

    def resid(x):
        qm = quality(np.exp(x[:n]), np.exp(x[n:]), freqs)
        return (np.log(np.maximum(qm, 1e-6)) - np.log(q)) * wts

    x0 = np.concatenate([np.log(tau0), np.log(d0)])
    sol = least_squares(resid, x0, bounds=(x0 - 3.0, x0 + 3.0), max_nfev=400,
                        xtol=1e-10, ftol=1e-10)
    tau, delta = np.exp(sol.x[:n]), np.exp(sol.x[n:])
    err = float(np.max(np.abs(np.log(np.maximum(quality(tau, delta, freqs), 1e-6)) - np.log(q))))
    return tau, delta, err


def relaxation(q_padded, f_ref, n, band_ratio, dt, nt, compensation=False, alpha_c=1.0,
               f_max=None, gain_budget_db=40.0):
    """Unique-Q tables, ertm/propagation.py:563-639.

    Returns dict(unique, tau (U,n), delta (U,n), modulus_ref (U,), alpha (U,n)).
    """
    vals = np.clip(np.asarray(q_padded, np.float64).ravel(), 2.5, 1.0e5)
    uniq = np.unique(vals)
    if uniq.size > 32:
        uniq = np.array([float(np.median(vals))])
    tau = np.zeros((uniq.size, n))
    delta = np.zeros((uniq.size, n))
    mref = np.zeros(uniq.size)
    for i, q in enumerate(uniq):
        if q > 1.0e4:
            tau[i, :] = 1.0e-3
            mref[i] = 1.0
            continue
        t, d, _ = fit(float(q), f_ref, n, band_ratio)
        tau[i, : t.size] = t
        if t.size < n:
            tau[i, t.size :] = t[-1]
        delta[i, : d.size] = d
        mref[i] = float(np.real(modulus(t, d, [f_ref])[0]))
    alpha = np.zeros((uniq.size, n))
    if compensation:
        total_t = max(nt * dt, dt)
        budget = 10.0 ** (float(gain_budget_db) / 20.0)
        for i, q in enumerate(uniq):
            fpk = np.sqrt(np.maximum(1.0 - delta[i], 1e-6)) / (2.0 * np.pi * tau[i])
            keep = np.ones(n, bool) if f_max is None else fpk <= float(f_max)
            need = np.log(budget) * max(float(q), 1.0) / (np.pi * np.maximum(fpk, 1e-6) * total_t)
            al = np.where(keep, np.minimum(float(alpha_c), need), 0.0)
            alpha[i] = al
            delta[i] = -al * delta[i]
    return dict(unique=uniq, tau=tau, delta=delta, modulus_ref=mref, alpha=alpha)
