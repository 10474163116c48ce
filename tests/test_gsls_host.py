"""The package's host-side GSLS fit (ertm/viscoelastic.py), which builds the
viscoacoustic relaxation tables the kernels read: pinned to the reference's
fitted mechanisms (golden vectors, written from the reference package) and
held to the reference's own fit-quality thresholds
(tests/test_rtm.py:377-389)."""

from __future__ import annotations

import numpy as np

from golden_cases import load

from ertm.viscoelastic import fit_gsls, single_mechanism


def test_fit_matches_reference_mechanisms():
    rows = load("gsls")["table"]
    assert len(rows) > 0
    for row in rows:
        q, f_ref, n, err = row[:4]
        k = int(n)
        p = fit_gsls(float(q), float(f_ref), n_mechanisms=k, band_ratio=1.5)
        np.testing.assert_allclose(p.tau_sigma, row[4:4 + k], rtol=1e-12)
        np.testing.assert_allclose(p.delta, row[7:7 + k], rtol=1e-12)
        assert abs(p.fit_error - err) <= 1e-12
        assert p.k_unrelaxed == 1.0 + float(np.sum(p.delta))


def test_constant_q_fit_over_band():
    """Q(f)/Q within 2 % over the band, 1 < K_U < 1.2 (tests/test_rtm.py:377-384)."""
    p = fit_gsls(40.0, 25.0, n_mechanisms=3, band_ratio=1.5)
    quality = p.quality(np.geomspace(25.0 / 1.5, 25.0 * 1.5, 9))
    assert np.max(np.abs(quality / 40.0 - 1.0)) < 0.02
    assert 1.0 < p.k_unrelaxed < 1.2


def test_single_mechanism_peaks_at_reference_frequency():
    """One mechanism has Q = 40 at f_ref and more away from it
    (tests/test_rtm.py:386-389)."""
    p = fit_gsls(40.0, 25.0, n_mechanisms=1)
    assert abs(float(p.quality([25.0])[0]) - 40.0) < 0.5
    assert float(p.quality([10.0])[0]) > 40.0
    tau, delta = single_mechanism(40.0, 25.0)
    assert p.tau_sigma == (tau,) and p.delta == (delta,)


def test_nearly_elastic_q_uses_one_mechanism():
    p = fit_gsls(2.0e4, 30.0, n_mechanisms=3)
    assert len(p.tau_sigma) == 1 and p.fit_error == 0.0


def test_cached_fit_equals_a_fresh_fit():
    """A repeated fit returns the cached parameters; they equal a fit made
    after the cache is cleared, value for value."""
    from ertm import viscoelastic as ve

    a = ve.fit_gsls(57.0, 30.0, n_mechanisms=3, band_ratio=1.5)
    assert ve.fit_gsls(57.0, 30.0, n_mechanisms=3, band_ratio=1.5) is a
    ve._FIT_CACHE.clear()
    b = ve.fit_gsls(57.0, 30.0, n_mechanisms=3, band_ratio=1.5)
    assert b is not a
    assert a == b
