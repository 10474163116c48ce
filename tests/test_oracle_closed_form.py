"""Closed-form and invariant checks on the oracle (independent of the golden
vectors): each is chosen so that a plausible slip -- a dropped stencil term,
a wrong sign or offset, a transposed axis -- fails it."""

from __future__ import annotations

import numpy as np

from oracle import gsls as ogsls
from oracle import numerics as onum
from oracle import rtm as ortm


def test_coefficients_are_consistent():
    for order, cs in onum.COEFFS.items():
        # sum c_k (2k - 1) == 1: exact for a linear function
        assert abs(sum(c * (2 * k - 1) for k, c in enumerate(cs, 1)) - 1.0) < 1e-15
        # higher moments vanish: sum c_k (2k-1)^(2m+1) == 0 for 1 <= m < K
        for mpow in range(1, len(cs)):
            assert abs(sum(c * (2 * k - 1) ** (2 * mpow + 1) for k, c in enumerate(cs, 1))) < 1e-12
    assert onum.ghost(8) == 5
    assert abs(onum.coef_l1(8) - 1.2863095238095) < 1e-12


def _loop_derivative(f, axis, order, h, plus):
    g = np.moveaxis(np.asarray(f, np.float64), axis, 0)
    out = np.zeros_like(g)
    cs = onum.COEFFS[order]
    kk = len(cs)
    n = g.shape[0]
    for i in range(kk, n - kk):
        acc = 0.0
        for k, c in enumerate(cs, 1):
            acc += c * ((g[i + k] - g[i - k + 1]) if plus else (g[i + k - 1] - g[i - k]))
        out[i] = acc / h
    return np.moveaxis(out, 0, axis)


def test_derivative_equals_brute_force_loops():
    rng = np.random.default_rng(5)
    f = rng.standard_normal((17, 14))
    for order in (2, 4, 6, 8):
        for axis in (0, 1):
            for plus in (True, False):
                got = onum.derivative(f, axis, order, 2.5, plus)
                ref = _loop_derivative(f, axis, order, 2.5, plus)
                np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_derivative_of_sine_is_eighth_order_accurate():
    # the reference's own accuracy test (tests/test_rtm.py:139-157)
    x = np.arange(401) * 1.0
    k = 2.0 * np.pi * 0.15
    f = np.tile(np.sin(k * x), (6, 1))
    exact = k * np.cos(k * (x + 0.5))
    e8 = np.max(np.abs(onum.derivative(f, 1, 8, 1.0, True)[0, 10:-10] - exact[10:-10])) / k
    e4 = np.max(np.abs(onum.derivative(f, 1, 4, 1.0, True)[0, 10:-10] - exact[10:-10])) / k
    assert e8 < 1e-3 and e4 < 1e-2 and e4 > e8
    # the minus derivative sits half a cell to the left
    em = onum.derivative(f, 1, 8, 1.0, False)[0, 10:-10] - k * np.cos(k * (x - 0.5))[10:-10]
    assert np.max(np.abs(em)) / k < 1e-3


def test_cpml_is_identity_outside_strips_and_decays_inside():
    a, b, w = onum.cpml_profile(100, 5.0, 20, 5e-4, 3000.0, 25.0)
    assert w == 20
    assert np.all(a[20:80] == 0) and np.all(b[20:80] == 1)
    assert np.all(b[:20] < 1) and np.all(a[:20] < 0)
    assert np.array_equal(a[:20], a[80:][::-1]) and np.array_equal(b[:20], b[80:][::-1])
    a2, b2, _ = onum.cpml_profile(100, 5.0, 20, 5e-4, 3000.0, 25.0, lo_side=False)
    assert np.all(a2[:80] == 0) and np.all(b2[:80] == 1)
    a3, b3, _ = onum.cpml_profile(100, 5.0, 20, 5e-4, 3000.0, 25.0, enabled=False)
    assert np.all(a3 == 0) and np.all(b3 == 1)
    # width is capped at a third of the axis
    assert onum.cpml_profile(30, 5.0, 20, 5e-4, 3000.0, 25.0)[2] == 10


def test_cfl_formula():
    assert abs(onum.cfl_dt(4000.0, 10.0, 10.0, 8, 0.55)
               - 0.55 / (4000.0 * np.sqrt(2.0) * onum.coef_l1(8) / 10.0)) < 1e-18


def test_single_mechanism_q_peaks_at_reference_frequency():
    tau, delta = ogsls.single(40.0, 25.0)
    assert abs(float(ogsls.quality([tau], [delta], [25.0])[0]) - 40.0) < 1e-9
    assert float(ogsls.quality([tau], [delta], [10.0])[0]) > 40.0


def test_three_mechanism_fit_is_constant_q():
    tau, delta, _ = ogsls.fit(40.0, 25.0, 3, 1.5)
    qf = ogsls.quality(tau, delta, np.geomspace(25.0 / 1.5, 25.0 * 1.5, 9))
    assert np.max(np.abs(qf / 40.0 - 1.0)) < 0.02
    assert 1.0 < 1.0 + delta.sum() < 1.2


def test_sample_is_transpose_of_inject():
    """<S f, v> == <f, S^T v> for the bilinear sample/inject pair (float64)."""
    from oracle import engines as oeng

    g = oeng.Grid(20, 24, 5.0, 5.0)
    rng = np.random.default_rng(1)
    vp = np.full((20, 24), 3000.0, np.float32)
    p = oeng.Acoustic(g, vp, np.full_like(vp, 2000.0), 1e-4, dtype="float64")
    p.set_receivers(rng.uniform(0, 110, 7), rng.uniform(0, 90, 7))
    f = rng.standard_normal(g.shape)
    v = rng.standard_normal(7)
    for comp in ("p", "vx", "vz"):
        lhs = float(np.dot(p._sample(f, comp), v))
        fld = np.zeros(g.shape)
        p._scatter(fld, v, comp)
        rhs = float(np.sum(f * fld))
        assert abs(lhs - rhs) < 1e-10 * max(1.0, abs(lhs))


def test_acoustic_moveout_matches_constant_velocity():
    """Reference behaviour test tests/test_rtm.py:221-235 on the oracle."""
    nz, nx = 121, 161
    m = ortm.Model(nz, nx, 10.0, 10.0, np.full((nz, nx), 3000.0, np.float32),
                   np.full((nz, nx), 1700.0, np.float32), np.full((nz, nx), 2400.0, np.float32),
                   np.full((nz, nx), 100.0, np.float32))
    p = ortm.build("acoustic", m, 4e-4, 700, 8, 15, "cpml")
    p.set_source(800.0, 600.0)
    p.set_receivers([1000.0, 1300.0], [600.0, 600.0])
    w = ortm.wavelet("ricker", 20.0, 0.06, 1.0, 4e-4, 700)
    rec = ortm.forward(p, w, ("p",))["p"]

    def lag(tr):
        c = np.correlate(tr, w, mode="full")
        return (int(np.argmax(np.abs(c))) - (w.size - 1)) * 4e-4

    assert abs((lag(rec[:, 1]) - lag(rec[:, 0])) - 300.0 / 3000.0) < 0.01
