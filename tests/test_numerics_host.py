"""The package's host-side numerics (ertm/numerics.py) that feed the kernels
and the host post-processing: CPML recursion coefficients against the
reference's (golden vectors), stencil constants against their closed forms,
and the image Laplacian and wavenumber low-pass against the oracle, bit for
bit; and the byte layout of the receiver tables' single upload."""

from __future__ import annotations

from fractions import Fraction

import numpy as np
import pytest

from golden_cases import load
from oracle import numerics as onum

from ertm import numerics as num


def test_cpml_profiles_match_reference():
    d = load("numerics")
    for k in range(3):
        n, h, npml, dt, vmax, f0, lo, hi = d[f"cpml{k}_params"]
        a, b, w = num.cpml_profile(int(n), h, int(npml), dt, vmax, f0, bool(lo), bool(hi))
        assert np.array_equal(a, d[f"cpml{k}_a"])
        assert np.array_equal(b, d[f"cpml{k}_b"])
        assert w == max(1, min(int(npml), int(n) // 3))


def test_cpml_is_a_no_op_outside_the_strips():
    a, b, w = num.cpml_profile(200, 5.0, 20, 4e-4, 4500.0, 25.0)
    assert np.all(a[w:200 - w] == 0.0) and np.all(b[w:200 - w] == 1.0)
    assert np.all(b[:w] < 1.0) and np.all(a[:w] < 0.0)
    a, b, _ = num.cpml_profile(200, 5.0, 20, 4e-4, 4500.0, 25.0, enabled=False)
    assert np.all(a == 0.0) and np.all(b == 1.0)


def test_staggered_coefficients_closed_form():
    """Order 2K weights satisfy sum c_k (2k-1)^(2m-1) = [m == 1] for m = 1..K
    (exact rationals: 1225/1024, -245/3072, 49/5120, -5/7168 at order 8)."""
    exact8 = (Fraction(1225, 1024), Fraction(-245, 3072), Fraction(49, 5120), Fraction(-5, 7168))
    assert num.staggered_coefficients(8) == tuple(float(c) for c in exact8)
    for order in (2, 4, 6, 8):
        cs = num.staggered_coefficients(order)
        for m in range(1, len(cs) + 1):
            s = sum(c * (2 * k - 1) ** (2 * m - 1) for k, c in enumerate(cs, 1))
            assert abs(s - (1.0 if m == 1 else 0.0)) < 1e-12
    assert num.ghost_width(8) == 5
    assert abs(num.coefficient_l1(8) - 1.2863095) < 1e-7
    with pytest.raises(ValueError):
        num.staggered_coefficients(10)


def test_laplacian_and_lowpass_match_oracle():
    rng = np.random.default_rng(3)
    img = rng.normal(size=(47, 61))
    assert np.array_equal(num.laplacian(img, 5.0, 4.0), onum.laplacian(img, 5.0, 4.0))
    for frac in (0.85, 0.5):
        assert np.array_equal(num.lowpass_wavenumber(img, 5.0, 4.0, frac),
                              onum.lowpass(img, 5.0, 4.0, frac))


def test_cfl_dt_formula():
    c = num.coefficient_l1(8)
    expected = 0.5 / (3000.0 * np.hypot(c / 5.0, c / 4.0))
    assert abs(num.cfl_dt(3000.0, 5.0, 4.0, 8) - expected) < 1e-15


def test_packed_table_upload_layout():
    """The receiver tables go to the device as one buffer: every array starts
    on a 256-byte boundary, holds its values in its own dtype, and the next
    array starts after it (empty arrays included)."""
    from ertm.device import pack_host

    rng = np.random.default_rng(7)
    items = [(rng.integers(-5, 5, 37), np.int32), (np.arange(0), np.int32),
             (rng.integers(0, 2 ** 40, 11), np.int64), (rng.standard_normal(65), np.float64),
             (rng.standard_normal(3), np.float32)]
    host, offs = pack_host(items)
    assert host.dtype == np.uint8 and all(o % 256 == 0 for o in offs)
    sizes = [np.asarray(a, dtype=dt).nbytes for a, dt in items]
    assert all(offs[k] + sizes[k] <= offs[k + 1] for k in range(len(items) - 1))
    assert offs[-1] + sizes[-1] <= host.size
    for (a, dt), o, nb in zip(items, offs, sizes):
        back = host[o:o + nb].view(dt)
        assert np.array_equal(back, np.asarray(a, dtype=dt))
