"""Medium palette (host logic, no GPU): one byte per cell indexing a table
of the model's distinct (1/rho, mu, lambda) tuples, which the single-launch
elastic step reads instead of the three float32 arrays.  The table must
give back every cell's float32 values bit for bit, index 0 must be the zero
medium, and media with too many tuples must be refused."""
from __future__ import annotations

import numpy as np

from ertm import device as dev


def _layered(nz, nx, seed, nlayers=6):
    rng = np.random.default_rng(seed)
    vals = rng.uniform(1.0, 5.0, (3, nlayers)).astype(np.float32)
    rows = np.sort(rng.integers(0, nz, nlayers - 1))
    layer = np.searchsorted(rows, np.arange(nz), side="right")
    arrs = [np.repeat(v[layer][:, None], nx, axis=1) for v in vals]
    # an elliptical anomaly and a single odd cell
    zz, xx = np.mgrid[0:nz, 0:nx]
    ell = ((zz - nz / 2) / (nz / 5)) ** 2 + ((xx - nx / 3) / (nx / 7)) ** 2 < 1
    for a in arrs:
        a[ell] *= np.float32(0.6)
        a[3, 4] = np.float32(7.25)
    return arrs


def _check_round_trip(arrs):
    pal = dev.medium_palette(arrs)
    assert pal is not None
    index, table, n = pal
    assert index.dtype == np.uint8 and index.shape == arrs[0].shape
    assert table.dtype == np.float32 and table.shape == (256 * len(arrs),)
    assert index.min() >= 1 and index.max() == n
    for a, arr in enumerate(arrs):
        got = table[a * 256 + index.astype(np.int64)]
        assert np.array_equal(got.view(np.uint32), arr.view(np.uint32))
        assert table[a * 256] == 0.0 and table[a * 256].view(np.uint32) == 0
        assert not np.any(table[a * 256 + n + 1:(a + 1) * 256])
    return n


def test_layered_medium_round_trips_bit_for_bit():
    arrs = _layered(90, 140, 1)
    ref = {tuple(t) for t in np.stack([a.view(np.uint32).reshape(-1) for a in arrs], 1)}
    assert _check_round_trip(arrs) == len(ref)


def test_tuples_differing_in_one_array_or_only_in_sign_bits_are_distinct():
    a = np.zeros((4, 5), np.float32)
    b = np.zeros((4, 5), np.float32)
    c = np.zeros((4, 5), np.float32)
    b[1, 1] = -0.0          # same value as +0.0, different bits
    c[2, 3] = np.float32(1e-30)
    a[0, 4] = np.float32(2.0)
    assert _check_round_trip([a, b, c]) == 4


def test_too_many_tuples_are_refused():
    rng = np.random.default_rng(3)
    many = rng.integers(0, 300, (40, 50)).astype(np.float32)
    assert len(np.unique(many)) > 255
    assert dev.medium_palette([many, many, many]) is None
    assert dev.medium_palette([rng.standard_normal((300, 400)).astype(np.float32)] * 3) is None
    # exactly 255 tuples fit
    v = np.arange(255, dtype=np.float32).reshape(15, 17)
    assert _check_round_trip([v, v * 2, v * 3]) == 255


def test_the_workload_models_have_palettes():
    import bench

    m = bench.layered_model()
    pad = 5
    st = lambda a: np.pad(np.asarray(a, np.float32), pad, mode="edge")  # noqa: E731
    rho, vp, vs = st(m.rho), st(m.vp), st(m.vs)
    mu = (rho * vs * vs).astype(np.float32)
    lam = (rho * (vp * vp - 2.0 * vs * vs)).astype(np.float32)
    irho = (1.0 / rho).astype(np.float32)
    assert _check_round_trip([irho, mu, lam]) == 13
