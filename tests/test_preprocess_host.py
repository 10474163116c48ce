"""Record preparation on the host (ertm/preprocess.py), the reference's own
preprocessing tests (tests/test_rtm.py:607-648) on synthetic records, plus
checks of the formulas."""

from __future__ import annotations

import numpy as np
import pytest

from ertm.preprocess import (
    mute_direct_wave,
    mute_weights,
    prepare_shot,
    resample_shot,
    rotate_components,
    taper_record_ends,
)
from ertm.types import ElasticModel, GridSpec, ReceiverSpec, ShotRecord, SourceSpec


def _shot(nt=400, nrec=21, dt=5e-4, seed=0):
    rng = np.random.default_rng(seed)
    x = np.linspace(0.0, 600.0, nrec)
    data = rng.normal(size=(nt, nrec)).astype(np.float32)
    return ShotRecord(data=data, dt=dt, t=np.arange(nt) * dt,
                      receivers=ReceiverSpec(x=x, z=np.full(nrec, 20.0)),
                      source=SourceSpec(x=300.0, z=20.0, f0=20.0, delay=0.06),
                      components={"p": data})


def test_rotation_preserves_energy():
    rng = np.random.default_rng(1)
    vx = rng.normal(size=(64, 5)).astype(np.float32)
    vz = rng.normal(size=(64, 5)).astype(np.float32)
    rot = rotate_components({"p": vx + vz, "vx": vx, "vz": vz}, 30.0)
    before = np.sum(vx ** 2) + np.sum(vz ** 2)
    after = np.sum(rot["vx"] ** 2) + np.sum(rot["vz"] ** 2)
    assert abs(float(before) - float(after)) < 5e-3
    assert np.array_equal(rot["p"], vx + vz)


def test_rotation_by_90_degrees_swaps_components():
    vx = np.array([[1.0], [2.0]], dtype=np.float32)
    vz = np.array([[3.0], [4.0]], dtype=np.float32)
    rot = rotate_components({"vx": vx, "vz": vz}, 90.0)
    assert np.allclose(rot["vx"], vz, atol=1e-6) and np.allclose(rot["vz"], -vx, atol=1e-6)


def test_resample_roundtrip_keeps_shape_and_dt():
    shot = _shot(nt=200)
    coarse = resample_shot(shot, 1.0e-3)
    fine = resample_shot(coarse, 5e-4)
    assert coarse.dt == pytest.approx(1e-3) and fine.dt == pytest.approx(5e-4)
    assert fine.data.shape[1] == shot.data.shape[1] and np.isfinite(fine.data).all()


def test_mute_weights_follow_the_direct_arrival():
    shot = _shot(nt=600)
    w = mute_weights(shot, 3000.0)
    period = 1.0 / 20.0
    offset = np.abs(shot.receivers.x - 300.0)
    t_mute = 0.06 + 1.2 * period + 1.05 * offset / 3000.0
    for i in (0, 10, 20):
        t = shot.t
        assert np.all(w[t <= t_mute[i] - 0.5 * period, i] == 0.0)
        assert np.all(w[t >= t_mute[i] + 0.5 * period, i] == 1.0)
        mid = np.argmin(np.abs(t - t_mute[i]))
        assert abs(w[mid, i] - 0.5) < 0.1


def test_direct_mute_removes_early_energy():
    shot = _shot()
    muted = mute_direct_wave(shot, v_reference=3000.0)
    early = shot.t < 0.08
    assert np.abs(muted.data[early]).max() < np.abs(shot.data[early]).max()
    assert np.abs(muted.data[early]).max() < 0.3 * np.abs(shot.data).max()


def test_end_taper_and_prepare_chain():
    shot = _shot(nt=100)
    tap = taper_record_ends(shot, 0.05)
    w = 5
    ramp = 0.5 - 0.5 * np.cos(np.pi * (np.arange(w) + 0.5) / w)
    assert np.array_equal(tap.data[:w], (shot.data[:w].astype(np.float64) * ramp[:, None]).astype(np.float32))
    assert np.array_equal(tap.data[w:-w], shot.data[w:-w])
    g = GridSpec(nx=61, nz=51, dx=10.0, dz=10.0)
    vp = np.full((51, 61), 3000.0, np.float32)
    model = ElasticModel(g, vp=vp, vs=vp / 1.8, rho=np.full_like(vp, 2000.0),
                         qp=np.full_like(vp, 100.0), qs=np.full_like(vp, 50.0))
    prep = prepare_shot(shot, model=model)
    manual = taper_record_ends(mute_direct_wave(shot, 3000.0), 0.05)
    assert np.array_equal(prep.components["p"], manual.components["p"])
