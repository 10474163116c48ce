"""Lanes (ertm_run_lanes) against the golden images and ``migrate`` alone.

Bar: bit-exact.  Shots whose nt, dt, grid, checkpoint interval, replay
segmentation, condition and imaging modes differ run as lanes of one runner;
each image equals the reference's golden image where the case's configs are
unchanged, and otherwise the image ``migrate`` gives the same shot alone.
Lanes that join a running set, and calls cut short by a tick limit in the
middle of passes, give the same images."""

from __future__ import annotations

import dataclasses
import sys

import numpy as np
import pytest

from golden_cases import migrate_case
from product_cases import migrate_objects
from serving_cases import make_request

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _engine():
    from ertm import _lib

    _lib.lib()


def _item(tag, nt=None, dt_scale=None, gpu=None, modes=None, **imaging_over):
    """(imager, one-shot job, expected images or None) for a golden case with
    optional changes; expected is the golden image when nothing changed."""
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    changed = bool(nt or dt_scale or imaging_over or modes or (gpu and gpu.replay_slots))
    if nt:
        solver = dataclasses.replace(solver, nt=int(nt))
    if dt_scale:
        solver = dataclasses.replace(solver, dt=solver.dt * float(dt_scale))
    if imaging_over:
        imaging = dataclasses.replace(imaging, **imaging_over)
    im = RTMImager(imaging, solver, q, gpu or GpuConfig())
    job = im.prepare_batch([shot], model, modes=modes)
    if changed:
        alone = RTMImager(imaging, solver, q, gpu or GpuConfig()).migrate_modes(
            shot, model, modes or (imaging.mode,))
        expected = {m: r.image for m, r in alone.items()}
    else:
        expected = {imaging.mode: c["image"]}
    return im, job, expected, tag


def _check(results, items):
    for res, (_im, _job, expected, tag) in zip(results, items):
        assert set(res) == set(expected), tag
        for m, img in expected.items():
            assert np.abs(img).max() > 0, (tag, m)
            assert np.array_equal(res[m].image, img), (tag, m)


def test_elastic_lanes_of_different_schedules_match_golden_and_alone():
    from ertm.config import GpuConfig
    from ertm.lanes import migrate_lanes

    _say("elastic lanes")
    items = [_item(t) for t in ("el_pp_divcurl", "el_ps_divcurl", "el_pp_stress", "el_ps_stress",
                                "el_sp_zl")]
    items += [
        _item("el_pp_divcurl", checkpoint_mb=0.02),            # more, shorter blocks
        _item("el_ps_stress", nt=253),                          # another nt
        _item("el_sp_zl", gpu=GpuConfig(replay_slots=7)),       # sub-checkpoint segments
        _item("el_pp_stress", dt_scale=0.8, nt=331, checkpoint_mb=0.05),
        _item("el_pp_divcurl", modes=("pp", "ps")),             # two receiver waves
    ]
    _check(migrate_lanes([(im, job) for im, job, _e, _t in items]), items)


def test_acoustic_lanes_of_different_schedules_match_golden_and_alone():
    from ertm.config import GpuConfig
    from ertm.lanes import migrate_lanes

    _say("acoustic lanes")
    items = [_item(t) for t in ("ac_zero_lag", "ac_normalized", "ac_raw_rn", "ac_cpml_zl",
                                "ac_resample")]
    items += [
        _item("ac_normalized", checkpoint_mb=0.01),
        _item("ac_zero_lag", nt=211, gpu=GpuConfig(replay_slots=5)),
        _item("ac_cpml_zl", dt_scale=0.75, nt=350),
    ]
    _check(migrate_lanes([(im, job) for im, job, _e, _t in items]), items)


def test_lanes_joining_a_running_set_and_resuming_mid_pass():
    """Requests of the serving kind on different grids and nt: two lanes
    start, two join after 57 ticks and one more after 200; every call is cut
    at 37 ticks, so passes stop and resume in the middle.  Launches are
    shared, and every image equals migrate alone."""
    from ertm.imaging import RTMImager
    from ertm.lanes import Lane, LaneRunner

    _say("lanes joining")
    reqs = [make_request(300 + k, engine="elastic", rid=k) for k in range(5)]
    runner = LaneRunner()
    lanes = {}

    def add(k):
        r = reqs[k]
        im = RTMImager(r.imaging, r.solver, r.q)
        lanes[k] = Lane(im, im.prepare_batch([r.shot], r.model), tag=k)
        runner.add(lanes[k])

    add(0)
    add(1)
    done = {}
    ticks = 0
    while True:
        if len(lanes) == 2 and (ticks >= 57 or not runner.lanes):
            add(2)
            add(3)
        elif len(lanes) == 4 and (ticks >= 200 or not runner.lanes):
            add(4)
        if not runner.lanes:
            break
        before = runner.ticks
        for ln in runner.run(max_ticks=37):
            done[ln.tag] = ln.result
        ticks += runner.ticks - before
        assert runner.ticks > before
    _say("ticks", runner.ticks, "lane steps", runner.lane_steps, "groups", runner.step_groups)
    assert runner.lane_steps > runner.step_groups      # lanes shared launches
    assert max(ln.mean_group for ln in lanes.values()) > 1.5
    for k, r in enumerate(reqs):
        alone = RTMImager(r.imaging, r.solver, r.q).migrate(r.shot, r.model)
        assert np.abs(alone.image).max() > 0, k
        assert np.array_equal(done[k][r.imaging.mode].image, alone.image), k
