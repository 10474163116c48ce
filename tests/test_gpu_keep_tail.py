"""Migrations whose checkpointing pass keeps the source fields of the last
steps (``GpuConfig.keep_tail``), on the GPU.

The golden migrate cases never reach a constant wavelet within their 300
steps, so these tests give them a wavelet that is the case's own up to step
``N_VARY`` and a constant after it.  From the first checkpoint block at or
after that step on, replaying a block repeats what the forward pass did one
step later, and the forward pass stores those source fields in the replay
slots instead; excitation imaging replays the blocks from there in one run
carried on to nt, so its forward pass ends at that block's checkpoint.
Each image must equal the oracle's (bit for bit, or within
the compensated viscoacoustic tolerance) and the image of the same migration
with the tail switched off, bit for bit, and the migration must have
executed fewer cell-updates than that one (the tail was used).  Cases: whole-block slots,
and slots fewer than a block (the tail starts mid-block and earlier blocks
are replayed in segments); one shot and a batch of three; lanes of
different schedules; the 64-shot workload's grid and launch shape at
nt = 600.
"""

from __future__ import annotations

import dataclasses
import sys

import numpy as np
import pytest

from golden_cases import migrate_case
from product_cases import migrate_objects, rel_l2

pytestmark = pytest.mark.gpu

N_VARY = 131
VISCO_COMP_TOL_IMAGE = 1e-4


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _engine():
    from ertm import _lib

    _lib.lib()


def _wavelet(c, tail_value=-0.0):
    from oracle import rtm as ortm

    s = c["src"]
    w = ortm.wavelet(s["kind"], s["f0"], s["delay"], s["amplitude"], c["dt"], c["nt"])
    w = np.array(w, np.float32)
    w[N_VARY:] = np.float32(tail_value)
    return w


def _oracle(c, w):
    from oracle import rtm as ortm

    m = ortm.Model(c["nz"], c["nx"], c["dx"], c["dz"], c["vp"], c["vs"], c["rho"], c["qp"],
                   c["x0"], c["z0"])
    q, q_enabled = None, False
    if c["engine"] == "viscoacoustic":
        qq = c["q"]
        q = ortm.QSpec(f_ref=qq["f_ref"], n=qq["n"], band_ratio=qq["band_ratio"],
                       alpha_c=qq["alpha_c"], amplitude_cap=qq["amplitude_cap"],
                       lowpass_interval=qq["lowpass_interval"], f_max=qq["f_max"],
                       gain_budget_db=qq["gain_budget_db"])
        q_enabled = qq["enabled"]
    img, _info = ortm.migrate(c["engine"], m, c["dt"], c["nt"], c["order"], c["boundary"],
                              ortm.Imaging(**c["imaging"]), c["shot"], c["src"], c["rx"],
                              c["rz"], w=w, q=q, q_enabled=q_enabled, rec_dt=c["shot_dt"])
    return img


def _moved(shot, dx):
    src = dataclasses.replace(shot.source, x=float(shot.source.x) + dx)
    return dataclasses.replace(shot, source=src)


# (tag, replay slots): 0 = whole blocks; elastic blocks are 12 steps, 23
# slots start the tail mid-block; acoustic blocks are 8 and visco 18 steps,
# 6 and 15 slots also replay the earlier blocks in segments
CASES = [(t, 0) for t in ("el_pp_divcurl", "el_ps_stress", "el_sp_zl", "ac_zero_lag",
                          "ac_raw_rn", "vi_noq", "vi_norm", "ac_excitation", "el_ss_exc")] + [
    ("el_pp_divcurl", 23), ("el_sp_zl", 23), ("ac_zero_lag", 6), ("vi_noq", 15),
    ("vi_norm", 15)]


@pytest.mark.parametrize("tag,slots", CASES)
def test_kept_tail_equals_replay_and_oracle(tag, slots):
    _say("kept tail", tag, "slots", slots)
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    w = _wavelet(c)
    mode = imaging.mode
    runs, work = {}, {}
    for keep in (True, False):
        imager = RTMImager(imaging, solver, q, gpu_config=GpuConfig(replay_slots=slots,
                                                                    keep_tail=keep))
        runs[keep] = imager.migrate(shot, model, wavelet=w)
        work[keep] = imager.last_timing["cell_updates"]
    kept, plain = runs[True], runs[False]
    _say("replayed", kept.meta["replayed_steps"], "against", plain.meta["replayed_steps"],
         "cell-updates", work[True], "against", work[False])
    assert plain.meta["replayed_steps"] >= c["nt"]
    assert work[True] < work[False]
    if imaging.condition == "excitation":
        # the peak pass replays every step; the forward pass ends early
        assert kept.meta["replayed_steps"] == plain.meta["replayed_steps"]
    else:
        assert kept.meta["replayed_steps"] < plain.meta["replayed_steps"]
    assert np.array_equal(kept.image, plain.image)
    assert kept.meta["checkpoints"] == plain.meta["checkpoints"]
    ref = _oracle(c, w)
    assert np.abs(ref).max() > 0
    if c["engine"] == "viscoacoustic" and c["q"]["enabled"]:
        # compensated viscoacoustic images carry the hipFFT low-pass
        assert rel_l2(kept.image, ref) < VISCO_COMP_TOL_IMAGE
    else:
        assert np.array_equal(kept.image, ref)


def test_kept_tail_in_a_batch_with_two_modes():
    _say("kept tail, batch of 3, pp + ps")
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case("el_pp_divcurl")
    shot, model, imaging, solver, q = migrate_objects(c)
    shots = [shot, _moved(shot, 2 * model.grid.dx), _moved(shot, -3 * model.grid.dx)]
    w = _wavelet(c)
    ws = [w, w, w]
    outs = {}
    for keep in (True, False):
        imager = RTMImager(imaging, solver, q, gpu_config=GpuConfig(keep_tail=keep))
        outs[keep] = imager.migrate_batch(shots, model, modes=("pp", "ps"), wavelets=ws)
        assert imager.last_timing["launch_groups"] == 1.0
    for a, b in zip(outs[True], outs[False]):
        for mode in ("pp", "ps"):
            assert a[mode].meta["replayed_steps"] < b[mode].meta["replayed_steps"]
            assert np.array_equal(a[mode].image, b[mode].image), mode
    assert np.array_equal(outs[True][0]["pp"].image, _oracle(c, w))


def test_kept_tail_on_the_64_shot_grid():
    """Two shots of the 64-shot workload's model (600 x 1200), PP and PS,
    normalized divcurl, its launch shape and in-step imaging; nt = 600 with
    blocks of 75 steps: the Ricker wavelet is bitwise -0.0 from step 378, so
    the forward pass keeps the steps from 450 on, or from 500 with 100 slots."""
    import bench
    from ertm.config import GpuConfig, ImagingConfig, SolverConfig
    from ertm.imaging import RTMImager

    nt = 600
    _say("64-shot grid, kept tail, nt", nt)
    model = bench.layered_model()
    sources, receivers = bench.acquisition(64)
    sources = [sources[5], sources[40]]
    shots = bench.model_records(model, sources, receivers, nt)
    solver = SolverConfig(nt=nt, dt=bench.DT, order=8, boundary="free", engine="elastic",
                          npml=20)
    imaging = ImagingConfig(mode="pp", condition="normalized", separation="divcurl",
                            checkpoint_mb=120.0)
    outs = {}
    for keep, slots in ((False, 0), (True, 0), (True, 100)):
        gpu = GpuConfig(strict_cfl=False, kernel_variant=(2, 48), image_ring=32,
                        keep_tail=keep, replay_slots=slots)
        imager = RTMImager(imaging, solver, gpu_config=gpu)
        outs[keep, slots] = imager.migrate_batch(shots, model, modes=("pp", "ps"))
        _say("keep", keep, "slots", slots, "replayed", outs[keep, slots][0]["pp"].meta[
            "replayed_steps"], "run", imager.last_timing.get("run_s"))
    base = outs[False, 0]
    assert base[0]["pp"].meta["checkpoints"]["interval"] == 75
    assert outs[True, 0][0]["pp"].meta["replayed_steps"] == 450
    assert outs[True, 100][0]["pp"].meta["replayed_steps"] == 500
    for key in ((True, 0), (True, 100)):
        for k in range(2):
            for mode in ("pp", "ps"):
                assert np.array_equal(outs[key][k][mode].image, base[k][mode].image), (key, k,
                                                                                       mode)


def test_kept_tail_in_lanes():
    """Lanes (ertm_run_lanes) of different schedules, each with a wavelet
    that turns constant: the kept tail's forward steps run as extracting
    lane steps.  Each image equals the same shot's migration without the
    tail, bit for bit, with fewer steps replayed."""
    _say("kept tail in lanes")
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager
    from ertm.lanes import migrate_lanes

    items, expected = [], []
    for tag, nt, slots in (("el_pp_divcurl", None, 0), ("el_ps_stress", 253, 0),
                           ("el_sp_zl", None, 23), ("el_pp_stress", None, 0),
                           ("ac_zero_lag", None, 6), ("ac_raw_rn", 277, 0)):
        c = migrate_case(tag)
        shot, model, imaging, solver, q = migrate_objects(c)
        if nt:
            solver = dataclasses.replace(solver, nt=int(nt))
            c = dict(c, nt=int(nt))
        w = _wavelet(c)
        im = RTMImager(imaging, solver, q, GpuConfig(replay_slots=slots))
        items.append((im, im.prepare_batch([shot], model, wavelets=[w])))
        plain = RTMImager(imaging, solver, q, GpuConfig(replay_slots=slots, keep_tail=False))
        expected.append((tag, imaging.mode, plain.migrate(shot, model, wavelet=w)))
    results = migrate_lanes(items)
    for res, (tag, mode, alone) in zip(results, expected):
        _say(tag, "replayed", res[mode].meta["replayed_steps"], "against",
             alone.meta["replayed_steps"])
        assert res[mode].meta["replayed_steps"] < alone.meta["replayed_steps"], tag
        assert np.abs(alone.image).max() > 0, tag
        assert np.array_equal(res[mode].image, alone.image), tag
