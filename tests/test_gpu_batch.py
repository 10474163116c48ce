"""Batched migration: shots advanced together give each shot its alone result.

``RTMImager.migrate_batch`` advances every shot of a launch group with one
kernel launch per time step.  These tests check, bit for bit, that a shot's
image inside a batch equals the reference golden image (for the golden shot)
and equals the image the same shot gets from a plain ``migrate`` call (for
the others), for equal and for different grid sizes in one launch.
"""

from __future__ import annotations

import dataclasses
import sys

import numpy as np
import pytest

from golden_cases import migrate_case
from product_cases import migrate_objects, rel_l2

pytestmark = pytest.mark.gpu

VISCO_COMP_TOL_IMAGE = 1e-4


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _engine():
    from ertm import _lib

    _lib.lib()


def _moved(shot, dx):
    src = dataclasses.replace(shot.source, x=float(shot.source.x) + dx)
    return dataclasses.replace(shot, source=src)


@pytest.mark.parametrize("tag", ["ac_zero_lag", "ac_excitation", "el_pp_divcurl", "el_ss_exc",
                                 "vi_noq", "vi_norm"])
def test_batch_equals_alone_and_golden(tag):
    _say("batch vs alone", tag)
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    imager = RTMImager(imaging, solver, q)
    shots = [shot, _moved(shot, 2 * model.grid.dx), _moved(shot, -3 * model.grid.dx)]
    out = imager.migrate_batch(shots, model)
    assert imager.last_timing["launch_groups"] == 1.0
    mode = imaging.mode
    if c["engine"] == "viscoacoustic" and c["q"]["enabled"]:
        # compensated viscoacoustic images carry the hipFFT low-pass
        assert rel_l2(out[0][mode].image, c["image"]) < VISCO_COMP_TOL_IMAGE
    else:
        assert np.array_equal(out[0][mode].image, c["image"])
    for s, o in zip(shots[1:], out[1:]):
        alone = RTMImager(imaging, solver, q).migrate(s, model)
        assert np.array_equal(o[mode].image, alone.image)
        assert o[mode].meta["checkpoints"] == alone.meta["checkpoints"]


def test_batch_of_uneven_grids_in_one_launch():
    """Two grids of different size with the same checkpoint interval."""
    _say("batch uneven grids")
    from ertm.imaging import RTMImager
    from ertm.types import ElasticModel, GridSpec

    c = migrate_case("ac_zero_lag")
    shot, model, imaging, solver, q = migrate_objects(c)
    nz, nx = model.shape
    g = model.grid
    pad = lambda a: np.pad(np.asarray(a), ((0, 1), (0, 2)), mode="edge")  # noqa: E731
    big = ElasticModel(grid=GridSpec(nx=nx + 2, nz=nz + 1, dx=g.dx, dz=g.dz, x0=g.x0, z0=g.z0),
                       vp=pad(model.vp), vs=pad(model.vs), rho=pad(model.rho), qp=pad(model.qp),
                       qs=pad(model.qs))
    imager = RTMImager(imaging, solver, q)
    from ertm.propagation import PropagationSpec

    k0 = imager._batch_key(model, PropagationSpec.from_config(solver, model))
    k1 = imager._batch_key(big, PropagationSpec.from_config(solver, big))
    if k0 != k1:
        pytest.skip(f"grids do not share a checkpoint interval: {k0} vs {k1}")
    out = imager.migrate_batch([shot, shot], [model, big])
    assert imager.last_timing["launch_groups"] == 1.0
    assert np.array_equal(out[0]["pp"].image, c["image"])
    alone = RTMImager(imaging, solver, q).migrate(shot, big)
    assert out[1]["pp"].image.shape == (nz + 1, nx + 2)
    assert np.array_equal(out[1]["pp"].image, alone.image)


def test_batch_nested_replay_and_pp_ps():
    """Segmented replay inside a batch and fused pp+ps keep alone results."""
    _say("batch nested pp+ps")
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case("el_pp_divcurl")
    shot, model, imaging, solver, q = migrate_objects(c)
    shots = [shot, _moved(shot, 4 * model.grid.dx)]
    out = RTMImager(imaging, solver, q, GpuConfig(replay_slots=5)).migrate_batch(
        shots, model, modes=("pp", "ps"))
    assert np.array_equal(out[0]["pp"].image, c["image"])
    assert np.array_equal(out[0]["ps"].image, migrate_case("el_ps_divcurl")["image"])
    alone = RTMImager(imaging, solver, q).migrate_modes(shots[1], model, ("pp", "ps"))
    for m in ("pp", "ps"):
        assert np.array_equal(out[1][m].image, alone[m].image)


def test_compensated_visco_ring_equals_imaging_after_the_stabiliser():
    """Compensated viscoacoustic backward passes store the pressure of each
    step in the receiver-field ring before the stabiliser runs; the flush
    multiplies each entry by its step's clip factor, and low-pass steps
    rewrite their entry.  With a small amplitude cap (clipping at most
    steps) and a low-pass every 7 steps, inside ring windows of 5 and 32
    steps, the images equal the separate imaging pass after the stabiliser
    bit for bit."""
    _say("compensated visco ring")
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case("vi_norm")
    shot, model, imaging, solver, q = migrate_objects(c)
    q = dataclasses.replace(q, amplitude_cap=1.5, lowpass_interval=7)
    shots = [shot, _moved(shot, 3 * model.grid.dx)]
    outs = {}
    for ring in (0, 5, 32):
        outs[ring] = RTMImager(imaging, solver, q, GpuConfig(kernel_variant=(2, 16),
                                                             image_ring=ring)).migrate_batch(
            shots, model)
    meta = outs[0][0]["pp"].meta["q"]
    assert meta["clipped_steps"] > 0 and meta["lowpass_steps"] > 0, meta
    for ring in (5, 32):
        for a, b in zip(outs[0], outs[ring]):
            assert np.array_equal(a["pp"].image, b["pp"].image), ring
            assert a["pp"].meta["q"]["clipped_steps"] == b["pp"].meta["q"]["clipped_steps"]
    assert np.abs(outs[0][0]["pp"].image).max() > 0


def test_batch_max_batch_splits_groups():
    _say("batch max_batch")
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case("ac_zero_lag")
    shot, model, imaging, solver, q = migrate_objects(c)
    shots = [_moved(shot, k * model.grid.dx) for k in range(3)]
    imager = RTMImager(imaging, solver, q, GpuConfig(max_batch=2))
    out = imager.migrate_batch(shots, model)
    assert imager.last_timing["launch_groups"] == 2.0
    ref = RTMImager(imaging, solver, q).migrate_batch(shots, model)
    for a, b in zip(out, ref):
        assert np.array_equal(a["pp"].image, b["pp"].image)


@pytest.mark.parametrize("tag", ["ac_cpml", "el_free", "vi_fwd"])
def test_forward_batch_equals_single_shots(tag):
    _say("forward batch", tag)
    from golden_cases import forward_case
    from product_cases import model_of, receivers_of, source_of
    from ertm.config import GpuConfig, QConfig, SolverConfig
    from ertm.solver import build_solver

    c = forward_case(tag)
    model = model_of(c)
    q = None
    if c["engine"] == "viscoacoustic":
        qq = c["q"]
        q = QConfig(enabled=True, model="gsls", f_ref=qq["f_ref"], n_mechanisms=qq["n"],
                    band_ratio=qq["band_ratio"], alpha_c=qq["alpha_c"],
                    amplitude_cap=qq["amplitude_cap"], lowpass_interval=qq["lowpass_interval"],
                    f_max=qq["f_max"], gain_budget_db=qq["gain_budget_db"])
    solver = build_solver(SolverConfig(nt=c["nt"], dt=c["dt"], order=c["order"],
                                       boundary=c["boundary"], engine=c["engine"],
                                       npml=c["npml"]), q, GpuConfig(strict_cfl=False))
    src = source_of(c)
    rcv = receivers_of(c)
    sources = [src] + [dataclasses.replace(src, x=float(src.x) + k * c["dx"]) for k in (1, -2)]
    comps = tuple(str(k) for k in c["components"])
    batch = solver.forward_batch(model, sources, rcv, components=comps)
    for s, rec in zip(sources, batch):
        alone = solver.forward(model, s, rcv, components=comps)
        for name in comps:
            assert np.array_equal(rec.components[name], alone.components[name])


@pytest.mark.parametrize("tag", ["ac_zero_lag", "el_ps_divcurl"])
def test_pipelined_batches_equal_run_batch(tag):
    """run_batches overlaps one group's host post-processing with the next
    group's GPU passes (here three groups: max_batch 2 over 3 shots, then a
    second job); every image and its metadata equal run_batch's."""
    _say("pipelined batches", tag)
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    imager = RTMImager(imaging, solver, q, GpuConfig(max_batch=2))
    modes = ("pp", "ps") if c["engine"] == "elastic" else ("pp",)
    job_a = imager.prepare_batch([shot, _moved(shot, 2 * model.grid.dx),
                                  _moved(shot, -3 * model.grid.dx)], model, modes=modes)
    job_b = imager.prepare_batch([_moved(shot, model.grid.dx)], model, modes=modes)
    alone = [imager.run_batch(job_a), imager.run_batch(job_b)]
    piped = imager.run_batches([job_a, job_b])
    assert imager.last_timing["launch_groups"] == 3.0
    for got_job, ref_job in zip(piped, alone):
        assert len(got_job) == len(ref_job)
        for got, ref in zip(got_job, ref_job):
            for m in modes:
                assert np.array_equal(got[m].image, ref[m].image)
                assert got[m].meta == ref[m].meta
    # the first shot is the golden one
    assert np.array_equal(piped[0][0][c["imaging"]["mode"]].image, c["image"])


def test_pipelined_batches_on_a_caller_stream_equal_default_stream():
    """Batched calls launched on a stream of the caller's keep set-up and
    host copies on that stream (no side streams); results equal those of
    the default stream, where side streams overlap them with the passes."""
    _say("pipelined batches on a caller stream")
    import torch

    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager, side_stream

    c = migrate_case("el_ps_divcurl")
    shot, model, imaging, solver, q = migrate_objects(c)
    imager = RTMImager(imaging, solver, q, GpuConfig(max_batch=2))
    job_a = imager.prepare_batch([shot, _moved(shot, 2 * model.grid.dx),
                                  _moved(shot, -3 * model.grid.dx)], model, modes=("pp", "ps"))
    job_b = imager.prepare_batch([_moved(shot, model.grid.dx)], model, modes=("pp", "ps"))
    ref = imager.run_batches([job_a, job_b])
    assert side_stream("setup") != torch.cuda.current_stream()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        assert side_stream("setup") == s and side_stream("fetch") == s
        got = imager.run_batches([job_a, job_b])
    s.synchronize()
    for got_job, ref_job in zip(got, ref):
        for g, r in zip(got_job, ref_job):
            for m in ("pp", "ps"):
                assert np.array_equal(g[m].image, r[m].image)
                assert g[m].meta == r[m].meta


@pytest.mark.parametrize("ring", [1, 7, 16])
@pytest.mark.parametrize("tag", ["el_pp_divcurl", "el_ps_stress", "ac_normalized", "ac_raw_rn"])
def test_ring_folds_beside_the_steps_equal_folds_between_them(tag, ring):
    """GpuConfig.overlap_flush: the receiver ring (2 x ring steps) is folded
    every ``ring`` steps on a stream of the library's own while the next
    backward steps run.  Each image equals the golden image and the one the
    folds on the launching stream give, bit for bit, alone and for a batch
    of three shots imaging PP and PS."""
    _say("folds beside the steps", tag, ring)
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    outs = {}
    for beside in (True, False):
        gpu = GpuConfig(kernel_variant=(2, 19), image_ring=ring, overlap_flush=beside)
        outs[beside] = RTMImager(imaging, solver, q, gpu_config=gpu).migrate(shot, model)
    assert np.array_equal(outs[True].image, c["image"])
    assert np.array_equal(outs[True].image, outs[False].image)
    if c["engine"] == "elastic":
        shots = [shot, _moved(shot, 2 * model.grid.dx), _moved(shot, -3 * model.grid.dx)]
        got = {}
        for beside in (True, False):
            gpu = GpuConfig(kernel_variant=(2, 19), image_ring=ring, overlap_flush=beside)
            got[beside] = RTMImager(imaging, solver, q, gpu_config=gpu).migrate_batch(
                shots, model, modes=("pp", "ps"))
        for a, b in zip(got[True], got[False]):
            for mode in ("pp", "ps"):
                assert np.abs(b[mode].image).max() > 0
                assert np.array_equal(a[mode].image, b[mode].image), mode
