"""The workloads at their own grid sizes on the GPU, against the oracle.

* 64-shot elastic workload (600 x 1200 at 5 m, free surface, the layered
  model with cavity and fault): two of its shots migrated as one batch for PP
  and PS in one pass, normalized condition, divcurl separation, with the
  benchmark's launch shape (128-lane blocks, 48-row chunks) and its imaging
  inside the step (source-P extraction in the replay pass, the receiver P/S
  ring in the backward pass); nt truncated to 150 with two checkpoint blocks
  and the cross term from step 0 (the pulse peaks at step 100).
  Each image equals the oracle's, bit for bit.
* Large elastic shot (2000 x 6000 at 2.5 m): nt truncated to 40, two
  checkpoint blocks, against the oracle bit for bit; and a 400-step
  migration whose blocks are replayed in segments from sub-checkpoints
  equals whole-block replay bit for bit.
* Borehole viscoacoustic workload (800 x 1600 at 4 m, four layers with Q
  30-200, GSLS L = 3), nt truncated to 400: the uncompensated forward
  wavefield of one of its shots equals the oracle's bit for bit; two shots'
  records (p, vx, vz), modelled on the GPU plus seeded noise (the first
  arrivals reach the well only after ~1200 steps, and the direct-wave mute
  zeroes the earlier samples), resampled to 36 us, rotated by the sensor
  azimuths and muted, are migrated as one batch with Q compensation in the
  backward pass (normalized condition, records aligned back to the solver's
  dt on the device, several checkpoint blocks); the first image agrees with
  the oracle's within the compensated viscoacoustic tolerance (relative L2
  1e-4: the stabiliser's float64 FFT low-pass is hipFFT here, pocketfft in
  NumPy).
* Serving workload at its real grid bounds (256 x 256 to 2000 x 4000, dx
  2.5-10 m, acoustic and elastic, nt capped at 1500, the bottom of its
  range): requests submitted together to the serving loop; each image equals
  the one its request gets alone, and the smallest acoustic and elastic
  requests equal the oracle bit for bit.

Records are modelled on the GPU (the migrations' input, not what is
checked).  The oracle runs on the host: these tests take a few minutes.
"""

from __future__ import annotations

import dataclasses
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _engine():
    from ertm import _lib

    _lib.lib()


def _oracle_image(engine, model, dt, nt, boundary, imaging, shot, mode):
    """oracle.rtm.migrate of one shot and one mode with the product's configs."""
    from oracle import rtm as ortm

    g = model.grid
    om = ortm.Model(g.nz, g.nx, g.dx, g.dz, model.vp, model.vs, model.rho, model.qp, g.x0, g.z0)
    ic = imaging
    cfg = ortm.Imaging(mode=mode, condition=ic.condition, separation=ic.separation,
                       illumination=ic.illumination,
                       receiver_normalization=ic.receiver_normalization,
                       laplacian=ic.laplacian, lowpass_fraction=ic.lowpass_fraction,
                       checkpoint_mb=ic.checkpoint_mb, eps=ic.eps,
                       start_periods=ic.start_periods)
    s = shot.source
    src = dict(x=s.x, z=s.z, f0=s.f0, delay=s.delay, amplitude=s.amplitude, kind=s.kind,
               mechanism=s.mechanism)
    img, info = ortm.migrate(engine, om, dt, nt, 8, boundary, cfg, dict(shot.components), src,
                             shot.receivers.x, shot.receivers.z, rec_dt=shot.dt)
    return img, info


def test_c2_shape_batched_pp_ps_matches_oracle():
    import bench
    from ertm.config import GpuConfig, ImagingConfig, SolverConfig
    from ertm.imaging import RTMImager

    nt = 150
    _say("64-shot workload shape: 2 shots, nt", nt)
    model = bench.layered_model()
    sources, receivers = bench.acquisition(64)
    sources = [sources[5], sources[40]]
    shots = bench.model_records(model, sources, receivers, nt)
    solver = SolverConfig(nt=nt, dt=bench.DT, order=8, boundary="free", engine="elastic",
                          npml=20)
    # fields-only state 14.8 MB: a 30 MB budget holds two checkpoints, blocks of 75
    # steps; the source pulse peaks at step 100, so the cross term starts at step 0
    imaging = ImagingConfig(mode="pp", condition="normalized", separation="divcurl",
                            checkpoint_mb=30.0, start_periods=0.0)
    gpu = GpuConfig(strict_cfl=False, kernel_variant=(2, 48), image_ring=32)
    out = RTMImager(imaging, solver, gpu_config=gpu).migrate_batch(shots, model,
                                                                  modes=("pp", "ps"))
    for k, shot in enumerate(shots):
        for mode in ("pp", "ps"):
            ref, info = _oracle_image("elastic", model, bench.DT, nt, "free", imaging, shot, mode)
            assert info["n_states"] == 2 and info["interval"] == 75
            got = out[k][mode]
            _say("shot", k, mode, "max|image|", float(np.abs(ref).max()))
            assert np.abs(ref).max() > 0
            assert got.meta["checkpoints"]["interval"] == 75
            assert np.array_equal(got.image, ref), (k, mode)


# compensated viscoacoustic images (hipFFT against pocketfft in the stabiliser)
VISCO_COMP_TOL_IMAGE = 1e-4


def test_borehole_visco_shape_matches_oracle():
    import bench
    from ertm.config import GpuConfig, ImagingConfig
    from ertm.imaging import RTMImager
    from ertm.preprocess import prepare_shot
    from ertm.propagation import PropagationSpec, build_propagator
    from ertm.solver import build_solver
    from oracle import rtm as ortm
    from product_cases import rel_l2

    nt = 400
    _say("borehole viscoacoustic workload: 2 shots, nt", nt)
    model = bench.c3_model()
    sources, receivers, az = bench.c3_acquisition(32)
    sources, az = [sources[15], sources[16]], [az[15], az[16]]
    solver, q = bench.c3_configs(nt)
    g = model.grid
    om = ortm.Model(g.nz, g.nx, g.dx, g.dz, model.vp, model.vs, model.rho, model.qp, g.x0, g.z0)
    qs = ortm.QSpec(f_ref=q.f_ref, n=q.n_mechanisms, band_ratio=q.band_ratio,
                    alpha_c=q.alpha_c, amplitude_cap=q.amplitude_cap,
                    lowpass_interval=q.lowpass_interval, f_max=q.f_max,
                    gain_budget_db=q.gain_budget_db)
    # forward modelling (no compensation): the whole wavefield after nt steps
    # against the oracle's, bit for bit (the wave has not reached the well yet)
    fwd = build_solver(solver, q, gpu_config=GpuConfig(strict_cfl=False))
    spec = PropagationSpec(dt=solver.dt, nt=nt, order=8, npml=solver.npml, boundary="cpml")
    prop = build_propagator(model, spec, engine="viscoacoustic", q_options=fwd.q_options())
    prop.set_source(sources[0])
    op = ortm.build("viscoacoustic", om, solver.dt, nt, 8, solver.npml, "cpml", q=qs)
    op.set_source(sources[0].x, sources[0].z)
    w = ortm.wavelet("ricker", sources[0].f0, sources[0].delay, 1.0, solver.dt, nt)
    for k in range(nt):
        prop.step(float(w[k]))
        op.step(float(w[k]))
    for f in ("p", "vx", "vz"):
        assert np.array_equal(getattr(prop, f), getattr(op, f)), f
    assert float(np.abs(op.p).max()) > 0
    # migration inputs: the modelled records plus seeded noise (so that data
    # survive the direct-wave mute), resampled to 36 us, rotated by the
    # sensor azimuths, muted and tapered
    raw = fwd.forward_batch(model, sources, receivers, components=("p", "vx", "vz"))
    rng = np.random.default_rng(33)
    for r in raw:
        for c in r.components:
            r.components[c] = (r.components[c]
                               + rng.normal(0.0, 1e-6, r.components[c].shape)).astype(np.float32)
    vref = float(np.percentile(model.vp, 25))
    shots = [prepare_shot(s, model=model, record_dt=bench.C3_RECORD_DT, rotation_deg=a,
                          mute=True, v_reference=vref) for s, a in zip(raw, az)]
    imaging = ImagingConfig(mode="pp", condition="normalized")
    out = RTMImager(imaging, solver, q, GpuConfig(strict_cfl=False)).migrate_batch(shots, model)
    cfg = ortm.Imaging(mode="pp", condition="normalized", separation=imaging.separation,
                       illumination=imaging.illumination,
                       receiver_normalization=imaging.receiver_normalization,
                       laplacian=imaging.laplacian, lowpass_fraction=imaging.lowpass_fraction,
                       checkpoint_mb=imaging.checkpoint_mb, eps=imaging.eps,
                       start_periods=imaging.start_periods)
    shot = shots[0]
    s = shot.source
    src = dict(x=s.x, z=s.z, f0=s.f0, delay=s.delay, amplitude=s.amplitude, kind=s.kind,
               mechanism=s.mechanism)
    ref, info = ortm.migrate("viscoacoustic", om, solver.dt, nt, 8, "cpml", cfg,
                             dict(shot.components), src, shot.receivers.x, shot.receivers.z,
                             q=qs, q_enabled=True, rec_dt=shot.dt)
    got = out[0]["pp"]
    err = rel_l2(got.image, ref)
    _say("azimuth", round(az[0], 2), "blocks", info["n_states"], "max|image|",
         float(np.abs(ref).max()), "rel L2", err)
    assert info["n_states"] >= 2
    assert np.abs(ref).max() > 0
    assert got.meta["checkpoints"]["interval"] == info["interval"]
    assert err <= VISCO_COMP_TOL_IMAGE, err
    assert np.isfinite(out[1]["pp"].image).all() and np.abs(out[1]["pp"].image).max() > 0


def _c5_shot(nt, delay):
    import bench
    from ertm.config import GpuConfig, SolverConfig
    from ertm.solver import build_solver
    from ertm.types import ReceiverSpec, SourceSpec

    model = bench.c5_model()
    nz, nx = model.shape
    src = SourceSpec(x=7500.0, z=5.0, f0=40.0, delay=delay, mechanism="explosive")
    recv = ReceiverSpec(x=np.arange(nx) * 2.5, z=np.full(nx, 5.0))
    solver = SolverConfig(nt=nt, dt=2.0e-4, order=8, boundary="free", engine="elastic", npml=20)
    shot = build_solver(solver, gpu_config=GpuConfig(strict_cfl=False)).forward(
        model, src, recv, components=("vx", "vz"))
    return model, solver, shot


def test_c5_shape_matches_oracle():
    from ertm.config import GpuConfig, ImagingConfig
    from ertm.imaging import RTMImager

    nt = 40
    _say("large single shot 2000 x 6000, nt", nt)
    # the pulse peaks at step 20; the cross term from step 0
    model, solver, shot = _c5_shot(nt, delay=0.004)
    # fields-only state 241.6 MB: a 500 MB budget holds two checkpoints
    imaging = ImagingConfig(mode="pp", condition="zero_lag", checkpoint_mb=500.0,
                            start_periods=0.0)
    got = RTMImager(imaging, solver, gpu_config=GpuConfig(strict_cfl=False)).migrate(shot, model)
    ref, info = _oracle_image("elastic", model, solver.dt, nt, "free", imaging, shot, "pp")
    assert info["n_states"] == 2
    assert np.abs(ref).max() > 0
    assert np.array_equal(got.image, ref)


def test_c5_shape_segment_replay_equals_whole_block_replay():
    from ertm.config import GpuConfig, ImagingConfig
    from ertm.imaging import RTMImager

    nt = 400
    _say("large single shot 2000 x 6000, segment replay, nt", nt)
    model, solver, shot = _c5_shot(nt, delay=0.03)
    # two checkpoints, blocks of 200 steps; 97 replay slots: segments of 6, 97, 97
    imaging = ImagingConfig(mode="pp", condition="zero_lag", checkpoint_mb=500.0)
    whole = RTMImager(imaging, solver, gpu_config=GpuConfig(strict_cfl=False)).migrate(shot, model)
    seg = RTMImager(imaging, solver,
                    gpu_config=GpuConfig(strict_cfl=False, replay_slots=97)).migrate(shot, model)
    assert whole.meta["checkpoints"]["interval"] == 200
    assert whole.meta["replay_overhead"] == pytest.approx(1.0)
    assert seg.meta["replay_overhead"] > 1.0
    assert np.abs(whole.image).max() > 0
    assert np.array_equal(seg.image, whole.image)


def test_serving_workload_at_its_sizes_alone_equals_mix_and_oracle():
    import bench
    from ertm.imaging import RTMImager
    from ertm.serving import ServingLoop

    nt_max = 1500
    reqs = [bench.serving_request(77, i, 1.0, nt_max=nt_max)[0] for i in range(5)]
    reqs.append(bench.serving_request(77, 100, 1.0, engine="elastic", nz=2000, nx=4000,
                                      nt_max=nt_max)[0])
    small = [bench.serving_request(77, 101, 1.0, engine="acoustic", nz=256, nx=256,
                                   nt_max=nt_max)[0],
             bench.serving_request(77, 102, 1.0, engine="elastic", nz=256, nx=256,
                                   nt_max=nt_max)[0]]
    reqs += small
    reqs = [dataclasses.replace(r, rid=k) for k, r in enumerate(reqs)]
    _say("serving at size:", [(r.solver.engine, r.model.shape, r.solver.nt) for r in reqs])
    with ServingLoop(hbm_budget_gb=200.0, streams=4) as loop:
        futs = [loop.submit(r) for r in reqs]
        res = [f.result(timeout=900) for f in futs]
        stats = loop.stats()
    _say("stats", stats)
    assert all(r.ok for r in res), [r.error for r in res if not r.ok]
    # requests of different nt, dt and grids shared launches (as lanes)
    assert stats["mean_group"] > 1.0 and stats["lanes_per_launch"] > 1.0, stats
    # arrivals overlapped: some requests started before others had finished
    assert any(a.start < b.finish and b.start < a.finish
               for i, a in enumerate(res) for b in res[i + 1:])
    for req, r in zip(reqs, res):
        alone = RTMImager(req.imaging, req.solver, req.q).migrate(req.shot, req.model)
        assert np.abs(alone.image).max() > 0, req.rid
        assert np.array_equal(r.images[req.imaging.mode].image, alone.image), req.rid
    for req, r in zip(reqs[-2:], res[-2:]):
        _say("oracle:", req.solver.engine, req.model.shape)
        ref, _ = _oracle_image(req.solver.engine, req.model, req.solver.dt, req.solver.nt,
                               req.solver.boundary, req.imaging, req.shot, req.imaging.mode)
        assert np.array_equal(r.images[req.imaging.mode].image, ref), req.solver.engine
