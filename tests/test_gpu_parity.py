"""GPU parity: the HIP engine against the reference's golden vectors and the
oracle.  Bar: bit-exact (``np.array_equal``) for every float32 propagation,
sampling, injection, replay and imaging result; the viscoacoustic
compensation path (hipFFT low-pass vs NumPy's pocketfft) is held to a
relative L2 tolerance of 1e-5 on records and 1e-4 on images.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

from golden_cases import forward_case, load, migrate_case, tags
from product_cases import migrate_objects, propagator_of, rel_l2

pytestmark = pytest.mark.gpu

VISCO_COMP_TOL_RECORD = 1e-5
VISCO_COMP_TOL_IMAGE = 1e-4


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _engine():
    from ertm import _lib

    _lib.lib()


@pytest.mark.parametrize("tag", tags("forward"))
def test_forward_matches_reference(tag):
    _say("forward golden", tag)
    from ertm.propagation import source_wavelet

    c = forward_case(tag)
    p, spec = propagator_of(c)
    from product_cases import source_of

    w = source_wavelet(source_of(c), spec.dt, spec.nt)
    rec = p.run_steps(w, c["components"])
    comp = c.get("q", {}).get("compensation", False)
    for name in c["components"]:
        if comp:
            assert rel_l2(rec[name], c["rec"][name]) < VISCO_COMP_TOL_RECORD, name
        else:
            assert np.array_equal(rec[name], c["rec"][name]), name
    if comp:
        s = p.quality_summary
        assert [s["clipped"], s["lowpass_steps"]] == list(c["visco_summary"])
        return
    for name, ref in c["state"].items():
        assert np.array_equal(getattr(p, name), ref), name
    psi = {k: p._core.expand_psi(k, v.cpu().numpy()) for k, v in p._core.psi_views().items()}
    for name, ref in c["psi"].items():
        got = psi.get(name, np.zeros_like(ref))
        assert np.array_equal(got, ref), name


def test_stepwise_api_matches_run_loop():
    """prop.step()+prop.sample() (the reference's loop) == the C-side loop."""
    _say("stepwise vs loop")
    from ertm.propagation import source_wavelet
    from product_cases import source_of

    c = forward_case("el_shear")
    p, spec = propagator_of(c)
    w = source_wavelet(source_of(c), spec.dt, spec.nt)
    rec = {n: np.zeros((spec.nt, len(c["rx"])), np.float32) for n in c["components"]}
    for s in range(spec.nt):
        p.step(float(w[s]))
        smp = p.sample(c["components"])
        for n in c["components"]:
            rec[n][s] = smp[n]
    for n in c["components"]:
        assert np.array_equal(rec[n], c["rec"][n]), n


@pytest.mark.parametrize("tag", ["ac", "el", "vi"])
def test_sample_and_inject_match_reference(tag):
    _say("inject golden", tag)
    from ertm.types import ReceiverSpec

    d = load("inject")
    case = {"ac": "ac_cpml", "el": "el_shear", "vi": "vi_fwd"}[tag]
    c = forward_case(case)
    p, _ = propagator_of(c)
    p.set_receivers(ReceiverSpec(x=d[f"{tag}/rx"], z=d[f"{tag}/rz"]))
    for name in p._field_names():
        p.set_field(name, d[f"{tag}/before_{name}"])
    comps = [str(v) for v in d[f"{tag}/components"]]
    for comp in comps:
        assert np.array_equal(p.sample((comp,))[comp], d[f"{tag}/sample_{comp}"]), comp
    for comp in comps:
        p.inject(comp, d[f"{tag}/values_{comp}"])
    for name in p._field_names():
        assert np.array_equal(getattr(p, name), d[f"{tag}/after_{name}"]), name


def test_replay_blocks_match_reference():
    _say("replay golden")
    from ertm.imaging import record_source_wavefield, replay_block
    from ertm.propagation import source_wavelet
    from product_cases import source_of

    d = load("replay")
    c = forward_case("ac_cpml")
    p, spec = propagator_of(c)
    w = source_wavelet(source_of(c), spec.dt, spec.nt)
    store = record_source_wavefield(p, w, budget_mb=1.0)
    assert store.interval == int(d["interval"])
    assert [s for s, _ in store.states] == list(d["starts"])
    for b in (0, 3, len(store.states) - 1):
        fields = replay_block(p, w, store, b, "stress")
        assert np.array_equal(np.stack([f["P"] for f in fields]), d[f"block{b}"]), b


def test_forward_to_last_checkpoint_saves_the_same_checkpoints():
    """A migration's checkpointing pass may end at the step that saves the
    last checkpoint: every checkpoint equals the full pass's, bit for bit,
    and the last block replays to the reference's fields."""
    _say("forward to the last checkpoint")
    from ertm.imaging import record_source_wavefields, replay_block
    from ertm.propagation import source_wavelet
    from product_cases import source_of

    d = load("replay")
    c = forward_case("ac_cpml")
    p_full, spec = propagator_of(c)
    p_cut, _spec = propagator_of(c)
    w = source_wavelet(source_of(c), spec.dt, spec.nt)
    full = record_source_wavefields([p_full], [w], budget_mb=1.0)[0]
    cut = record_source_wavefields([p_cut], [w], budget_mb=1.0, to_last_checkpoint=True)[0]
    last = cut.states[-1][0]
    assert full.forward_steps == spec.nt and cut.forward_steps == last + 1 < spec.nt
    assert p_cut.step_index == last + 1
    assert cut.states == full.states and cut.interval == full.interval
    assert np.array_equal(cut.arena.cpu().numpy(), full.arena.cpu().numpy())
    b = len(cut.states) - 1
    fields = replay_block(p_cut, w, cut, b, "stress")
    assert np.array_equal(np.stack([f["P"] for f in fields]), d[f"block{b}"])


@pytest.mark.parametrize("tag", tags("migrate"))
def test_migrate_matches_reference(tag):
    _say("migrate golden", tag)
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    res = RTMImager(imaging, solver, q).migrate(shot, model)
    assert (res.meta["checkpoints"]["interval"], res.meta["checkpoints"]["n_states"]) == \
        c["checkpoints"]
    assert res.meta["start_step"] == c["start_step"]
    if c["engine"] == "viscoacoustic" and c["q"]["enabled"]:
        assert rel_l2(res.image, c["image"]) < VISCO_COMP_TOL_IMAGE
    else:
        assert np.array_equal(res.image, c["image"])


@pytest.mark.parametrize("tag", ["ac_zero_lag", "el_ps_divcurl"])
def test_nested_replay_is_bit_identical(tag):
    """Segmented replay from sub-checkpoints reproduces the same image."""
    _say("nested replay", tag)
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    res = RTMImager(imaging, solver, q, GpuConfig(replay_slots=3)).migrate(shot, model)
    assert res.meta["replay_overhead"] > 1.0
    assert np.array_equal(res.image, c["image"])


@pytest.mark.parametrize("kv,ring", [((2, 19), 16), ((1, 10), 0), ((3, 64), 7), ((2, 8), 1),
                                     ((2, 48), 0), ((4, 11), 32)])
@pytest.mark.parametrize("tag", ["el_pp_divcurl", "el_ps_divcurl", "el_pp_stress", "el_ps_stress",
                                 "el_sp_zl", "el_ss_exc", "ac_zero_lag", "ac_cpml_zl",
                                 "ac_normalized", "ac_raw_rn", "ac_excitation", "vi_noq",
                                 "vi_norm"])
def test_migrate_with_fused_step_imaging_matches_reference(tag, kv, ring):
    """Fused step kernels that extract and correlate inside the step, or store
    the receiver fields in a ring folded every `ring` steps, give the
    reference images bit for bit."""
    _say("migrate fused imaging", tag, kv, ring)
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    res = RTMImager(imaging, solver, q, GpuConfig(kernel_variant=kv, image_ring=ring)).migrate(
        shot, model)
    assert np.array_equal(res.image, c["image"])


def test_fused_imaging_pp_ps_nested_batch():
    _say("fused imaging pp+ps nested batch")
    import dataclasses

    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case("el_pp_divcurl")
    shot, model, imaging, solver, q = migrate_objects(c)
    src2 = dataclasses.replace(shot.source, x=float(shot.source.x) + 3 * model.grid.dx)
    shots = [shot, dataclasses.replace(shot, source=src2)]
    out = RTMImager(imaging, solver, q, GpuConfig(kernel_variant=(2, 16), replay_slots=5)) \
        .migrate_batch(shots, model, modes=("pp", "ps"))
    assert np.array_equal(out[0]["pp"].image, c["image"])
    assert np.array_equal(out[0]["ps"].image, migrate_case("el_ps_divcurl")["image"])
    alone = RTMImager(imaging, solver, q).migrate_modes(shots[1], model, ("pp", "ps"))
    for m in ("pp", "ps"):
        assert np.array_equal(out[1][m].image, alone[m].image)


def test_fused_pp_ps_equals_separate_runs():
    _say("fused pp+ps")
    from ertm.imaging import RTMImager

    c = migrate_case("el_pp_divcurl")
    shot, model, imaging, solver, q = migrate_objects(c)
    both = RTMImager(imaging, solver, q).migrate_modes(shot, model, ("pp", "ps"))
    assert np.array_equal(both["pp"].image, c["image"])
    assert np.array_equal(both["ps"].image, migrate_case("el_ps_divcurl")["image"])


def test_checkpoint_restore_is_bit_exact():
    """state()/load_state() then continue == uninterrupted run (tests/test_rtm.py:188-214)."""
    _say("checkpoint restore")
    from ertm.propagation import source_wavelet
    from product_cases import source_of

    c = forward_case("ac_cpml")
    a, spec = propagator_of(c)
    b, _ = propagator_of(c)
    w = source_wavelet(source_of(c), spec.dt, spec.nt)
    for s in range(60):
        a.step(float(w[s]))
    b.step(float(w[0]))
    st = b.state()
    b.load_state(st)
    for s in range(1, 60):
        b.step(float(w[s]))
    for name in ("p", "vx", "vz"):
        assert np.array_equal(getattr(a, name), getattr(b, name))
    expected = sum(getattr(b, n).nbytes for n in ("p", "vx", "vz")) + b.pml.memory_bytes()
    assert b.state_bytes() == expected


def test_derivatives_match_reference():
    _say("derivatives")
    from ertm import numerics as num

    d = load("numerics")
    f = d["deriv_field"]
    for order in (2, 4, 6, 8):
        for axis in (0, 1):
            cs = num.staggered_coefficients(order)
            assert np.array_equal(num.dplus(f, axis, cs, 3.7), d[f"dplus_o{order}_a{axis}"])
            assert np.array_equal(num.dminus(f, axis, cs, 3.7), d[f"dminus_o{order}_a{axis}"])


def test_acoustic_c1_shape_forward_matches_oracle():
    """Config-1-shaped acoustic run (200 x 300, CPML 20) against the oracle, 300 steps."""
    _say("C1-shaped forward vs oracle")
    from ertm.propagation import PropagationSpec, build_propagator, source_wavelet
    from ertm.types import ElasticModel, GridSpec, ReceiverSpec, SourceSpec
    from oracle import rtm as ortm

    nz, nx = 200, 300
    z = np.arange(nz)[:, None] * 5.0
    vp = np.select([z < 250, z < 500, z < 750], [2000.0, 2500.0, 3000.0], 3500.0) * np.ones((1, nx))
    vp = vp.astype(np.float32)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    m = ElasticModel(GridSpec(nx=nx, nz=nz, dx=5.0, dz=5.0), vp=vp, vs=(vp / 1.8).astype(np.float32),
                     rho=rho, qp=np.full_like(vp, 100.0), qs=np.full_like(vp, 50.0))
    spec = PropagationSpec(dt=5e-4, nt=300, order=8, npml=20, boundary="cpml")
    src = SourceSpec(x=750.0, z=10.0, f0=25.0, delay=0.048)
    rec = ReceiverSpec(x=np.arange(300) * 5.0, z=np.full(300, 10.0))
    p = build_propagator(m, spec, engine="acoustic")
    p.set_source(src)
    p.set_receivers(rec)
    w = source_wavelet(src, spec.dt, spec.nt)
    got = p.run_steps(w, ("p",))["p"]
    om = ortm.Model(nz, nx, 5.0, 5.0, vp, m.vs, rho, m.qp)
    op = ortm.build("acoustic", om, spec.dt, spec.nt, 8, 20, "cpml")
    op.set_source(src.x, src.z)
    op.set_receivers(rec.x, rec.z)
    ref = ortm.forward(op, w, ("p",))["p"]
    assert np.array_equal(got, ref)
    assert np.array_equal(p.p, op.p)


@pytest.mark.parametrize("variant,zc", [(1, 64), (1, 10), (2, 19), (3, 10), (3, 64), (2, 4096),
                                        (2, 8), (1, 9), (3, 13), (2, 20), (4, 48), (4, 13)])
@pytest.mark.parametrize("tag", ["el_free", "el_shear", "el_force8", "ac_cpml", "ac_free",
                                 "vi_fwd", "ac_gauss"])
def test_fused_elastic_variants_bit_identical(tag, variant, zc):
    """Every fused launch shape (elastic, acoustic, viscoacoustic) reproduces
    the two-launch kernels (and thus the reference) bit for bit, across chunk
    seams, CPML strips, free surfaces and all sources."""
    _say("fused variant", tag, variant, zc)
    from ertm import device as dev
    from ertm.propagation import source_wavelet
    from product_cases import source_of

    c = forward_case(tag)
    assert c["order"] == 8  # the fused kernels' order
    p, spec = propagator_of(c)
    p._batches["step"] = dev.Batch([p._core], variant=(variant, zc))
    w = source_wavelet(source_of(c), spec.dt, spec.nt)
    for s in range(spec.nt):
        p.step(float(w[s]))
    for name, ref in c["state"].items():
        assert np.array_equal(getattr(p, name), ref), name
    psi = {k: p._core.expand_psi(k, v.cpu().numpy()) for k, v in p._core.psi_views().items()}
    for name, ref in c["psi"].items():
        assert np.array_equal(psi.get(name, np.zeros_like(ref)), ref), name


@pytest.mark.parametrize("engine", ["elastic", "acoustic", "viscoacoustic"])
def test_fused_variant_large_random_model_matches_two_launch(engine):
    """600 x 1200 (config-2 shape), random medium, CPML on all sides: fused
    (every variant) == two-launch after 60 steps, elastic and acoustic."""
    _say("fused vs two-launch C2 shape", engine)
    import torch
    from ertm import device as dev
    from ertm.propagation import PropagationSpec, build_propagator
    from ertm.types import ElasticModel, GridSpec, SourceSpec

    rng = np.random.default_rng(4)
    nz, nx = 600, 1200
    vp = rng.uniform(2200, 4500, (nz, nx)).astype(np.float32)
    m = ElasticModel(GridSpec(nx=nx, nz=nz, dx=5.0, dz=5.0), vp=vp,
                     vs=(vp / rng.uniform(1.6, 2.0, (nz, nx))).astype(np.float32),
                     rho=rng.uniform(1800, 2600, (nz, nx)).astype(np.float32),
                     qp=np.full_like(vp, 100.0), qs=np.full_like(vp, 50.0))
    spec = PropagationSpec(dt=4e-4, nt=60, order=8, npml=20, boundary="cpml")
    src = SourceSpec(x=3000.0, z=1500.0, f0=30.0, delay=0.01)
    w = np.sin(np.arange(60) * 0.3).astype(np.float32) * 1e6
    outs = []
    for v in [(0, 64), (1, 64), (1, 10), (2, 28), (3, 127), (2, 19), (2, 8), (3, 33), (4, 48),
              (4, 9)]:
        p = build_propagator(m, spec, engine=engine)
        p.set_source(src)
        p._batches["step"] = dev.Batch([p._core], variant=v)
        assert p._batches["step"].variant == v
        for s in range(60):
            p.step(float(w[s]))
        torch.cuda.synchronize()
        outs.append(p._core.state[p._core.cur].clone())
    for o in outs[1:]:
        assert torch.equal(o, outs[0])


# Device post-processing: normalisation, Laplacian and mask round exactly like
# NumPy; only the float64 FFT of the low-pass differs from pocketfft, by a few
# units in the last float64 place, so a float32 pixel may differ by one unit
# in the last place.  Tolerance: |device - host| <= 2**-22 * max|host|.
POST_TOL = 2.0 ** -22


@pytest.mark.parametrize("tag", ["ac_zero_lag", "ac_normalized", "ac_raw_rn", "ac_cpml_zl",
                                 "el_pp_divcurl", "el_ps_stress", "vi_noq"])
def test_device_postprocess_matches_host(tag):
    _say("device postprocess", tag)
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    res = RTMImager(imaging, solver, q, GpuConfig(device_postprocess=True)).migrate(shot, model)
    ref = c["image"]
    scale = float(np.abs(ref).max())
    assert scale > 0
    diff = np.abs(res.image.astype(np.float64) - ref.astype(np.float64))
    assert diff.max() <= POST_TOL * scale, (diff.max(), scale)
    if imaging.lowpass_fraction >= 1.0:
        assert np.array_equal(res.image, ref)


def test_device_postprocess_without_lowpass_is_bit_exact():
    _say("device postprocess no lowpass")
    import dataclasses

    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case("el_pp_divcurl")
    shot, model, imaging, solver, q = migrate_objects(c)
    imaging = dataclasses.replace(imaging, lowpass_fraction=1.0)
    host = RTMImager(imaging, solver, q).migrate(shot, model)
    devi = RTMImager(imaging, solver, q, GpuConfig(device_postprocess=True)).migrate(shot, model)
    assert np.array_equal(host.image, devi.image)


def test_device_postprocess_on_two_streams_at_once_equals_serial():
    """Two streams post-process different images of one shape at the same
    time (low-pass on, so both run hipFFT transforms): each result equals the
    serial one bit for bit.  Each stream gets its own plan, so the transforms
    never share a plan's work area."""
    _say("device postprocess, two streams")
    import torch

    from ertm import device as dev

    nz, nx = 600, 1200
    rng = np.random.default_rng(4)
    imgs = [torch.from_numpy(rng.standard_normal((nz, nx))).cuda() for _ in range(4)]
    ills = [torch.from_numpy(rng.uniform(0.5, 2.0, (nz, nx))).cuda() for _ in range(4)]
    kw = dict(dx=5.0, dz=5.0, eps=1e-10, div_ill=True, div_rill=False, laplacian=True,
              lowpass=True, fraction=0.85, npml=20, mask_top=False)
    serial = [dev.postprocess(a, b, None, **kw) for a, b in zip(imgs, ills)]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [None] * 4
    for rep in range(3):
        for k in range(4):
            with torch.cuda.stream(streams[k % 2]):
                outs[k] = dev.postprocess(imgs[k], ills[k], None, **kw)
        torch.cuda.synchronize()
        for k in range(4):
            assert torch.equal(outs[k], serial[k]), (rep, k)


@pytest.mark.parametrize("kv", [(0, 64), (2, 19), (1, 64)])
@pytest.mark.parametrize("boundary", ["cpml", "free"])
def test_subnormal_quotients_match_oracle(kv, boundary):
    """Derivative sums in the subnormal range, where s/h can be an exact
    rounding midpoint (dx = 10: a sum of 5 * 2^-149 gives 2^-150), next to
    normal-range cells in the same waves: every kernel path divides like
    NumPy's float32 division (ties to even), checked against the oracle."""
    _say("subnormal quotients", kv, boundary)
    from ertm import device as dev
    from ertm.propagation import PropagationSpec, build_propagator
    from ertm.types import ElasticModel, GridSpec
    from oracle import rtm as ortm

    nz, nx, h = 45, 70, 10.0
    rng = np.random.default_rng(7)
    vp = rng.uniform(2500.0, 3500.0, (nz, nx)).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = rng.uniform(2000.0, 2600.0, (nz, nx)).astype(np.float32)
    m = ElasticModel(GridSpec(nx=nx, nz=nz, dx=h, dz=h), vp=vp, vs=vs, rho=rho,
                     qp=np.full_like(vp, 100.0), qs=np.full_like(vp, 50.0))
    dt = 5e-4
    spec = PropagationSpec(dt=dt, nt=4, order=8, npml=10, boundary=boundary)
    p = build_propagator(m, spec, engine="elastic")
    p._batches["step"] = dev.Batch([p._core], variant=kv)
    op = ortm.build("elastic", ortm.Model(nz, nx, h, h, vp, vs, rho, m.qp), dt, 4, 8, 10, boundary)
    shape = op.vx.shape
    tiny = np.float32(2.0 ** -149)
    for name in ("vx", "vz", "sxx", "szz", "sxz"):
        # small multiples of 2^-149 (subnormal sums, many exact midpoints)
        # in most cells, normal-range values in a band of columns
        f = rng.integers(-40, 41, shape).astype(np.float32) * tiny
        f[:, 20:26] = rng.normal(0.0, 1e-3, (shape[0], 6)).astype(np.float32)
        getattr(op, name)[...] = f
        p.set_field(name, f)
    for _ in range(4):
        op.step(0.0)
        p.step(0.0)
    for name in ("vx", "vz", "sxx", "szz", "sxz"):
        got, want = getattr(p, name), getattr(op, name)
        assert np.array_equal(got, want), (name, int(np.sum(got != want)))
    assert np.any((np.abs(op.sxx) > 0) & (np.abs(op.sxx) < 2.0 ** -126))


@pytest.mark.parametrize("kv,ring,gsteps", [((2, 19), 16, 16), ((0, 64), 0, 4), ((1, 10), 4, 8),
                                            ((3, 64), 0, 2)])
@pytest.mark.parametrize("tag", ["el_pp_divcurl", "el_ps_divcurl", "el_sp_zl", "ac_zero_lag",
                                 "ac_normalized", "ac_cpml_zl", "vi_noq"])
def test_migrate_with_captured_time_loop_matches_reference(tag, kv, ring, gsteps):
    """The time loops captured as hipGraphs (chunks of gsteps steps replayed
    many times per pass, with eager tail steps, checkpoint restores and
    segment boundaries between the replays) give the reference images bit
    for bit, for fused and two-launch kernels, ring and per-step imaging."""
    _say("captured loop", tag, kv, ring, gsteps)
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    gpu = GpuConfig(kernel_variant=kv, image_ring=ring, graph_steps=gsteps)
    imager = RTMImager(imaging, solver, q, gpu)
    first = imager.migrate(shot, model)
    assert np.array_equal(first.image, c["image"])
    # the cached graphs are replayed again by a second migration
    again = imager.migrate(shot, model)
    assert np.array_equal(again.image, c["image"])


@pytest.mark.parametrize("seed", range(12))
def test_fused_variants_on_random_small_grids_match_two_launch(seed):
    """Odd grid sizes (down to a few chunks of 8 rows), every boundary, every
    engine and elastic source mechanism, sources near edges and corners: the
    whole state (fields and CPML memories) of every fused launch shape equals
    the two-launch kernels' after 40 steps."""
    import torch
    from ertm import device as dev
    from ertm.propagation import PropagationSpec, QOptions, build_propagator
    from ertm.types import ElasticModel, GridSpec, SourceSpec

    rng = np.random.default_rng(100 + seed)
    engine = ("elastic", "acoustic", "viscoacoustic")[seed % 3]
    boundary = ("cpml", "free", "none")[(seed // 3 + seed) % 3]
    nz, nx = int(rng.integers(23, 90)), int(rng.integers(19, 260))
    mech = ("explosive", "shear", "force_z")[seed % 3] if engine == "elastic" else "explosive"
    _say("random small grid", seed, engine, boundary, nz, nx, mech)
    vp = rng.uniform(2000, 4000, (nz, nx)).astype(np.float32)
    m = ElasticModel(GridSpec(nx=nx, nz=nz, dx=5.0, dz=4.0), vp=vp,
                     vs=(vp / rng.uniform(1.6, 2.0, (nz, nx))).astype(np.float32),
                     rho=rng.uniform(1800, 2600, (nz, nx)).astype(np.float32),
                     qp=rng.choice([30.0, 60.0, 120.0, 200.0], (nz, nx)).astype(np.float32),
                     qs=np.full_like(vp, 50.0))
    npml = int(rng.integers(5, 12))
    spec = PropagationSpec(dt=3e-4, nt=40, order=8, npml=npml, boundary=boundary)
    sx = float(rng.choice([0.0, (nx - 1) * 5.0, rng.uniform(0, (nx - 1) * 5.0)]))
    # one cell below the top: a source on a free-surface row is zeroed every step
    sz = float(rng.choice([4.0, (nz - 1) * 4.0, rng.uniform(4.0, (nz - 1) * 4.0)]))
    src = SourceSpec(x=sx, z=sz, f0=30.0, delay=0.01, mechanism=mech)
    q = QOptions(enabled=True, f_ref=30.0) if engine == "viscoacoustic" else None
    w = (np.sin(np.arange(40) * 0.4) * 1e5).astype(np.float32)
    shapes = [(0, 64), (1, 9), (2, 8), (3, 13), (4, 48), (2, 31), (4, 10)]
    outs = []
    for v in shapes:
        p = build_propagator(m, spec, engine=engine, q_options=q)
        p.set_source(src)
        p._batches["step"] = dev.Batch([p._core], variant=v)
        for s in range(40):
            p.step(float(w[s]))
        torch.cuda.synchronize()
        outs.append(p._core.state[p._core.cur].clone())
    assert float(outs[0].abs().max()) > 0.0
    for v, o in zip(shapes[1:], outs[1:]):
        assert torch.equal(o, outs[0]), v


@pytest.mark.parametrize("seed", range(6))
def test_fused_imaging_on_random_small_grids_matches_two_launch(seed):
    """Random odd-sized grids: migration with imaging inside the fused step
    (replay-slot extraction, receiver ring of random length, flush) equals
    migration on the two-launch kernels with a separate imaging pass, bit
    for bit, for PP and PS (elastic) or PP (acoustic) and every condition."""
    from ertm.config import GpuConfig, ImagingConfig, SolverConfig
    from ertm.imaging import RTMImager
    from ertm.propagation import PropagationSpec, forward_shot
    from ertm.types import ElasticModel, GridSpec, ReceiverSpec, SourceSpec

    rng = np.random.default_rng(200 + seed)
    engine = ("elastic", "acoustic")[seed % 2]
    boundary = ("cpml", "free")[(seed // 2) % 2]
    condition = ("normalized", "zero_lag", "excitation")[seed % 3]
    nz, nx = int(rng.integers(30, 80)), int(rng.integers(25, 200))
    _say("random small migration", seed, engine, boundary, condition, nz, nx)
    vp = rng.uniform(2000, 3500, (nz, nx)).astype(np.float32)
    vp[nz // 2:] += 800.0  # an interface to image
    m = ElasticModel(GridSpec(nx=nx, nz=nz, dx=5.0, dz=5.0), vp=vp,
                     vs=(vp / 1.8).astype(np.float32),
                     rho=rng.uniform(1900, 2500, (nz, nx)).astype(np.float32),
                     qp=np.full_like(vp, 100.0), qs=np.full_like(vp, 50.0))
    nt = 400
    solver = SolverConfig(nt=nt, dt=3e-4, order=8, boundary=boundary, engine=engine, npml=10)
    src = SourceSpec(x=float(rng.uniform(0.2, 0.8) * (nx - 1) * 5.0), z=10.0, f0=30.0,
                     delay=0.035)
    x = np.linspace(0.0, (nx - 1) * 5.0, max(5, nx // 3))
    comps = ("vx", "vz") if engine == "elastic" else ("p",)
    rec = ReceiverSpec(x=x, z=np.full_like(x, 10.0), component=comps[-1])
    shot = forward_shot(m, src, rec, PropagationSpec.from_config(solver, m), engine=engine,
                        components=comps)
    modes = ("pp", "ps") if engine == "elastic" and condition != "excitation" else ("pp",)
    # a budget of four fields-only states: four replay blocks
    per_state = (5 if engine == "elastic" else 3) * (nz + 10) * (nx + 10) * 4
    imaging = ImagingConfig(mode="pp", condition=condition, checkpoint_mb=4 * per_state / 2**20)
    ref = RTMImager(imaging, solver, None, GpuConfig(kernel_variant=(0, 64))).migrate_batch(
        [shot], m, modes=modes)[0]
    ring = int(rng.choice([1, 5, 16, 32]))
    for kv in [(2, 8), (4, 13), (1, 30)]:
        got = RTMImager(imaging, solver, None,
                        GpuConfig(kernel_variant=kv, image_ring=ring)).migrate_batch(
            [shot], m, modes=modes)[0]
        for md in modes:
            assert np.array_equal(got[md].image, ref[md].image), (kv, ring, md)
        assert np.abs(ref[modes[0]].image).max() > 0.0


@pytest.mark.parametrize("seed", range(6))
def test_fused_on_mid_size_grids_match_two_launch(seed):
    """Grids of 150-260 x 200-420 cells, where most blocks of a fused launch
    stream chunks away from every edge, beside blocks at CPML strips, the
    free surface, grid edges and the source: the whole state after 60 steps,
    the state and checkpoints of a checkpointing forward pass two steps per
    launch, and migrated images with imaging inside the step (one or two
    steps per launch) equal the two-launch kernels' bit for bit for every
    launch shape, boundary, source mechanism, mode pair and separation."""
    import torch
    from ertm import device as dev
    from ertm.config import GpuConfig, ImagingConfig, SolverConfig
    from ertm.imaging import RTMImager
    from ertm.propagation import PropagationSpec, build_propagator, forward_shot
    from ertm.types import ElasticModel, GridSpec, ReceiverSpec, SourceSpec

    rng = np.random.default_rng(300 + seed)
    boundary = ("none", "free", "cpml")[seed % 3]
    mech = ("explosive", "shear", "force_z")[(seed // 3 + seed) % 3]
    nz, nx = int(rng.integers(150, 260)), int(rng.integers(200, 420))
    _say("mid-size grid", seed, boundary, mech, nz, nx)
    vp = rng.uniform(2000, 3600, (nz, nx)).astype(np.float32)
    vp[nz // 2:] += 700.0
    m = ElasticModel(GridSpec(nx=nx, nz=nz, dx=5.0, dz=4.0), vp=vp,
                     vs=(vp / rng.uniform(1.6, 2.0, (nz, nx))).astype(np.float32),
                     rho=rng.uniform(1800, 2600, (nz, nx)).astype(np.float32),
                     qp=np.full_like(vp, 100.0), qs=np.full_like(vp, 50.0))
    spec = PropagationSpec(dt=3e-4, nt=60, order=8, npml=int(rng.integers(6, 14)),
                           boundary=boundary)
    src = SourceSpec(x=float(rng.uniform(0.1, 0.9) * (nx - 1) * 5.0),
                     z=float(rng.uniform(0.1, 0.9) * (nz - 1) * 4.0), f0=30.0, delay=0.01,
                     mechanism=mech)
    w = (np.sin(np.arange(60) * 0.4) * 1e5).astype(np.float32)
    shapes = [(0, 64), (1, 16), (2, 8), (2, 48), (3, 20), (4, 48), (4, 11)]
    outs = []
    for v in shapes:
        p = build_propagator(m, spec, engine="elastic")
        p.set_source(src)
        p._batches["step"] = dev.Batch([p._core], variant=v)
        for s in range(60):
            p.step(float(w[s]))
        torch.cuda.synchronize()
        outs.append(p._core.state[p._core.cur].clone())
    assert float(outs[0].abs().max()) > 0.0
    for v, o in zip(shapes[1:], outs[1:]):
        assert torch.equal(o, outs[0]), v

    # checkpointing forward pass two steps per launch (every built shape)
    # against one step per launch: final state and every checkpoint
    interval = 7
    ref = None
    for ts in [(0, 0), (2, 16), (2, 24), (2, 48), (2, 37)]:
        p = build_propagator(m, spec, engine="elastic")
        p.set_source(src)
        c = p._core
        arena = torch.zeros(((60 + interval - 1) // interval, c.state_floats),
                            dtype=torch.float32, device="cuda")
        b = dev.Batch([c], variant=(2, 48), two_step=ts)
        with dev.StepTimer():
            cur = b.run(0, 0, 60, c.cur, wavelets=[w], ck_interval=interval, ck_origin=0,
                        save=[arena], save_stride=c.state_floats)
            torch.cuda.synchronize()
        kinds = dev.StepTimer.read_kinds()
        b.close()
        assert (kinds["plain2"]["launches"] > 0) == (ts[0] != 0), (ts, kinds["plain2"])
        out = (c.state[cur].clone(), arena)
        if ref is None:
            ref = out
            assert torch.equal(out[0], outs[0])
            continue
        assert torch.equal(out[0], ref[0]), (ts, "state")
        assert torch.equal(out[1], ref[1]), (ts, "checkpoints")

    # migration: replay extraction and receiver ring inside the step
    modes = (("pp", "ps"), ("sp", "ss"), ("pp", "ps"))[seed % 3]
    separation = ("divcurl", "divcurl", "stress")[(seed + 1) % 3]
    bnd = "free" if boundary == "none" else boundary
    solver = SolverConfig(nt=300, dt=3e-4, order=8, boundary=bnd, engine="elastic", npml=10)
    ssrc = SourceSpec(x=float(rng.uniform(0.3, 0.7) * (nx - 1) * 5.0), z=12.0, f0=30.0,
                      delay=0.035)
    x = np.linspace(0.0, (nx - 1) * 5.0, nx // 4)
    rec = ReceiverSpec(x=x, z=np.full_like(x, 12.0), component="vz")
    shot = forward_shot(m, ssrc, rec, PropagationSpec.from_config(solver, m), engine="elastic",
                        components=("vx", "vz"))
    per_state = 5 * (nz + 20) * (nx + 20) * 4
    imaging = ImagingConfig(mode=modes[0], condition="normalized", separation=separation,
                            checkpoint_mb=3 * per_state / 2**20)
    ref = RTMImager(imaging, solver, None, GpuConfig(kernel_variant=(0, 64))).migrate_batch(
        [shot], m, modes=modes)[0]
    for kv, ts in [((2, 16), (0, 0)), ((4, 48), (0, 0)), ((1, 24), (0, 0)),
                   ((2, 48), (2, 16)), ((2, 48), (2, 40)), ((4, 24), (2, 48))]:
        got = RTMImager(imaging, solver, None,
                        GpuConfig(kernel_variant=kv, image_ring=int(rng.choice([1, 7, 32])),
                                  two_step=ts)).migrate_batch([shot], m, modes=modes)[0]
        for md in modes:
            assert np.array_equal(got[md].image, ref[md].image), (kv, ts, md)
    assert np.abs(ref[modes[0]].image).max() > 0.0
