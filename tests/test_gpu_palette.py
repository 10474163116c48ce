"""Medium palette on the GPU: the single-launch elastic and acoustic steps
reading one byte of medium per cell (an index into the model's distinct
(1/rho, mu, lambda) or (1/rho, kappa) tuples) instead of float32 arrays.  Bar: bit-exact.  On layered models
(where the palette applies) the whole state after 40 steps, fields and CPML
memories, equals the same launch shape without the palette and the
two-launch kernels (which always read the float32 arrays), for every launch
shape, boundary and source mechanism; and PP+PS migrations with the in-step
imaging variants give the same images with and without it.  The workloads'
own models are layered too, so the oracle tests at their sizes
(`test_gpu_configs.py`) run the palette path as well."""

from __future__ import annotations

import contextlib
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


@contextlib.contextmanager
def _palette(on: bool):
    from ertm import device as dev

    prev = dev.MEDIUM_PALETTE
    dev.MEDIUM_PALETTE = on
    dev._HOST_MEDIA.clear()
    try:
        yield
    finally:
        dev.MEDIUM_PALETTE = prev
        dev._HOST_MEDIA.clear()


def _layered_model(nz, nx, rng, dx=5.0, dz=4.0):
    from ertm.types import ElasticModel, GridSpec

    nl = int(rng.integers(3, 9))
    tops = np.sort(rng.integers(1, nz, nl - 1))
    layer = np.searchsorted(tops, np.arange(nz), side="right")
    vps = np.sort(rng.uniform(1800, 4500, nl)).astype(np.float32)
    vp = np.repeat(vps[layer][:, None], nx, axis=1)
    vs = (vp / np.float32(1.8)).astype(np.float32)
    rho = (310.0 * vp ** 0.25).astype(np.float32)
    zz, xx = np.mgrid[0:nz, 0:nx]
    for _ in range(int(rng.integers(0, 3))):
        cz, cx = rng.uniform(0, nz), rng.uniform(0, nx)
        ell = ((zz - cz) / rng.uniform(3, 20)) ** 2 + ((xx - cx) / rng.uniform(3, 40)) ** 2 < 1
        vp[ell] *= np.float32(0.6)
        vs[ell] *= np.float32(0.6)
        rho[ell] *= np.float32(0.85)
    return ElasticModel(GridSpec(nx=nx, nz=nz, dx=dx, dz=dz), vp=vp, vs=vs, rho=rho,
                        qp=np.full_like(vp, 100.0), qs=np.full_like(vp, 50.0))


@pytest.mark.parametrize("engine", ["elastic", "acoustic"])
@pytest.mark.parametrize("seed", range(8))
def test_palette_step_equals_float_media_on_every_launch_shape(seed, engine):
    import torch
    from ertm import device as dev
    from ertm.propagation import PropagationSpec, build_propagator
    from ertm.types import SourceSpec

    rng = np.random.default_rng(700 + seed)
    boundary = ("cpml", "free", "none")[seed % 3]
    mech = (("explosive", "shear", "force_z")[(seed // 3 + seed) % 3] if engine == "elastic"
            else "explosive")
    if seed == 7:
        nz, nx = 600, 1200
    else:
        nz, nx = int(rng.integers(23, 150)), int(rng.integers(19, 330))
    _say("palette", engine, seed, boundary, mech, nz, nx)
    m = _layered_model(nz, nx, rng)
    npml = int(rng.integers(5, 21))
    spec = PropagationSpec(dt=3e-4, nt=40, order=8, npml=npml, boundary=boundary)
    src = SourceSpec(x=float(rng.uniform(0, (nx - 1) * 5.0)),
                     z=float(rng.choice([4.0, rng.uniform(4.0, (nz - 1) * 4.0)])),
                     f0=30.0, delay=0.01, mechanism=mech)
    w = (np.sin(np.arange(40) * 0.4) * 1e5).astype(np.float32)
    shapes = [(0, 64), (1, 9), (2, 8), (3, 13), (4, 48), (2, 31), (4, 10), (2, 48)]
    outs = {}
    for on in (True, False):
        with _palette(on):
            for v in shapes:
                p = build_propagator(m, spec, engine=engine)
                assert (p._core.media.get("midx") is not None) == on
                p.set_source(src)
                p._batches["step"] = dev.Batch([p._core], variant=v)
                for s in range(40):
                    p.step(float(w[s]))
                torch.cuda.synchronize()
                outs[on, v] = p._core.state[p._core.cur].clone()
    ref = outs[False, (0, 64)]
    assert float(ref.abs().max()) > 0.0
    for key, o in outs.items():
        assert torch.equal(o, ref), key


@pytest.mark.parametrize("kv,ring", [((2, 19), 16), ((1, 10), 0), ((4, 48), 32)])
def test_palette_migration_equals_float_media(kv, ring):
    """Two shots on a layered 90 x 170 model, PP and PS, normalized divcurl
    imaging inside the step (extraction, receiver ring or direct
    correlation), with and without the palette."""
    from ertm.config import GpuConfig, ImagingConfig, SolverConfig
    from ertm.imaging import RTMImager
    from ertm.solver import build_solver
    from ertm.types import ReceiverSpec, SourceSpec

    rng = np.random.default_rng(31)
    m = _layered_model(90, 170, rng, dx=5.0, dz=5.0)
    # the cross term starts at (delay + 1.5 / f0) / dt = step 200
    solver = SolverConfig(nt=400, dt=4e-4, order=8, boundary="free", engine="elastic", npml=12)
    recv = ReceiverSpec(x=np.arange(0, 170, 2) * 5.0, z=np.full(85, 5.0))
    shots = [build_solver(solver, gpu_config=GpuConfig(strict_cfl=False)).forward(
        m, SourceSpec(x=x, z=10.0, f0=25.0, delay=0.02), recv, components=("vx", "vz"))
        for x in (200.0, 610.0)]
    imaging = ImagingConfig(mode="pp", condition="normalized", separation="divcurl",
                            checkpoint_mb=1.0)
    outs = {}
    for on in (True, False):
        with _palette(on):
            gpu = GpuConfig(strict_cfl=False, kernel_variant=kv, image_ring=ring)
            outs[on] = RTMImager(imaging, solver, gpu_config=gpu).migrate_batch(
                shots, m, modes=("pp", "ps"))
    for a, b in zip(outs[True], outs[False]):
        for mode in ("pp", "ps"):
            assert np.abs(b[mode].image).max() > 0
            assert np.array_equal(a[mode].image, b[mode].image), mode


def test_acoustic_palette_migration_equals_float_media():
    """Acoustic zero_lag migration of a shot on a layered model with the
    replay extraction and the receiver ring inside the step."""
    from ertm.config import GpuConfig, ImagingConfig, SolverConfig
    from ertm.imaging import RTMImager
    from ertm.solver import build_solver
    from ertm.types import ReceiverSpec, SourceSpec

    rng = np.random.default_rng(47)
    m = _layered_model(110, 150, rng, dx=5.0, dz=5.0)
    solver = SolverConfig(nt=400, dt=4e-4, order=8, boundary="cpml", engine="acoustic", npml=12)
    recv = ReceiverSpec(x=np.arange(0, 150, 2) * 5.0, z=np.full(75, 10.0))
    shot = build_solver(solver, gpu_config=GpuConfig(strict_cfl=False)).forward(
        m, SourceSpec(x=370.0, z=10.0, f0=25.0, delay=0.02), recv)
    imaging = ImagingConfig(mode="pp", condition="zero_lag", checkpoint_mb=1.0)
    outs = {}
    for on in (True, False):
        with _palette(on):
            gpu = GpuConfig(strict_cfl=False, kernel_variant=(2, 19), image_ring=16)
            outs[on] = RTMImager(imaging, solver, gpu_config=gpu).migrate(shot, m)
    assert np.abs(outs[False].image).max() > 0
    assert np.array_equal(outs[True].image, outs[False].image)
