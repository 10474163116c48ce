"""The two-steps-per-launch elastic kernel against one step per launch.

Bar: bit-exact.  The whole state (fields and CPML memories) after a
checkpointing forward pass, and every checkpoint it saved, equal the
one-step kernels' on random odd-sized grids (every boundary and source
mechanism, sources at edges and corners, batches of uneven grids, odd
checkpoint intervals so that some steps run alone); migrations with the
two-step forward and replay passes give the golden images.
"""

from __future__ import annotations

import sys

import numpy as np
import pytest

from golden_cases import migrate_case
from product_cases import migrate_objects

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _engine():
    from ertm import _lib

    _lib.lib()


def _random_props(seed, n):
    from ertm.propagation import PropagationSpec, build_propagator
    from ertm.types import ElasticModel, GridSpec, SourceSpec

    rng = np.random.default_rng(700 + seed)
    boundary = ("cpml", "free", "none")[seed % 3]
    mech = ("explosive", "shear", "force_z")[(seed // 3) % 3]
    npml = int(rng.integers(5, 12))
    props, desc = [], []
    for k in range(n):
        nz, nx = int(rng.integers(23, 140)), int(rng.integers(19, 420))
        vp = rng.uniform(2000, 4000, (nz, nx)).astype(np.float32)
        m = ElasticModel(GridSpec(nx=nx, nz=nz, dx=5.0, dz=4.0), vp=vp,
                         vs=(vp / rng.uniform(1.6, 2.0, (nz, nx))).astype(np.float32),
                         rho=rng.uniform(1800, 2600, (nz, nx)).astype(np.float32),
                         qp=np.full_like(vp, 100.0), qs=np.full_like(vp, 50.0))
        spec = PropagationSpec(dt=3e-4, nt=41, order=8, npml=npml, boundary=boundary)
        sx = float(rng.choice([0.0, (nx - 1) * 5.0, rng.uniform(0, (nx - 1) * 5.0)]))
        sz = float(rng.choice([4.0, (nz - 1) * 4.0, rng.uniform(4.0, (nz - 1) * 4.0)]))
        p = build_propagator(m, spec, engine="elastic")
        p.set_source(SourceSpec(x=sx, z=sz, f0=30.0, delay=0.01, mechanism=mech))
        props.append(p)
        desc.append((nz, nx))
    return props, desc, boundary, mech


def _forward(props, variant, two_step, w, interval):
    """Checkpointing forward pass of ertm_run over len(w) steps; returns the
    final states and the checkpoint arenas."""
    import torch

    from ertm import device as dev

    cores = [p._core for p in props]
    stride = max(c.state_floats for c in cores)
    nck = (len(w) + interval - 1) // interval
    arenas = [torch.zeros((nck, stride), dtype=torch.float32, device="cuda") for _ in cores]
    b = dev.Batch(cores, variant=variant, two_step=two_step)
    with dev.StepTimer():
        cur = b.run(0, 0, len(w), cores[0].cur, wavelets=[w] * len(cores), ck_interval=interval,
                    ck_origin=0, save=arenas, save_stride=stride)
        torch.cuda.synchronize()
    kinds = dev.StepTimer.read_kinds()
    b.close()
    # one-step launches for the steps that end at a checkpoint, pairs for the rest
    pairs = kinds["plain2"]["launches"]
    if two_step[0]:
        singles = kinds["plain"]["launches"]
        assert pairs > 0 and singles + 2 * pairs == len(w), (pairs, singles)
    else:
        assert pairs == 0
    return [c.state[cur].clone() for c in cores], arenas


@pytest.mark.parametrize("seed", range(9))
def test_two_step_forward_equals_one_step_on_random_grids(seed):
    n = 1 + seed % 3
    w = (np.sin(np.arange(41) * 0.4) * 1e5).astype(np.float32)
    interval = (7, 4, 9)[seed % 3]
    ref = None
    for two_step in ((0, 0), (2, 16), (2, 8), (2, 13), (2, 300), (2, 37)):
        props, desc, boundary, mech = _random_props(seed, n)
        _say("two-step", seed, desc, boundary, mech, "interval", interval, two_step)
        out = _forward(props, (2, 48), two_step, w, interval)
        if ref is None:
            ref = out
            assert float(out[0][0].abs().max()) > 0.0
            continue
        for k in range(n):
            assert bool((out[0][k] == ref[0][k]).all()), (two_step, k, "state")
            assert bool((out[1][k] == ref[1][k]).all()), (two_step, k, "checkpoints")


@pytest.mark.parametrize("two_step", [(2, 16), (2, 8), (2, 48)])
@pytest.mark.parametrize("tag", ["el_pp_divcurl", "el_ps_divcurl", "el_pp_stress",
                                 "el_ps_stress", "el_sp_zl", "el_ss_exc"])
def test_migrate_with_two_step_passes_matches_reference(tag, two_step):
    """Forward and replay passes two steps per launch (the replay extracting
    the source field of both steps; excitation replay stays one step per
    launch): the golden images bit for bit."""
    _say("two-step migrate", tag, two_step)
    from ertm.config import GpuConfig
    from ertm.imaging import RTMImager

    c = migrate_case(tag)
    shot, model, imaging, solver, q = migrate_objects(c)
    from ertm import device as dev

    for gpu in (GpuConfig(kernel_variant=(2, 19), two_step=two_step),
                GpuConfig(kernel_variant=(2, 19), two_step=two_step, replay_slots=5)):
        with dev.StepTimer():
            res = RTMImager(imaging, solver, q, gpu).migrate(shot, model)
        kinds = dev.StepTimer.read_kinds()
        assert np.array_equal(res.image, c["image"]), gpu
        assert kinds["plain2"]["launches"] > 0, kinds          # forward pass
        if imaging.condition != "excitation":
            assert kinds["extract2"]["launches"] > 0, kinds    # replay pass
