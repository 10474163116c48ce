"""Seeded synthetic migration requests (small grids) for serving tests.

Each request follows the serving workload's rules at test scale: random
layered model with 2-4 layers and up to 2 elliptical anomalies, dt from the
CFL rule with safety 0.5, a surface source at a random x, receivers every two
cells at z = 2 dz, zero_lag imaging, and a seeded random record.  nt is long
enough for the imaging window (start step = (delay + 1.5 / f0) / dt) to
hold about 150 steps."""

from __future__ import annotations

import numpy as np


def make_request(seed, engine=None, nz=None, nx=None, nt=None, amplitude=1e-6, rid=None):
    from ertm import numerics as num
    from ertm.config import ImagingConfig, SolverConfig
    from ertm.serving import ShotRequest
    from ertm.types import ElasticModel, GridSpec, ReceiverSpec, ShotRecord, SourceSpec

    rng = np.random.default_rng(seed)
    engine = engine or ("acoustic" if rng.random() < 0.5 else "elastic")
    nz = nz or int(rng.integers(40, 72))
    nx = nx or int(rng.integers(48, 96))
    nt = nt or int(rng.integers(380, 460))
    dx = 10.0
    z = np.arange(nz)[:, None] * dx * np.ones((1, nx))
    x = np.ones((nz, 1)) * np.arange(nx)[None, :] * dx
    nl = int(rng.integers(2, 5))
    tops = np.sort(rng.uniform(0, nz * dx, nl - 1))
    vps = np.sort(rng.uniform(1800, 4000, nl))
    vp = np.full((nz, nx), vps[0])
    for t, v in zip(tops, vps[1:]):
        vp[z >= t] = v
    for _ in range(int(rng.integers(0, 3))):
        cx, cz = rng.uniform(0, nx * dx), rng.uniform(0, nz * dx)
        rx, rz = rng.uniform(20, 80, 2)
        vp[((x - cx) / rx) ** 2 + ((z - cz) / rz) ** 2 <= 1.0] = rng.uniform(1800, 4000)
    vp = vp.astype(np.float32)
    model = ElasticModel(GridSpec(nx=nx, nz=nz, dx=dx, dz=dx), vp=vp,
                         vs=(vp / 1.8).astype(np.float32),
                         rho=(310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32),
                         qp=np.full_like(vp, 100.0), qs=np.full_like(vp, 50.0))
    dt = float(num.cfl_dt(float(vp.max()), dx, dx, 8, 0.5))
    f0 = float(np.clip(float(vp.min()) / (12.5 * dx), 5.0, 40.0))
    src = SourceSpec(x=float(rng.uniform(0.2, 0.8) * nx * dx), z=2 * dx, f0=f0, delay=1.2 / f0)
    rxs = np.arange(0, nx, 2) * dx
    rec = ReceiverSpec(x=rxs, z=np.full(rxs.size, 2 * dx))
    comps = ("p",) if engine == "acoustic" else ("vx", "vz")
    data = {c: (rng.standard_normal((nt, rxs.size)) * amplitude).astype(np.float32) for c in comps}
    shot = ShotRecord(data=data[comps[0]], dt=dt, t=np.arange(nt) * dt, receivers=rec,
                      source=src, components=data)
    solver = SolverConfig(nt=nt, dt=dt, order=8, boundary="cpml", engine=engine, npml=10)
    imaging = ImagingConfig(condition="zero_lag", checkpoint_mb=1.0)
    return ShotRequest(shot=shot, model=model, solver=solver, imaging=imaging, rid=rid)
