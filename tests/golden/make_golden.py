"""Write the golden vectors that pin the oracle to the reference implementation.

This script is the only place that imports the reference package.  It runs in
the development container (where the reference tree is mounted read-only) and
writes small ``.npz`` fixtures next to itself.  Nothing on the GPU path reads
the reference; tests only read the fixtures this script wrote.

Every case is tiny (at most 51 x 47 padded cells, at most 300 steps) so the
fixtures stay a few hundred KB.  Run:

    python tests/golden/make_golden.py [--reference /root/reference] [--store-only]

What is pinned (reference file:line of the function whose output is stored):
  * fit_gsls tau/delta/fit_error            ertm/viscoelastic.py:132-205
  * dplus/dminus, CPMLAxis a/b              ertm/numerics.py:68-137, 143-227
  * forward_shot records + final states     ertm/propagation.py:1018-1093
    (acoustic/elastic/viscoacoustic, cpml/free, all source mechanisms)
  * viscoacoustic stabiliser (compensation) ertm/propagation.py:726-749
  * sample/inject with duplicate receivers  ertm/propagation.py:331-391, 472-494
  * record_source_wavefield + replay_block  ertm/imaging.py:113-162
  * RTMImager.migrate images                ertm/imaging.py:205-349
    (one case with a shot sampled at 36 us, aligned by _prepare_record's
    resample_sinc, ertm/imaging.py:396-415)
  * gaussian_pulse, source_wavelet           ertm/numerics.py:472-475,
                                            ertm/propagation.py:1004-1015
  * resample_sinc                           ertm/numerics.py:337-392
  * cfl_dt                                  ertm/numerics.py:314-326
  * NumpyDataStore model/shot/image files   ertm/data.py:84-282
    (written by the reference's own store into store/; the raw inputs go
    beside them in store/inputs.npz, written with plain numpy)
"""

from __future__ import annotations

import argparse
import shutil
import sys
import tempfile
from dataclasses import replace
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def _import_reference(root: str):
    if root not in sys.path:
        sys.path.insert(0, root)
    import ertm.numerics as num  # noqa: F401
    from ertm import imaging, numerics, propagation, viscoelastic
    from ertm.config import ImagingConfig, QConfig, SolverConfig
    from ertm.data import NumpyDataStore
    from ertm.types import ElasticModel, GridSpec, ImageResult, ReceiverSpec, ShotRecord, SourceSpec

    return dict(
        NumpyDataStore=NumpyDataStore,
        ImageResult=ImageResult,
        ShotRecord=ShotRecord,
        imaging=imaging,
        numerics=numerics,
        propagation=propagation,
        viscoelastic=viscoelastic,
        ImagingConfig=ImagingConfig,
        QConfig=QConfig,
        SolverConfig=SolverConfig,
        ElasticModel=ElasticModel,
        GridSpec=GridSpec,
        ReceiverSpec=ReceiverSpec,
        SourceSpec=SourceSpec,
    )


# ---------------------------------------------------------------------------
# model / geometry builders (plain numpy; identical arrays are rebuilt by the
# tests from the stored inputs, so nothing here needs to be re-derived)
# ---------------------------------------------------------------------------
def layered(nz, nx, dz, z_tops, vps, ratio=1.8, qps=None):
    z = np.arange(nz) * dz
    vp = np.zeros((nz, nx), np.float32)
    qp = np.full((nz, nx), 100.0, np.float32)
    for k, top in enumerate(z_tops):
        vp[z >= top, :] = vps[k]
        if qps is not None:
            qp[z >= top, :] = qps[k]
    vs = (vp / ratio).astype(np.float32)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    qs = (qp * 0.5).astype(np.float32)
    return vp, vs, rho, qp, qs


def random_model(rng, nz, nx):
    vp = rng.uniform(2000.0, 3500.0, size=(nz, nx)).astype(np.float32)
    vs = (vp / rng.uniform(1.6, 2.2, size=(nz, nx))).astype(np.float32)
    rho = rng.uniform(1800.0, 2600.0, size=(nz, nx)).astype(np.float32)
    qp = rng.choice(np.array([25.0, 60.0, 140.0], np.float32), size=(nz, nx)).astype(np.float32)
    qs = (qp * 0.5).astype(np.float32)
    return vp, vs, rho, qp, qs


def main() -> None:
    parser = argparse.ArgumentParser()
    parser.add_argument("--reference", default="/root/reference")
    parser.add_argument("--store-only", action="store_true",
                        help="write only the NumpyDataStore files under store/")
    args = parser.parse_args()
    R = _import_reference(args.reference)
    if args.store_only:
        write_store_files(R)
        return
    num = R["numerics"]
    prop_mod = R["propagation"]
    img_mod = R["imaging"]
    GridSpec, ElasticModel = R["GridSpec"], R["ElasticModel"]
    ReceiverSpec, SourceSpec = R["ReceiverSpec"], R["SourceSpec"]
    SolverConfig, ImagingConfig, QConfig = R["SolverConfig"], R["ImagingConfig"], R["QConfig"]

    def model_of(arrays, nx, nz, dx, dz, x0=0.0, z0=0.0):
        vp, vs, rho, qp, qs = arrays
        return ElasticModel(
            grid=GridSpec(nx=nx, nz=nz, dx=dx, dz=dz, x0=x0, z0=z0),
            vp=vp, vs=vs, rho=rho, qp=qp, qs=qs,
        )

    # ------------------------------------------------------------ numerics
    rng = np.random.default_rng(20240)
    out = {}
    field = rng.standard_normal((23, 19)).astype(np.float32)
    out["deriv_field"] = field
    for order in (2, 4, 6, 8):
        coeffs = num.staggered_coefficients(order)
        for axis in (0, 1):
            out[f"dplus_o{order}_a{axis}"] = num.dplus(field, axis, coeffs, 3.7)
            out[f"dminus_o{order}_a{axis}"] = num.dminus(field, axis, coeffs, 3.7)
    cp_cases = [(51, 5.0, 8, 5e-4, 3500.0, 25.0, (True, True)),
                (41, 7.0, 20, 3e-4, 2800.0, 25.0, (False, True)),
                (300, 2.5, 20, 2e-4, 5500.0, 25.0, (True, True))]
    for k, (n, h, npml, dt, vmax, f0, sides) in enumerate(cp_cases):
        axis = num.CPMLAxis(n, h, npml, dt, vmax, axis=1, alpha_max=2.0 * np.pi * f0, sides=sides)
        out[f"cpml{k}_a"] = axis.a.ravel()
        out[f"cpml{k}_b"] = axis.b.ravel()
        out[f"cpml{k}_params"] = np.array([n, h, npml, dt, vmax, f0, sides[0], sides[1]], np.float64)
    # source wavelets: Gaussian pulse and Ricker through source_wavelet
    t = np.arange(400, dtype=np.float64) * 2.5e-4
    for k, (f0, t0) in enumerate([(25.0, 0.048), (40.0, 0.03), (7.5, 0.0)]):
        out[f"gauss{k}_params"] = np.array([f0, t0], np.float64)
        out[f"gauss{k}"] = num.gaussian_pulse(t, f0, t0)
    for k, (kind, f0, delay, amp, dt, nt) in enumerate([
            ("gaussian", 30.0, 0.04, 2.5, 4e-4, 300), ("gaussian", 50.0, 0.01, 1.0, 2.5e-4, 150),
            ("ricker", 30.0, 0.04, 2.5, 4e-4, 300)]):
        srcw = R["SourceSpec"](x=0.0, z=0.0, f0=f0, delay=delay, amplitude=amp, kind=kind)
        out[f"wavelet{k}_params"] = np.array([f0, delay, amp, dt, nt], np.float64)
        out[f"wavelet{k}_kind"] = np.array(kind)
        out[f"wavelet{k}"] = prop_mod.source_wavelet(srcw, dt, nt)
    # sinc resampling: C3's 36 us records to the 0.35 ms solver step and
    # back, the reference test's 0.4 ms -> 0.1 ms, a non-divisible pair, and
    # the three-sample minimum
    rngS = np.random.default_rng(36)
    for k, (dt_in, dt_out, n_in, ntr) in enumerate([
            (3.6e-5, 3.5e-4, 1500, 5), (3.5e-4, 3.6e-5, 120, 3), (4.0e-4, 1.0e-4, 500, 1),
            (2.5e-4, 3.3e-4, 257, 4), (1.0e-3, 4.0e-4, 3, 2)]):
        tr = np.cumsum(rngS.standard_normal((n_in, ntr)), axis=0).astype(np.float32)
        out[f"sinc{k}_params"] = np.array([dt_in, dt_out], np.float64)
        out[f"sinc{k}_in"] = tr
        out[f"sinc{k}_out"] = num.resample_sinc(tr, dt_in, dt_out)
        out[f"sinc{k}_in1"] = tr[:, 0].astype(np.float64)
        out[f"sinc{k}_out1"] = num.resample_sinc(tr[:, 0].astype(np.float64), dt_in, dt_out)
    # CFL limits at the workloads' media and steps (and a few orders)
    cfl = []
    for vmax, dx, dz, order, safety in [(3500.0, 5.0, 5.0, 8, 0.55), (4500.0, 5.0, 5.0, 8, 0.55),
                                        (5000.0, 4.0, 4.0, 8, 0.55), (5500.0, 2.5, 2.5, 8, 0.55),
                                        (4000.0, 10.0, 10.0, 8, 0.55), (3000.0, 5.0, 7.0, 4, 0.5),
                                        (2500.0, 10.0, 10.0, 2, 1.0), (4500.0, 5.0, 5.0, 6, 1.0)]:
        cfl.append([vmax, dx, dz, order, safety, num.cfl_dt(vmax, dx, dz, order, safety=safety)])
    out["cfl_table"] = np.array(cfl, np.float64)
    np.savez_compressed(HERE / "numerics.npz", **out)

    # ------------------------------------------------------------ GSLS fit
    visco = R["viscoelastic"]
    rows = []
    for q in (15.0, 25.0, 30.0, 40.0, 60.0, 120.0, 200.0):
        for f_ref in (20.0, 25.0, 30.0):
            for nmech in (1, 3):
                params = visco.fit_gsls(q, f_ref, n_mechanisms=nmech, band_ratio=1.5)
                tau = np.zeros(3)
                dlt = np.zeros(3)
                tau[: len(params.tau_sigma)] = params.tau_sigma
                dlt[: len(params.delta)] = params.delta
                rows.append(np.concatenate([[q, f_ref, nmech, params.fit_error], tau, dlt]))
    np.savez_compressed(HERE / "gsls.npz", table=np.array(rows))

    # ------------------------------------------------------------ forward shots
    PropagationSpec = prop_mod.PropagationSpec
    QOptions = prop_mod.QOptions
    fwd = {}

    def run_case(tag, engine, model, spec, source, receivers, components, q_options=None):
        p = prop_mod.build_propagator(model, spec, engine=engine, q_options=q_options)
        data = prop_mod.run_forward(
            model, source, receivers, spec, engine=engine, components=components,
            propagator=p, q_options=q_options,
        )
        for name in components:
            fwd[f"{tag}/rec_{name}"] = data[name]
        for name in p._field_names():
            fwd[f"{tag}/state_{name}"] = getattr(p, name).copy()
        for key, value in p.pml._memory.items():
            fwd[f"{tag}/psi_{key}"] = value.copy()
        if engine == "viscoacoustic":
            fwd[f"{tag}/visco_a"] = p.a
            fwd[f"{tag}/visco_b"] = p.b
            fwd[f"{tag}/visco_kr"] = p.k_relaxed
            fwd[f"{tag}/visco_summary"] = np.array(
                [p.quality_summary["clipped"], p.quality_summary["lowpass_steps"]], np.int64
            )
            fwd[f"{tag}/visco_alpha"] = np.array(p.quality_summary["alpha_mechanisms"])
        m = model
        fwd[f"{tag}/vp"], fwd[f"{tag}/vs"], fwd[f"{tag}/rho"] = m.vp, m.vs, m.rho
        fwd[f"{tag}/qp"], fwd[f"{tag}/qs"] = m.qp, m.qs
        g = m.grid
        fwd[f"{tag}/grid"] = np.array([g.nx, g.nz, g.dx, g.dz, g.x0, g.z0], np.float64)
        fwd[f"{tag}/spec"] = np.array([spec.dt, spec.nt, spec.order, spec.npml], np.float64)
        fwd[f"{tag}/boundary"] = np.array(spec.boundary)
        fwd[f"{tag}/engine"] = np.array(engine)
        fwd[f"{tag}/source"] = np.array(
            [source.x, source.z, source.f0, source.amplitude, source.delay], np.float64
        )
        fwd[f"{tag}/mechanism"] = np.array(source.mechanism)
        fwd[f"{tag}/kind"] = np.array(source.kind)
        fwd[f"{tag}/rx"], fwd[f"{tag}/rz"] = receivers.x, receivers.z
        fwd[f"{tag}/components"] = np.array(list(components))
        if q_options is not None:
            fwd[f"{tag}/qopt"] = np.array([
                q_options.f_ref, q_options.n_mechanisms, q_options.band_ratio,
                float(q_options.compensation), q_options.alpha_c, q_options.amplitude_cap,
                q_options.lowpass_interval, -1.0 if q_options.f_max is None else q_options.f_max,
                q_options.gain_budget_db,
            ], np.float64)

    # A: acoustic, free surface, layered (the reference's tiny self-check geometry)
    arrs = layered(31, 41, 10.0, [0.0, 150.0], [2400.0, 3000.0])
    mA = model_of(arrs, 41, 31, 10.0, 10.0)
    spA = PropagationSpec(dt=2.5e-4, nt=150, order=8, npml=6, boundary="free")
    srcA = SourceSpec(x=200.0, z=20.0, f0=50.0, delay=0.01)
    recA = ReceiverSpec(x=np.linspace(0.0, 400.0, 9), z=np.full(9, 20.0))
    run_case("ac_free", "acoustic", mA, spA, srcA, recA, ("p", "vx", "vz"))

    # B: acoustic, cpml, random medium, anisotropic spacing, off-grid geometry
    rngB = np.random.default_rng(7)
    arrs = random_model(rngB, 29, 37)
    mB = model_of(arrs, 37, 29, 5.0, 7.0, x0=-3.0, z0=2.0)
    spB = PropagationSpec(dt=3e-4, nt=120, order=8, npml=7, boundary="cpml")
    srcB = SourceSpec(x=89.5, z=97.0, f0=45.0, delay=0.02, amplitude=2.5)
    rx = np.array([10.0, 10.0, 33.3, 71.25, 140.0, 140.0, -20.0, 177.0, 55.5])
    rz = np.array([20.0, 20.0, 5.5, 60.0, 90.1, 90.1, 30.0, 199.0, 150.25])
    recB = ReceiverSpec(x=rx, z=rz)
    run_case("ac_cpml", "acoustic", mB, spB, srcB, recB, ("p", "vx", "vz", "sxz"))

    # B4: same medium, 4th order, boundary none
    spB4 = PropagationSpec(dt=3e-4, nt=80, order=4, npml=7, boundary="none")
    run_case("ac_o4_none", "acoustic", mB, spB4, srcB, recB, ("p",))

    # C: elastic, free surface, explosive, layered
    arrs = layered(31, 41, 10.0, [0.0, 120.0, 220.0], [2400.0, 3000.0, 3600.0])
    mC = model_of(arrs, 41, 31, 10.0, 10.0)
    spC = PropagationSpec(dt=2.5e-4, nt=150, order=8, npml=6, boundary="free")
    run_case("el_free", "elastic", mC, spC, srcA, recA, ("p", "vx", "vz", "sxz"))

    # C2: elastic, cpml, random medium, shear source, anisotropic spacing
    rngC = np.random.default_rng(11)
    arrs = random_model(rngC, 29, 37)
    mC2 = model_of(arrs, 37, 29, 5.0, 7.0, x0=-3.0, z0=2.0)
    spC2 = PropagationSpec(dt=2e-4, nt=110, order=8, npml=7, boundary="cpml")
    srcC2 = SourceSpec(x=89.5, z=97.0, f0=45.0, delay=0.02, amplitude=1e3, mechanism="shear")
    run_case("el_shear", "elastic", mC2, spC2, srcC2, recB, ("vx", "vz", "sxz", "p", "sxx", "szz"))

    # C3: elastic, free, force_z source, 6th order
    spC3 = PropagationSpec(dt=2e-4, nt=90, order=6, npml=6, boundary="free")
    srcC3 = SourceSpec(x=151.0, z=35.0, f0=40.0, delay=0.02, amplitude=1.0, mechanism="force_z")
    run_case("el_force", "elastic", mC, spC3, srcC3, recA, ("vx", "vz"))

    # C4: elastic, free, force_z source, 8th order (the fused kernels' order)
    spC4 = PropagationSpec(dt=2e-4, nt=120, order=8, npml=6, boundary="free")
    run_case("el_force8", "elastic", mC, spC4, srcC3, recA, ("vx", "vz", "sxz"))

    # B5: acoustic, cpml, random medium, Gaussian source pulse
    srcG = SourceSpec(x=89.5, z=97.0, f0=45.0, delay=0.02, amplitude=2.5, kind="gaussian")
    run_case("ac_gauss", "acoustic", mB, spB, srcG, recB, ("p", "vx"))

    # D: viscoacoustic, forward (no compensation), layered Q
    arrs = layered(31, 41, 10.0, [0.0, 150.0], [2400.0, 3000.0], qps=[30.0, 80.0])
    mD = model_of(arrs, 41, 31, 10.0, 10.0)
    spD = PropagationSpec(dt=2.5e-4, nt=150, order=8, npml=6, boundary="cpml")
    qD = QOptions(enabled=True, f_ref=30.0, n_mechanisms=3)
    run_case("vi_fwd", "viscoacoustic", mD, spD, srcA, recA, ("p", "vx", "vz"), qD)

    # D2: viscoacoustic with compensation (stabiliser: clip + low-pass every 10 steps)
    qD2 = QOptions(enabled=True, f_ref=30.0, n_mechanisms=3, compensation=True,
                   alpha_c=1.0, amplitude_cap=1.5, lowpass_interval=10, f_max=60.0,
                   gain_budget_db=40.0)
    spD2 = PropagationSpec(dt=2.5e-4, nt=100, order=8, npml=6, boundary="free")
    run_case("vi_comp", "viscoacoustic", mD, spD2, srcA, recA, ("p", "vx", "vz"), qD2)

    np.savez_compressed(HERE / "forward.npz", **fwd)

    # ------------------------------------------------------------ inject
    inj = {}
    for tag, engine, model, spec in (
        ("ac", "acoustic", mB, spB),
        ("el", "elastic", mC2, spC2),
        ("vi", "viscoacoustic", mD, spD),
    ):
        p = prop_mod.build_propagator(model, spec, engine=engine,
                                      q_options=qD if engine == "viscoacoustic" else None)
        p.set_source(srcB)
        recv = recB if engine != "viscoacoustic" else ReceiverSpec(
            x=np.array([10.0, 10.0, 33.3, 71.25, 140.0, 140.0, 277.0, 395.0]),
            z=np.array([20.0, 20.0, 5.5, 60.0, 90.1, 90.1, 30.0, 260.0]))
        p.set_receivers(recv)
        names = p._field_names()
        rngI = np.random.default_rng(3)
        for name in names:
            arr = getattr(p, name)
            arr[...] = rngI.standard_normal(arr.shape).astype(arr.dtype)
            inj[f"{tag}/before_{name}"] = arr.copy()
        comps = {"acoustic": ("p", "vx", "vz"), "elastic": ("vx", "vz", "sxz", "p", "sxx"),
                 "viscoacoustic": ("p", "vx", "vz")}[engine]
        vals = {}
        for c in comps:
            vals[c] = rngI.standard_normal(recv.nrec).astype(np.float32)
            inj[f"{tag}/values_{c}"] = vals[c]
            inj[f"{tag}/sample_{c}"] = p.sample((c,))[c]
        for c in comps:
            p.inject(c, vals[c])
        for name in names:
            inj[f"{tag}/after_{name}"] = getattr(p, name).copy()
        inj[f"{tag}/rx"], inj[f"{tag}/rz"] = recv.x, recv.z
        inj[f"{tag}/components"] = np.array(list(comps))
    np.savez_compressed(HERE / "inject.npz", **inj)

    # ------------------------------------------------------------ checkpoint/replay
    rep = {}
    prop = prop_mod.build_propagator(mB, spB, engine="acoustic")
    prop.set_source(srcB)
    wave = prop_mod.source_wavelet(srcB, spB.dt, spB.nt)
    store = img_mod.record_source_wavefield(prop, wave, budget_mb=1.0)
    rep["interval"] = np.array(store.interval)
    rep["starts"] = np.array([s for s, _ in store.states])
    for b in (0, 3, len(store.states) - 1):
        fields = img_mod.replay_block(prop, wave, store, b, "stress")
        rep[f"block{b}"] = np.stack([f["P"] for f in fields])
    np.savez_compressed(HERE / "replay.npz", **rep)

    # ------------------------------------------------------------ migrate
    mig = {}

    def migrate_case(tag, engine, model, solver_cfg, imaging_cfg, source, receivers, comps,
                     q_cfg=None, fwd_dt=None):
        spec_fwd = PropagationSpec.from_config(solver_cfg, model, npml=solver_cfg.npml)
        if fwd_dt is not None:  # shot sampled at another step than the solver's
            duration = (solver_cfg.nt - 1) * solver_cfg.dt
            spec_fwd = replace(spec_fwd, dt=fwd_dt, nt=int(np.floor(duration / fwd_dt)) + 2)
        qo = None
        if engine == "viscoacoustic":
            qo = QOptions(enabled=True, f_ref=q_cfg.f_ref, n_mechanisms=q_cfg.n_mechanisms,
                          band_ratio=q_cfg.band_ratio)
        shot = prop_mod.forward_shot(model, source, receivers, spec_fwd, engine=engine,
                                     components=comps, q_options=qo)
        imager = img_mod.RTMImager(imaging_cfg, solver_cfg, q_cfg)
        result = imager.migrate(shot, model)
        mig[f"{tag}/image"] = result.image
        mig[f"{tag}/checkpoints"] = np.array(
            [result.meta["checkpoints"]["interval"], result.meta["checkpoints"]["n_states"]])
        mig[f"{tag}/start_step"] = np.array(result.meta["start_step"])
        for name, arr in shot.components.items():
            mig[f"{tag}/shot_{name}"] = arr
        mig[f"{tag}/shot_data"] = shot.data
        g = model.grid
        mig[f"{tag}/grid"] = np.array([g.nx, g.nz, g.dx, g.dz, g.x0, g.z0], np.float64)
        mig[f"{tag}/vp"], mig[f"{tag}/vs"], mig[f"{tag}/rho"] = model.vp, model.vs, model.rho
        mig[f"{tag}/qp"], mig[f"{tag}/qs"] = model.qp, model.qs
        mig[f"{tag}/solver"] = np.array([solver_cfg.nt, solver_cfg.dt, solver_cfg.order,
                                         solver_cfg.npml], np.float64)
        mig[f"{tag}/boundary"] = np.array(solver_cfg.boundary)
        mig[f"{tag}/engine"] = np.array(engine)
        mig[f"{tag}/imaging"] = np.array([
            imaging_cfg.mode, imaging_cfg.condition, imaging_cfg.separation,
            str(imaging_cfg.illumination), str(imaging_cfg.receiver_normalization),
            str(imaging_cfg.laplacian), repr(imaging_cfg.lowpass_fraction),
            repr(imaging_cfg.checkpoint_mb), repr(imaging_cfg.eps), repr(imaging_cfg.start_periods),
        ])
        mig[f"{tag}/source"] = np.array(
            [source.x, source.z, source.f0, source.amplitude, source.delay], np.float64)
        mig[f"{tag}/mechanism"] = np.array(source.mechanism)
        mig[f"{tag}/kind"] = np.array(source.kind)
        mig[f"{tag}/shot_dt"] = np.array(float(shot.dt))
        mig[f"{tag}/rx"], mig[f"{tag}/rz"] = receivers.x, receivers.z
        mig[f"{tag}/components"] = np.array(list(comps))
        if q_cfg is not None:
            mig[f"{tag}/q"] = np.array([
                float(q_cfg.enabled), q_cfg.alpha_c, q_cfg.f_ref, q_cfg.n_mechanisms,
                q_cfg.band_ratio, q_cfg.amplitude_cap, q_cfg.lowpass_interval,
                -1.0 if q_cfg.f_max is None else q_cfg.f_max, q_cfg.gain_budget_db,
            ])

    solA = SolverConfig(nt=300, dt=2.5e-4, order=8, boundary="free", engine="acoustic", npml=6)
    recM = ReceiverSpec(x=np.linspace(0.0, 400.0, 21), z=np.full(21, 20.0))
    srcM = SourceSpec(x=200.0, z=20.0, f0=50.0, delay=0.01)
    arrs = layered(31, 41, 10.0, [0.0, 150.0], [2400.0, 3000.0])
    mM = model_of(arrs, 41, 31, 10.0, 10.0)
    for cond in ("zero_lag", "normalized", "excitation"):
        migrate_case(f"ac_{cond}", "acoustic", mM, solA,
                     ImagingConfig(condition=cond, checkpoint_mb=1.0), srcM, recM, ("p",))
    migrate_case("ac_raw_rn", "acoustic", mM, solA,
                 ImagingConfig(condition="normalized", checkpoint_mb=1.0, laplacian=False,
                               lowpass_fraction=1.0, receiver_normalization=True),
                 srcM, recM, ("p",))
    # the shot sampled at 36 us: _prepare_record aligns it to the solver's
    # 0.25 ms with resample_sinc before the backward pass
    migrate_case("ac_resample", "acoustic", mM, solA,
                 ImagingConfig(condition="zero_lag", checkpoint_mb=1.0), srcM, recM, ("p",),
                 fwd_dt=3.6e-5)
    solAc = replace(solA, boundary="cpml")
    migrate_case("ac_cpml_zl", "acoustic", mM, solAc,
                 ImagingConfig(condition="zero_lag", checkpoint_mb=1.0, laplacian=False,
                               lowpass_fraction=1.0), srcM, recM, ("p",))

    arrs = layered(31, 41, 10.0, [0.0, 120.0, 220.0], [2400.0, 3000.0, 3600.0])
    mE = model_of(arrs, 41, 31, 10.0, 10.0)
    solE = SolverConfig(nt=300, dt=2.5e-4, order=8, boundary="free", engine="elastic", npml=6)
    for mode in ("pp", "ps"):
        for sep in ("divcurl", "stress"):
            migrate_case(f"el_{mode}_{sep}", "elastic", mE, solE,
                         ImagingConfig(mode=mode, condition="normalized", separation=sep,
                                       checkpoint_mb=1.0), srcM, recM, ("vx", "vz"))
    migrate_case("el_sp_zl", "elastic", mE, solE,
                 ImagingConfig(mode="sp", condition="zero_lag", separation="divcurl",
                               checkpoint_mb=1.0, laplacian=False, lowpass_fraction=1.0),
                 srcM, recM, ("p", "vx", "vz"))
    migrate_case("el_ss_exc", "elastic", mE, solE,
                 ImagingConfig(mode="ss", condition="excitation", separation="stress",
                               checkpoint_mb=1.0), srcM, recM, ("vx", "vz"))

    arrs = layered(31, 41, 10.0, [0.0, 150.0], [2400.0, 3000.0], qps=[30.0, 80.0])
    mV = model_of(arrs, 41, 31, 10.0, 10.0)
    solV = SolverConfig(nt=300, dt=2.5e-4, order=8, boundary="free", engine="viscoacoustic",
                        npml=6)
    qV = QConfig(enabled=True, model="gsls", alpha_c=1.0, f_ref=30.0, amplitude_cap=50.0,
                 lowpass_interval=20, f_max=60.0, gain_budget_db=40.0)
    migrate_case("vi_norm", "viscoacoustic", mV, solV,
                 ImagingConfig(condition="normalized", checkpoint_mb=1.0), srcM, recM, ("p",), qV)
    qV0 = replace(qV, enabled=False)
    migrate_case("vi_noq", "viscoacoustic", mV, solV,
                 ImagingConfig(condition="zero_lag", checkpoint_mb=1.0, laplacian=False,
                               lowpass_fraction=1.0), srcM, recM, ("p", "vx", "vz"), qV0)
    np.savez_compressed(HERE / "migrate.npz", **mig)
    write_store_files(R)
    print("golden vectors written to", HERE)


def write_store_files(R) -> None:
    """A model, a three-component shot and an image saved by the reference's
    NumpyDataStore (ertm/data.py:84-114, 149-184, 231-253), copied to
    store/{model,shot,image}.npz; the inputs they were made from go to
    store/inputs.npz so the loader test compares against data the reference
    did not write."""
    rng = np.random.default_rng(2024)
    GridSpec = R["GridSpec"]
    grid = GridSpec(nx=9, nz=7, dx=5.0, dz=4.0, x0=10.0, z0=2.0)
    vp = rng.uniform(2000, 4000, (7, 9)).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = rng.uniform(1800, 2600, (7, 9)).astype(np.float32)
    qp = rng.uniform(30, 200, (7, 9)).astype(np.float32)
    qs = (qp * 0.5).astype(np.float32)
    model = R["ElasticModel"](grid=grid, vp=vp, vs=vs, rho=rho, qp=qp, qs=qs, name="layers",
                              meta={"builder": "golden", "seed": 2024})
    nt, nrec, dt = 12, 5, 2.5e-4
    t = np.arange(nt) * dt
    rx = np.linspace(10.0, 50.0, nrec)
    rz = np.full(nrec, 6.0)
    comps = {c: rng.standard_normal((nt, nrec)).astype(np.float32) for c in ("p", "vx", "vz")}
    rec = R["ReceiverSpec"](x=rx, z=rz, component="vz")
    src = R["SourceSpec"](x=30.0, z=8.0, f0=25.0, kind="ricker", amplitude=2.0, delay=0.04,
                          mechanism="force_z")
    shot = R["ShotRecord"](data=comps["vz"], dt=dt, t=t, receivers=rec, source=src,
                           meta={"shot_index": 3}, components=comps)
    image = rng.standard_normal((7, 9)).astype(np.float32)
    img = R["ImageResult"](image=image, grid=grid, mode="ps", meta={"condition": "normalized"})
    out = HERE / "store"
    out.mkdir(exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        store = R["NumpyDataStore"](tmp)
        for kind, path in (("model", store.save_model(model, "m")),
                           ("shot", store.save_shot(shot, "s")),
                           ("image", store.save_image(img, "i"))):
            shutil.copyfile(path, out / f"{kind}.npz")
    np.savez_compressed(out / "inputs.npz", vp=vp, vs=vs, rho=rho, qp=qp, qs=qs, t=t, rx=rx,
                        rz=rz, image=image, **{f"comp_{c}": a for c, a in comps.items()})


if __name__ == "__main__":
    main()
