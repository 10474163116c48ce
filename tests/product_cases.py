"""Turn golden / oracle test cases into the product engine's objects."""

from __future__ import annotations

import numpy as np


def model_of(c):
    from ertm.types import ElasticModel, GridSpec

    return ElasticModel(
        grid=GridSpec(nx=c["nx"], nz=c["nz"], dx=c["dx"], dz=c["dz"], x0=c["x0"], z0=c["z0"]),
        vp=c["vp"], vs=c["vs"], rho=c["rho"], qp=c["qp"], qs=c["qs"])


def source_of(c):
    from ertm.types import SourceSpec

    s = c["src"]
    return SourceSpec(x=s["x"], z=s["z"], f0=s["f0"], amplitude=s["amplitude"], delay=s["delay"],
                      mechanism=s["mechanism"], kind=s.get("kind", "ricker"))


def receivers_of(c):
    from ertm.types import ReceiverSpec

    return ReceiverSpec(x=np.asarray(c["rx"]), z=np.asarray(c["rz"]))


def propagator_of(c):
    from ertm.propagation import PropagationSpec, QOptions, build_propagator

    spec = PropagationSpec(dt=c["dt"], nt=c["nt"], order=c["order"], npml=c["npml"],
                           boundary=c["boundary"])
    q = None
    if c["engine"] == "viscoacoustic":
        qq = c["q"]
        q = QOptions(enabled=True, f_ref=qq["f_ref"], n_mechanisms=qq["n"],
                     band_ratio=qq["band_ratio"], compensation=qq["compensation"],
                     alpha_c=qq["alpha_c"], amplitude_cap=qq["amplitude_cap"],
                     lowpass_interval=qq["lowpass_interval"], f_max=qq["f_max"],
                     gain_budget_db=qq["gain_budget_db"])
    p = build_propagator(model_of(c), spec, engine=c["engine"], q_options=q)
    p.set_source(source_of(c))
    p.set_receivers(receivers_of(c))
    return p, spec


def migrate_objects(c):
    """(shot, model, imaging cfg, solver cfg, q cfg) for a golden migrate case."""
    from ertm.config import ImagingConfig, QConfig, SolverConfig
    from ertm.types import ShotRecord

    model = model_of(c)
    solver = SolverConfig(nt=c["nt"], dt=c["dt"], order=c["order"], boundary=c["boundary"],
                          engine=c["engine"], npml=c["npml"])
    imaging = ImagingConfig(**c["imaging"])
    q = None
    if "q" in c:
        qq = c["q"]
        q = QConfig(enabled=qq["enabled"], model="gsls", alpha_c=qq["alpha_c"], f_ref=qq["f_ref"],
                    n_mechanisms=qq["n"], band_ratio=qq["band_ratio"],
                    amplitude_cap=qq["amplitude_cap"], lowpass_interval=qq["lowpass_interval"],
                    f_max=qq["f_max"], gain_budget_db=qq["gain_budget_db"])
    nt = c["shot_data"].shape[0]
    sdt = c.get("shot_dt", c["dt"])
    shot = ShotRecord(data=c["shot_data"], dt=sdt, t=np.arange(nt) * sdt,
                      receivers=receivers_of(c), source=source_of(c),
                      components={k: v for k, v in c["shot"].items()})
    return shot, model, imaging, solver, q


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (den if den > 0 else 1.0)
