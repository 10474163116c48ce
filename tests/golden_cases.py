"""Rebuild the golden cases written by ``tests/golden/make_golden.py``.

Shared by the oracle tests (CPU) and the GPU parity tests: each helper turns
one stored case into plain arrays and scalars.  Nothing here imports the
reference; the ``.npz`` fixtures hold everything.
"""

from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

_cache: dict = {}


def load(name: str):
    if name not in _cache:
        with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
            _cache[name] = {k: z[k] for k in z.files}
    return _cache[name]


def tags(name: str):
    return sorted({k.split("/")[0] for k in load(name) if "/" in k})


def _kind(d, tag: str) -> str:
    """Source wavelet kind of a case (cases written before kinds were stored
    are Ricker)."""
    return str(d[f"{tag}/kind"]) if f"{tag}/kind" in d else "ricker"


def forward_case(tag: str) -> dict:
    d = load("forward")
    g = d[f"{tag}/grid"]
    sp = d[f"{tag}/spec"]
    src = d[f"{tag}/source"]
    case = dict(
        tag=tag,
        engine=str(d[f"{tag}/engine"]),
        boundary=str(d[f"{tag}/boundary"]),
        nx=int(g[0]), nz=int(g[1]), dx=float(g[2]), dz=float(g[3]), x0=float(g[4]), z0=float(g[5]),
        dt=float(sp[0]), nt=int(sp[1]), order=int(sp[2]), npml=int(sp[3]),
        vp=d[f"{tag}/vp"], vs=d[f"{tag}/vs"], rho=d[f"{tag}/rho"], qp=d[f"{tag}/qp"],
        qs=d[f"{tag}/qs"],
        src=dict(x=float(src[0]), z=float(src[1]), f0=float(src[2]), amplitude=float(src[3]),
                 delay=float(src[4]), kind=_kind(d, tag), mechanism=str(d[f"{tag}/mechanism"])),
        rx=d[f"{tag}/rx"], rz=d[f"{tag}/rz"],
        components=tuple(str(c) for c in d[f"{tag}/components"]),
        rec={str(c): d[f"{tag}/rec_{c}"] for c in d[f"{tag}/components"]},
        state={k[len(tag) + 7:]: v for k, v in d.items() if k.startswith(f"{tag}/state_")},
        psi={k[len(tag) + 5:]: v for k, v in d.items() if k.startswith(f"{tag}/psi_")},
    )
    if f"{tag}/qopt" in d:
        q = d[f"{tag}/qopt"]
        case["q"] = dict(f_ref=float(q[0]), n=int(q[1]), band_ratio=float(q[2]),
                         compensation=bool(q[3]), alpha_c=float(q[4]), amplitude_cap=float(q[5]),
                         lowpass_interval=int(q[6]), f_max=None if q[7] < 0 else float(q[7]),
                         gain_budget_db=float(q[8]))
        case["visco_a"] = d[f"{tag}/visco_a"]
        case["visco_b"] = d[f"{tag}/visco_b"]
        case["visco_kr"] = d[f"{tag}/visco_kr"]
        case["visco_summary"] = d[f"{tag}/visco_summary"]
        case["visco_alpha"] = d[f"{tag}/visco_alpha"]
    return case


def migrate_case(tag: str) -> dict:
    d = load("migrate")
    g = d[f"{tag}/grid"]
    so = d[f"{tag}/solver"]
    im = [str(v) for v in d[f"{tag}/imaging"]]
    src = d[f"{tag}/source"]
    comps = tuple(str(c) for c in d[f"{tag}/components"])
    case = dict(
        tag=tag,
        engine=str(d[f"{tag}/engine"]),
        boundary=str(d[f"{tag}/boundary"]),
        nx=int(g[0]), nz=int(g[1]), dx=float(g[2]), dz=float(g[3]), x0=float(g[4]), z0=float(g[5]),
        nt=int(so[0]), dt=float(so[1]), order=int(so[2]), npml=int(so[3]),
        vp=d[f"{tag}/vp"], vs=d[f"{tag}/vs"], rho=d[f"{tag}/rho"], qp=d[f"{tag}/qp"],
        qs=d[f"{tag}/qs"],
        imaging=dict(mode=im[0], condition=im[1], separation=im[2], illumination=im[3] == "True",
                     receiver_normalization=im[4] == "True", laplacian=im[5] == "True",
                     lowpass_fraction=float(im[6]), checkpoint_mb=float(im[7]), eps=float(im[8]),
                     start_periods=float(im[9])),
        src=dict(x=float(src[0]), z=float(src[1]), f0=float(src[2]), amplitude=float(src[3]),
                 delay=float(src[4]), kind=_kind(d, tag), mechanism=str(d[f"{tag}/mechanism"])),
        rx=d[f"{tag}/rx"], rz=d[f"{tag}/rz"],
        components=comps,
        shot_dt=float(d[f"{tag}/shot_dt"]) if f"{tag}/shot_dt" in d else float(so[1]),
        shot={c: d[f"{tag}/shot_{c}"] for c in comps},
        shot_data=d[f"{tag}/shot_data"],
        image=d[f"{tag}/image"],
        checkpoints=tuple(int(v) for v in d[f"{tag}/checkpoints"]),
        start_step=int(d[f"{tag}/start_step"]),
    )
    if f"{tag}/q" in d:
        q = d[f"{tag}/q"]
        case["q"] = dict(enabled=bool(q[0]), alpha_c=float(q[1]), f_ref=float(q[2]), n=int(q[3]),
                         band_ratio=float(q[4]), amplitude_cap=float(q[5]),
                         lowpass_interval=int(q[6]), f_max=None if q[7] < 0 else float(q[7]),
                         gain_budget_db=float(q[8]))
    return case
