"""Oracle RTM driver: forward shot, checkpoints, replay, imaging, post-processing.

Test infrastructure only.  Restates ertm/propagation.py:1004-1093 and
ertm/imaging.py:61-502 on the oracle engines, keeping every reference quirk:

* checkpoint interval from fields-only bytes (CPML memories are created on
  the first step), ertm/imaging.py:113-137;
* replay loads the state saved *after* step s and steps w[s] again,
  ertm/imaging.py:140-162;
* migration uses the adaptive npml, ertm/propagation.py:84-87 via
  ertm/imaging.py:224;
* float32 products accumulated in float64, illumination over all steps, the
  cross term only from ``start_step``, ertm/imaging.py:263-331.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import engines as oeng
from . import gsls as ogsls
from . import numerics as onum

MODES = {"pp": ("P", "P"), "ps": ("P", "S"), "sp": ("S", "P"), "ss": ("S", "S")}


def wavelet(kind, f0, delay, amplitude, dt, nt):
    """ertm/propagation.py:1004-1015."""
    t = np.arange(nt, dtype=np.float64) * dt
    w = onum.ricker(t, f0, delay) if kind == "ricker" else onum.gaussian(t, f0, delay)
    return (float(amplitude) * w).astype(np.float32)


def adaptive_npml(nz, nx):
    """ertm/propagation.py:84-87."""
    return int(min(20, max(8, min(nx, nz) // 8)))


@dataclass
class Model:
    """Plain model arrays on the physical grid (ertm/types.py:85-122)."""

    nz: int
    nx: int
    dx: float
    dz: float
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    qp: np.ndarray
    x0: float = 0.0
    z0: float = 0.0


@dataclass
class QSpec:
    """Viscoacoustic options (ertm/propagation.py:130-159)."""

    f_ref: float = 30.0
    n: int = 3
    band_ratio: float = 1.5
    compensation: bool = False
    alpha_c: float = 1.0
    amplitude_cap: float = 50.0
    lowpass_interval: int = 20
    f_max: float | None = None
    gain_budget_db: float = 40.0


def build(engine, m: Model, dt, nt, order=8, npml=20, boundary="cpml", dtype="float32",
          q: QSpec | None = None):
    """ertm/propagation.py:986-1001."""
    g = oeng.Grid(m.nz, m.nx, m.dx, m.dz, m.x0, m.z0, order)
    if engine == "acoustic":
        return oeng.Acoustic(g, m.vp, m.rho, dt, npml, boundary, dtype)
    if engine == "elastic":
        return oeng.Elastic(g, m.vp, m.vs, m.rho, dt, npml, boundary, dtype)
    if engine == "viscoacoustic":
        q = q or QSpec()
        qpad = np.pad(np.asarray(m.qp).astype(dtype, copy=False), g.pad, mode="edge")
        rel = ogsls.relaxation(qpad, q.f_ref, int(q.n), q.band_ratio, dt, nt, q.compensation,
                               q.alpha_c, q.f_max, q.gain_budget_db)
        return oeng.Visco(g, m.vp, m.rho, m.qp, dt, rel, npml, boundary, dtype,
                          q.compensation, q.amplitude_cap, q.lowpass_interval)
    raise ValueError(engine)


def forward(prop, w, comps):
    """ertm/propagation.py:1018-1052: step then sample, every step."""
    nt = len(w)
    rec = {c: np.zeros((nt, prop.rx.size), np.float32) for c in comps}
    for s in range(nt):
        prop.step(float(w[s]))
        smp = prop.sample(comps)
        for c in comps:
            rec[c][s] = smp[c]
    return rec


@dataclass
class Store:
    """ertm/imaging.py:93-110."""

    interval: int
    per_state: int
    states: list = field(default_factory=list)


def record(prop, w, budget_mb):
    """ertm/imaging.py:113-137."""
    nt = len(w)
    per = max(int(prop.state_bytes()), 1)
    budget = max(float(budget_mb), 1.0) * 1024.0 ** 2
    interval = max(1, int(np.ceil(nt / max(1, int(budget // per)))))
    st = Store(interval, per)
    for s in range(nt):
        prop.step(float(w[s]))
        if s % interval == 0:
            st.states.append((s, prop.state()))
    return st


def block_bounds(st: Store, b, nt):
    start = st.states[b][0]
    end = st.states[b + 1][0] if b + 1 < len(st.states) else nt
    return start, end


def replay(prop, w, st: Store, b, sep):
    """ertm/imaging.py:140-162 (image-field copies for the whole block)."""
    start, end = block_bounds(st, b, len(w))
    prop.load_state(st.states[b][1])
    out = []
    for s in range(start, end):
        prop.step(float(w[s]))
        out.append({k: np.array(v, copy=True) for k, v in prop.image_fields(sep).items()})
    return out


def inject_record(back, engine, rec, s):
    """ertm/imaging.py:425-448 (routing of record names, zero rows skipped)."""
    for name, data in rec.items():
        if name in ("primary", "p") and engine == "acoustic":
            tgt = "p"
        elif name == "primary":
            tgt = "vz"
        else:
            tgt = name
        v = np.asarray(data[s], np.float32)
        if not np.any(v):
            continue
        back.inject(tgt, v)


@dataclass
class Imaging:
    """ImagingConfig fields used by migrate (ertm/config.py:310-351)."""

    mode: str = "pp"
    condition: str = "normalized"
    separation: str = "divcurl"
    illumination: bool = True
    receiver_normalization: bool = False
    laplacian: bool = True
    lowpass_fraction: float = 0.85
    checkpoint_mb: float = 256.0
    eps: float = 1.0e-10
    start_periods: float = 1.5


def start_step(delay, f0, start_periods, dt, nt):
    """ertm/imaging.py:266-276."""
    s = int(np.floor((float(delay) + float(start_periods) / max(float(f0), 1e-6)) / dt))
    return int(np.clip(s, 0, nt))


def align_record(rec, rec_dt, dt):
    """ertm/imaging.py:396-415 without the caller's preprocess: each
    component in float64, sinc-resampled to ``dt`` when its step differs by
    more than 1e-12 s (and it has more than one sample)."""
    out = {}
    for k, v in rec.items():
        d = np.asarray(v, np.float64)
        if abs(float(rec_dt) - float(dt)) > 1e-12 and d.shape[0] > 1:
            d = onum.resample_sinc(d, float(rec_dt), float(dt))
        out[k] = d
    return out


def migrate_raw(engine, m: Model, dt, nt, order, boundary, cfg: Imaging, rec, src, rx, rz,
                w=None, q: QSpec | None = None, q_enabled=False, dtype="float32", rec_dt=None):
    """ertm/imaging.py:205-331 up to (and including) normalisation.

    ``src`` = dict(x, z, f0, delay, amplitude, kind, mechanism); ``rec`` maps
    component name -> (nt_rec, nrec) array sampled at ``rec_dt`` (default:
    ``dt``), aligned here to (nt, dt).
    Returns (image float64 before post-processing, info dict).
    """
    if rec_dt is not None:
        rec = align_record(rec, rec_dt, dt)
    npml = adaptive_npml(m.nz, m.nx)
    if w is None:
        w = wavelet(src.get("kind", "ricker"), src["f0"], src["delay"], src["amplitude"], dt, nt)
    w = np.asarray(w, np.float32)
    qf = qb = None
    if engine == "viscoacoustic":
        base = q or QSpec()
        qf = QSpec(**{**base.__dict__, "compensation": False})
        qb = QSpec(**{**base.__dict__, "compensation": bool(q_enabled)})
    aligned = {}
    for k, v in rec.items():
        d = np.asarray(v, np.float64)
        if d.shape[0] >= nt:
            d = d[:nt]
        else:
            pad = np.zeros((nt, d.shape[1]))
            pad[: d.shape[0]] = d
            d = pad
        aligned[k] = d
    fwd = build(engine, m, dt, nt, order, npml, boundary, dtype=dtype, q=qf)
    fwd.set_source(src["x"], src["z"], src.get("mechanism", "explosive"))
    fwd.set_receivers(rx, rz)
    st = record(fwd, w, cfg.checkpoint_mb)
    back = build(engine, m, dt, nt, order, npml, boundary, dtype=dtype, q=qb)
    back.set_source(src["x"], src["z"], src.get("mechanism", "explosive"))
    back.set_receivers(rx, rz)
    sw, rw = MODES[cfg.mode]
    s0 = start_step(src["delay"], src["f0"], cfg.start_periods, dt, nt)
    num = np.zeros((m.nz, m.nx))
    ill = np.zeros((m.nz, m.nx))
    rill = np.zeros((m.nz, m.nx))
    cond = cfg.condition
    if cond == "excitation":
        # ertm/imaging.py:450-474: forward replay of every block for the peak maps
        pk_step = np.full((m.nz, m.nx), -1, np.int32)
        pk_amp = np.zeros((m.nz, m.nx))
        run = np.zeros((m.nz, m.nx))
        for b in range(len(st.states)):
            start, _ = block_bounds(st, b, nt)
            for off, f in enumerate(replay(fwd, w, st, b, cfg.separation)):
                amp = np.abs(np.asarray(f["P"]))
                better = amp > run
                if not np.any(better):
                    continue
                run = np.where(better, amp, run)
                pk_amp = np.where(better, amp, pk_amp)
                pk_step = np.where(better, start + off, pk_step)
        for b in range(len(st.states) - 1, -1, -1):
            start, end = block_bounds(st, b, nt)
            for s in range(end - 1, start - 1, -1):
                inject_record(back, engine, aligned, s)
                back.step(0.0)
                if s < s0:
                    continue
                mask = pk_step == s
                if np.any(mask):
                    r = back.image_fields(cfg.separation)[rw]
                    num += np.where(mask, pk_amp * np.asarray(r), 0.0)
                    rill += np.where(mask, np.asarray(r) ** 2, 0.0)
        ill = pk_amp ** 2
    else:
        for b in range(len(st.states) - 1, -1, -1):
            fields = replay(fwd, w, st, b, cfg.separation)
            start = st.states[b][0]
            for off in range(len(fields) - 1, -1, -1):
                s = start + off
                inject_record(back, engine, aligned, s)
                back.step(0.0)
                sv = np.asarray(fields[off][sw])
                rv = np.asarray(back.image_fields(cfg.separation)[rw])
                if cond == "normalized" or cfg.receiver_normalization:
                    ill += sv * sv
                    rill += rv * rv
                if s < s0:
                    continue
                num += sv * rv
    img = num
    if cond == "normalized" and cfg.illumination:
        img = img / (ill + float(cfg.eps))
    if cfg.receiver_normalization:
        img = img / (rill + float(cfg.eps))
    info = dict(interval=st.interval, n_states=len(st.states), start_step=s0, npml=npml,
                num=num, ill=ill, rill=rill, back=back, fwd=fwd)
    return np.asarray(img, np.float64), info


def postprocess(img, m: Model, npml, boundary, cfg: Imaging):
    """ertm/imaging.py:476-502."""
    d = np.asarray(img, np.float64)
    if cfg.laplacian:
        d = onum.laplacian(d, m.dx, m.dz)
    if cfg.lowpass_fraction < 1.0:
        d = onum.lowpass(d, m.dx, m.dz, cfg.lowpass_fraction)
    mask = np.ones((m.nz, m.nx))
    mask[:, :npml] = 0.0
    mask[:, -npml:] = 0.0
    mask[-npml:, :] = 0.0
    if boundary == "cpml":
        mask[:npml, :] = 0.0
    return d * mask


def migrate(engine, m, dt, nt, order, boundary, cfg, rec, src, rx, rz, w=None, q=None,
            q_enabled=False, dtype="float32", rec_dt=None):
    """Full single-shot migration; returns the float32 image (ImageResult.image).
    dtype="float64" runs the propagation in float64 (the reference's
    PropagationSpec.dtype), the accuracy reference for the float32 engine."""
    img, info = migrate_raw(engine, m, dt, nt, order, boundary, cfg, rec, src, rx, rz, w, q,
                            q_enabled, dtype, rec_dt)
    out = postprocess(img, m, info["npml"], boundary, cfg)
    return np.asarray(out, np.float32), info


def stack(images, normalize=True):
    """ertm/pipeline.py:133-161 (per-shot max-abs normalised mean)."""
    if len(images) == 1:
        return np.asarray(images[0])
    acc = np.zeros(np.asarray(images[0]).shape)
    for im in images:
        d = np.asarray(im, np.float64)
        if normalize:
            sc = float(np.max(np.abs(d)))
            if sc > 0:
                d = d / sc
        acc += d
    acc /= len(images)
    return acc.astype(np.float32)


def cell_updates(nz, nx, nt, passes=3):
    """Work unit: physical cells advanced one step per propagator pass."""
    return int(passes) * int(nt) * int(nz) * int(nx)


def ceil_div(a, b):
    return -(-a // b)


def checkpoint_interval(nt, per_state_bytes, budget_mb):
    """Interval rule of ertm/imaging.py:121-124."""
    per = max(int(per_state_bytes), 1)
    budget = max(float(budget_mb), 1.0) * 1024.0 ** 2
    return max(1, int(math.ceil(nt / max(1, int(budget // per)))))
