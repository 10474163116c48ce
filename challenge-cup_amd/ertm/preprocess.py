"""Record preparation before migration (ertm/preprocess.py).

The reference's four data-shaping steps, with the same arithmetic (float64,
results cast to float32 where the reference casts):

* ``rotate_components`` / ``rotate_shot``: rotate (vx, vz) by the receiver
  azimuth into the source-receiver plane (ertm/preprocess.py:36-74);
* ``resample_shot``: Kaiser-windowed sinc resampling to the solver's dt
  (ertm/preprocess.py:79-109);
* ``mute_direct_wave``: cosine-tapered top mute at the direct-arrival time
  delay + margin/f0 + factor*|x - xs|/v_ref (ertm/preprocess.py:121-169);
* ``taper_record_ends``: cosine taper of the first and last samples
  (ertm/preprocess.py:172-189);
* ``prepare_shot`` chains them (ertm/preprocess.py:195-224).

``prepare_records_on_device`` applies the mute, the end taper and the
rotation to records already in HBM (one kernel per record, float64
arithmetic in the same order), for callers that keep records on the device.
"""

from __future__ import annotations

import logging
from dataclasses import replace
from typing import Sequence

import numpy as np

from . import numerics as num
from .exceptions import DataError
from .types import ElasticModel, ShotRecord

logger = logging.getLogger("ertm.preprocess")


def rotate_components(components: dict[str, np.ndarray], angle_deg: float) -> dict[str, np.ndarray]:
    """vx' = cos*vx + sin*vz, vz' = -sin*vx + cos*vz (float64, cast to float32);
    other components are copied (ertm/preprocess.py:36-63)."""
    if not components:
        raise DataError("component rotation needs at least one component")
    angle = np.deg2rad(float(angle_deg))
    cos, sin = float(np.cos(angle)), float(np.sin(angle))
    out = {name: np.array(arr, copy=True) for name, arr in components.items()}
    if "vx" in out and "vz" in out:
        vx = np.asarray(components["vx"], dtype=np.float64)
        vz = np.asarray(components["vz"], dtype=np.float64)
        out["vx"] = (cos * vx + sin * vz).astype(np.float32)
        out["vz"] = (-sin * vx + cos * vz).astype(np.float32)
    else:
        logger.warning("vx/vz missing: rotation skipped")
    return out


def rotate_shot(shot: ShotRecord, angle_deg: float) -> ShotRecord:
    """ertm/preprocess.py:66-74."""
    if angle_deg == 0.0:
        return shot
    comps = rotate_components(shot.components, angle_deg)
    primary = "p" if "p" in comps else ("vz" if "vz" in comps else None)
    data = comps[primary] if primary else shot.data
    return replace(shot, data=data, components=comps,
                   meta={**shot.meta, "rotation_deg": angle_deg})


def resample_shot(shot: ShotRecord, dt_target: float) -> ShotRecord:
    """Sinc-resample every component to ``dt_target`` (ertm/preprocess.py:79-109)."""
    if abs(dt_target - shot.dt) < 1e-12:
        return shot
    if dt_target <= 0:
        raise DataError(f"dt_target must be positive, got {dt_target}")
    comps = dict(shot.components) or {"primary": shot.data}
    conv = {name: num.resample_sinc(np.asarray(arr, dtype=np.float64), shot.dt, dt_target)
            for name, arr in comps.items()}
    primary = "primary" if "primary" in conv else next(iter(conv))
    nt = conv[primary].shape[0]
    new_comps = {} if "primary" in conv else {n: a.astype(np.float32) for n, a in conv.items()}
    return replace(shot, data=conv[primary].astype(np.float32), dt=float(dt_target),
                   t=np.arange(nt, dtype=np.float64) * dt_target, components=new_comps,
                   meta={**shot.meta, "input_dt": float(shot.dt),
                         "resampled_dt": float(dt_target), "resampled": True})


def required_dt(model: ElasticModel, dx: float | None = None, *, order: int = 8) -> float:
    """CFL-safe dt with safety 0.5 (ertm/preprocess.py:112-115)."""
    dx = model.grid.dx if dx is None else dx
    return num.cfl_dt(float(np.max(model.vp)), dx, model.grid.dz, order, safety=0.5)


def mute_weights(shot: ShotRecord, v_reference: float, *, margin_periods: float = 1.2,
                 offset_factor: float = 1.05, taper_periods: float = 0.5) -> np.ndarray:
    """(nt, nrec) float64 mute weights: 0.5 - 0.5 cos(pi * ramp), ramp rising
    linearly over +-taper around t_mute = delay + margin/f0 + factor*offset/v
    (ertm/preprocess.py:144-156)."""
    if v_reference <= 0:
        raise DataError(f"v_reference must be positive, got {v_reference}")
    t = np.asarray(shot.t, dtype=np.float64)
    offset = np.abs(shot.receivers.x - shot.source.x)
    period = 1.0 / max(float(shot.source.f0), 1e-6)
    t_mute = float(shot.source.delay) + margin_periods * period + offset_factor * offset / v_reference
    taper = max(float(taper_periods) * period, 1e-9)
    nt = np.asarray(shot.data).shape[0]
    w = np.empty((nt, t_mute.size), dtype=np.float64)
    for i in range(t_mute.size):
        start, end = t_mute[i] - taper, t_mute[i] + taper
        ramp = np.clip((t - start) / (end - start), 0.0, 1.0)
        w[:, i] = 0.5 - 0.5 * np.cos(np.pi * ramp)
    return w


def mute_direct_wave(shot: ShotRecord, v_reference: float, *, margin_periods: float = 1.2,
                     offset_factor: float = 1.05, taper_periods: float = 0.5) -> ShotRecord:
    """Top mute of the direct arrival (ertm/preprocess.py:121-169)."""
    w = mute_weights(shot, v_reference, margin_periods=margin_periods,
                     offset_factor=offset_factor, taper_periods=taper_periods)
    data = np.array(shot.data, dtype=np.float64, copy=True) * w
    comps = {n: (np.asarray(a, dtype=np.float64) * w).astype(np.float32)
             for n, a in shot.components.items()}
    primary = comps.get("p", data.astype(np.float32)) if comps else data.astype(np.float32)
    return replace(shot, data=primary, components=comps,
                   meta={**shot.meta, "direct_mute": True, "mute_velocity": float(v_reference)})


def end_taper(n: int, fraction: float) -> tuple[int, np.ndarray]:
    """(w, ramp): the first w samples are scaled by ramp, the last w by ramp
    reversed (ertm/preprocess.py:177-184)."""
    w = max(int(fraction * n), 1)
    return w, 0.5 - 0.5 * np.cos(np.pi * (np.arange(w) + 0.5) / w)


def taper_record_ends(shot: ShotRecord, fraction: float = 0.05) -> ShotRecord:
    """Cosine taper of the record ends (ertm/preprocess.py:172-189)."""
    if not 0 < fraction < 0.5:
        raise DataError(f"fraction must lie in (0, 0.5), got {fraction}")

    def taper(arr):
        out = np.array(arr, dtype=np.float64, copy=True)
        w, ramp = end_taper(out.shape[0], fraction)
        out[:w] *= ramp[:, None]
        out[-w:] *= ramp[::-1][:, None]
        return out

    data = taper(shot.data)
    comps = {n: taper(a).astype(np.float32) for n, a in shot.components.items()}
    primary = comps.get("p", data.astype(np.float32)) if comps else data.astype(np.float32)
    return replace(shot, data=primary, components=comps,
                   meta={**shot.meta, "taper_fraction": fraction})


def prepare_shot(shot: ShotRecord, *, model: ElasticModel, record_dt: float | None = None,
                 rotation_deg: float = 0.0, mute: bool = True, v_reference: float | None = None,
                 taper_fraction: float = 0.05) -> ShotRecord:
    """rotate -> resample -> mute -> end taper (ertm/preprocess.py:195-224)."""
    out = shot
    if rotation_deg:
        out = rotate_shot(out, rotation_deg)
    if record_dt is not None and abs(record_dt - out.dt) > 1e-12:
        out = resample_shot(out, record_dt)
    if mute:
        v = float(v_reference) if v_reference is not None else float(np.percentile(model.vp, 25))
        out = mute_direct_wave(out, v)
    if taper_fraction > 0:
        out = taper_record_ends(out, taper_fraction)
    return out


def build_shot_sequence(base_shot: ShotRecord, source_positions: Sequence[float]) -> list[ShotRecord]:
    """Copies of a shot at other source x (ertm/preprocess.py:227-241)."""
    return [replace(base_shot, source=replace(base_shot.source, x=float(x)),
                    meta={**base_shot.meta, "shot_x": float(x)}) for x in source_positions]


def resample_record_on_device(record: np.ndarray, dt_in: float, dt_out: float, nt: int):
    """``resample_sinc`` (ertm/numerics.py:337-392) of a (n_in, nrec) float32
    record onto ``dt_out``, then cut or zero-padded to ``nt`` rows as
    ``RTMImager._prepare_record`` does (ertm/imaging.py:396-423): a (nt, nrec)
    float32 device tensor.  The weights come from ``numerics.sinc_taps`` on
    the host; the GPU sums the taps in order in float64, the order NumPy's
    einsum takes for two or more traces, so the result equals the host
    path's.  The record is complete when the call returns."""
    import ctypes as C

    import torch

    from . import _lib as L
    from . import device as dev

    rec = np.ascontiguousarray(record, dtype=np.float32)
    if rec.ndim != 2:
        raise DataError("resample_record_on_device takes a (nt, nrec) record")
    n_in, nrec = rec.shape
    w, idx = num.sinc_taps(n_in, float(dt_in), float(dt_out))
    n_out, taps = w.shape
    if idx.max(initial=0) >= 2 ** 31:
        raise DataError("record too long to resample on the device")
    src = dev.to_device(rec)
    wd = torch.from_numpy(np.ascontiguousarray(w)).to("cuda")
    od = torch.from_numpy(np.ascontiguousarray(idx, dtype=np.int32)).to("cuda")
    out = torch.empty((int(nt), nrec), dtype=torch.float32, device="cuda")
    L.check(L.lib().ertm_resample_record(C.c_void_p(L.ptr(src)), n_in, nrec,
                                         C.c_void_p(L.ptr(wd)), C.c_void_p(L.ptr(od)), n_out,
                                         taps, C.c_void_p(L.ptr(out)), int(nt),
                                         L.stream_handle()), "ertm_resample_record")
    # complete on return, like an uploaded record: callers (the serving
    # loop) may read it from other streams
    torch.cuda.current_stream().synchronize()
    return out


def prepare_records_on_device(records: dict, shot: ShotRecord, *, model: ElasticModel,
                              rotation_deg: float = 0.0, mute: bool = True,
                              v_reference: float | None = None,
                              taper_fraction: float = 0.05) -> dict:
    """``prepare_shot`` without resampling, on records already in HBM:
    ``records`` maps component name -> (nt, nrec) float32 device tensor of the
    shot (its own dt and length), prepared in place and returned.  Rotation,
    then mute, then end taper, each in float64 with float32 results as the
    reference; cosines come from the device's float64 cos, so a sample may
    differ from the host path in the last float32 bit."""
    import ctypes as C

    import torch

    from . import _lib as L

    lib = L.lib()
    stream = L.stream_handle()
    if rotation_deg and "vx" in records and "vz" in records:
        angle = np.deg2rad(float(rotation_deg))
        vx, vz = records["vx"], records["vz"]
        L.check(lib.ertm_rotate_components(C.c_void_p(L.ptr(vx)), C.c_void_p(L.ptr(vz)),
                                           vx.numel(), float(np.cos(angle)),
                                           float(np.sin(angle)), stream),
                "ertm_rotate_components")
    t_dev = t_mute = None
    taper = 0.0
    if mute:
        v = float(v_reference) if v_reference is not None else float(np.percentile(model.vp, 25))
        if v <= 0:
            raise DataError(f"v_reference must be positive, got {v}")
        period = 1.0 / max(float(shot.source.f0), 1e-6)
        offset = np.abs(shot.receivers.x - shot.source.x)
        tm = float(shot.source.delay) + 1.2 * period + 1.05 * offset / v
        taper = max(0.5 * period, 1e-9)
        t_dev = torch.from_numpy(np.asarray(shot.t, dtype=np.float64)).to("cuda")
        t_mute = torch.from_numpy(np.ascontiguousarray(tm, dtype=np.float64)).to("cuda")
    for name, rec in records.items():
        nt, nrec = rec.shape
        w_end = 0
        if taper_fraction > 0:
            if not 0 < taper_fraction < 0.5:
                raise DataError(f"fraction must lie in (0, 0.5), got {taper_fraction}")
            w_end = end_taper(nt, taper_fraction)[0]
        L.check(lib.ertm_prepare_record(C.c_void_p(L.ptr(rec)), nt, nrec,
                                        C.c_void_p(L.ptr(t_dev)), C.c_void_p(L.ptr(t_mute)),
                                        float(taper), int(w_end), stream),
                "ertm_prepare_record")
    return records

