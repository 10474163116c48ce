"""SEG-Y input and output (ertm/segy.py:57-554) without segyio.

The reference reads and writes SEG-Y through segyio, which is not part of
this environment; this module implements the SEG-Y rev 1 layout directly
with NumPy, vectorised over whole files:

* file = 3200-byte textual header (EBCDIC) + 400-byte binary header
  + optional 3200-byte extended textual headers + traces;
* trace = 240-byte header + samples, all big-endian;
* binary-header fields used: sample interval (bytes 3217-3218, microseconds),
  samples per trace (3221-3222), sample format (3225-3226), trace sorting
  (3229-3230), revision (3501-3502), fixed-length flag (3503-3504),
  extended header count (3505-3506);
* trace-header fields used (1-based bytes): TRACE_SEQUENCE_LINE 1-4,
  TRACE_SEQUENCE_FILE 5-8, FieldRecord 9-12, TraceNumber 13-16, CDP 21-24,
  offset 37-40, SourceGroupScalar 71-72, SourceX 73-76, SourceY 77-80,
  GroupX 81-84, GroupY 85-88, TRACE_SAMPLE_COUNT 115-116,
  TRACE_SAMPLE_INTERVAL 117-118;
* sample formats read: 1 (IBM float), 2 (int32), 3 (int16), 5 (IEEE float),
  8 (int8); written: 1 and 5.  IBM values convert exactly to float32 (their
  <= 24 significant bits fit); float32 to IBM rounds the hexadecimal
  fraction to nearest, ties to even.

Reading rules follow the reference: the sample interval comes from the
binary header (a 0 there reads as 1000 us); coordinates are scaled by
``SourceGroupScalar`` (positive multiplies, negative divides) or by
``SegYReadOptions.coordinate_scalar``; traces are grouped into shots by
source x, by FieldRecord or per file, and sorted by receiver x within a
shot.  The writers put the same header fields the reference's writers put,
so files written by either side read back on the other.
"""

from __future__ import annotations

import glob as _glob
import logging
from dataclasses import dataclass
from pathlib import Path
from typing import Any, Sequence

import numpy as np

from .exceptions import DataError
from .types import ImageResult, ReceiverSpec, ShotRecord, SourceSpec

logger = logging.getLogger("ertm.segy")

TEXT_BYTES, BIN_BYTES, TRACE_HEADER_BYTES = 3200, 400, 240

# binary header: name -> (byte offset from the file start, big-endian dtype)
BIN_FIELDS = {
    "Interval": (3216, ">u2"),
    "Samples": (3220, ">u2"),
    "Format": (3224, ">i2"),
    "SortingCode": (3228, ">i2"),
    "SEGYRevision": (3500, ">u2"),
    "FixedLengthTrace": (3502, ">i2"),
    "ExtendedHeaders": (3504, ">i2"),
}

# trace header: name -> (byte offset inside the 240-byte header, dtype)
TRACE_FIELDS = {
    "TRACE_SEQUENCE_LINE": (0, ">i4"),
    "TRACE_SEQUENCE_FILE": (4, ">i4"),
    "FieldRecord": (8, ">i4"),
    "TraceNumber": (12, ">i4"),
    "CDP": (20, ">i4"),
    "offset": (36, ">i4"),
    "SourceGroupScalar": (70, ">i2"),
    "SourceX": (72, ">i4"),
    "SourceY": (76, ">i4"),
    "GroupX": (80, ">i4"),
    "GroupY": (84, ">i4"),
    "TRACE_SAMPLE_COUNT": (114, ">u2"),
    "TRACE_SAMPLE_INTERVAL": (116, ">u2"),
}

SAMPLE_DTYPES = {1: ">u4", 2: ">i4", 3: ">i2", 5: ">f4", 8: "i1"}


# ----------------------------------------------------------------------------
# IBM hexadecimal floating point
# ----------------------------------------------------------------------------
def ibm_to_float32(words: np.ndarray) -> np.ndarray:
    """IBM single-precision words (uint32) -> float32:
    (-1)^s * 0.F * 16^(E - 64) with a 7-bit exponent E and 24-bit fraction F."""
    w = np.asarray(words, dtype=np.uint32)
    sign = np.where(w >> 31, -1.0, 1.0)
    expo = ((w >> 24) & 0x7F).astype(np.int64)
    frac = (w & 0x00FFFFFF).astype(np.float64)
    return (sign * np.ldexp(frac, 4 * (expo - 64) - 24)).astype(np.float32)


def float32_to_ibm(values: np.ndarray) -> np.ndarray:
    """float32 -> IBM single-precision words (uint32), fraction rounded to
    nearest (ties to even); magnitudes beyond IBM's range saturate, below it
    flush to zero; NaN/Inf are rejected."""
    a = np.asarray(values, dtype=np.float64)
    if not np.all(np.isfinite(a)):
        raise DataError("IBM float cannot hold NaN/Inf")
    sign = (a < 0).astype(np.uint32) << 31
    mag = np.abs(a)
    m, e = np.frexp(mag)                          # mag = m * 2^e, m in [0.5, 1)
    e16 = -((-e) // 4)                            # ceil(e / 4): mag < 16^e16
    frac = np.rint(np.ldexp(m, 24 + e - 4 * e16))  # 0.F * 2^24, exact before rounding
    carry = frac >= 2.0 ** 24
    frac = np.where(carry, 2.0 ** 20, frac)
    e16 = e16 + carry
    biased = e16 + 64
    over, under = biased > 127, (biased < 0) | (mag == 0)
    frac = np.where(over, 2.0 ** 24 - 1, np.where(under, 0.0, frac))
    biased = np.clip(biased, 0, 127)
    biased = np.where(under, 0, biased)
    return sign | (biased.astype(np.uint32) << 24) | frac.astype(np.uint32)


# ----------------------------------------------------------------------------
# file access
# ----------------------------------------------------------------------------
@dataclass
class _SegyFile:
    ns: int
    interval: int
    fmt: int
    headers: np.ndarray   # structured, one row per trace (the TRACE_FIELDS)
    samples: np.ndarray   # (n_trace, ns) raw samples in the file's dtype


def _trace_dtype(ns: int, fmt: int) -> np.dtype:
    names = list(TRACE_FIELDS) + ["samples"]
    formats = [TRACE_FIELDS[n][1] for n in TRACE_FIELDS] + [(SAMPLE_DTYPES[fmt], (ns,))]
    offsets = [TRACE_FIELDS[n][0] for n in TRACE_FIELDS] + [TRACE_HEADER_BYTES]
    size = np.dtype(SAMPLE_DTYPES[fmt]).itemsize
    return np.dtype({"names": names, "formats": formats, "offsets": offsets,
                     "itemsize": TRACE_HEADER_BYTES + ns * size})


def _bin_field(raw: np.ndarray, name: str) -> int:
    off, dt = BIN_FIELDS[name]
    return int(np.frombuffer(raw[off:off + np.dtype(dt).itemsize].tobytes(), dtype=dt)[0])


def _open(path: Path, *, headers_only: bool = False) -> _SegyFile:
    try:
        raw = np.memmap(path, dtype=np.uint8, mode="r")
    except (OSError, ValueError) as exc:
        raise DataError(f"cannot read SEG-Y {path}: {exc}") from exc
    if raw.size < TEXT_BYTES + BIN_BYTES:
        raise DataError(f"{path}: shorter than the SEG-Y file headers")
    ns = _bin_field(raw, "Samples")
    interval = _bin_field(raw, "Interval") or 1000
    fmt = _bin_field(raw, "Format")
    if fmt not in SAMPLE_DTYPES:
        raise DataError(f"{path}: sample format {fmt} not supported (1, 2, 3, 5, 8)")
    n_ext = max(_bin_field(raw, "ExtendedHeaders"), 0)
    start = TEXT_BYTES + BIN_BYTES + TEXT_BYTES * n_ext
    dtype = _trace_dtype(ns, fmt)
    body = raw.size - start
    if ns <= 0 or body < 0 or body % dtype.itemsize:
        raise DataError(f"{path}: {body} trace bytes are not a whole number of "
                        f"{dtype.itemsize}-byte traces (ns={ns}); variable-length traces "
                        "are not supported")
    traces = np.ndarray((body // dtype.itemsize,), dtype=dtype, buffer=raw, offset=start)
    counts = traces["TRACE_SAMPLE_COUNT"]
    if np.any((counts != 0) & (counts != ns)):
        raise DataError(f"{path}: trace sample counts differ from the binary header's {ns}")
    heads = np.empty(traces.shape, dtype=[(n, TRACE_FIELDS[n][1]) for n in TRACE_FIELDS])
    for n in TRACE_FIELDS:
        heads[n] = traces[n]
    samples = np.empty((0, ns)) if headers_only else np.array(traces["samples"])
    return _SegyFile(ns, interval, fmt, heads, samples)


def _float_samples(f: _SegyFile) -> np.ndarray:
    """(ns, n_trace) float64 samples."""
    s = f.samples
    data = ibm_to_float32(s) if f.fmt == 1 else s
    return np.asarray(data, dtype=np.float64).T


def _apply_scalar(values: np.ndarray, scalar: np.ndarray | int | None) -> np.ndarray:
    """SEG-Y coordinate scalar: positive multiplies, negative divides, 0 = 1
    (ertm/segy.py:93-97)."""
    v = np.asarray(values, dtype=np.float64)
    s = np.broadcast_to(np.asarray(0 if scalar is None else scalar, dtype=np.float64), v.shape)
    return np.where(s > 0, v * np.abs(s), np.where(s < 0, v / np.where(s < 0, np.abs(s), 1.0), v))


def _write(path: Path, text: str, ns: int, interval: int, fmt: int,
           heads: dict[str, np.ndarray], samples: np.ndarray) -> Path:
    """Write a fixed-length SEG-Y rev 1 file; ``samples`` is (n_trace, ns)."""
    if fmt not in (1, 5):
        raise DataError(f"writing supports sample formats 1 (IBM) and 5 (IEEE), got {fmt}")
    if not 0 < ns < 65536 or not 0 < interval < 65536:
        raise DataError(f"samples ({ns}) and interval ({interval} us) must fit 16 bits")
    n = samples.shape[0]
    lines = [text[i:i + 80] for i in range(0, len(text), 80)][:40] or [""]
    card = "".join(f"C{k + 1:2d} {ln}"[:80].ljust(80) for k, ln in enumerate(lines))
    head = np.zeros(TEXT_BYTES + BIN_BYTES, dtype=np.uint8)
    head[:TEXT_BYTES] = np.frombuffer(card.ljust(TEXT_BYTES).encode("cp500"), dtype=np.uint8)
    for name, value in (("Interval", interval), ("Samples", ns), ("Format", fmt),
                        ("SortingCode", 1), ("SEGYRevision", 0x0100), ("FixedLengthTrace", 1),
                        ("ExtendedHeaders", 0)):
        off, dt = BIN_FIELDS[name]
        head[off:off + np.dtype(dt).itemsize] = np.frombuffer(
            np.array(value, dtype=dt).tobytes(), dtype=np.uint8)
    traces = np.zeros((n,), dtype=_trace_dtype(ns, fmt))
    traces["TRACE_SEQUENCE_LINE"] = np.arange(1, n + 1)
    traces["TRACE_SEQUENCE_FILE"] = np.arange(1, n + 1)
    traces["TRACE_SAMPLE_COUNT"] = ns
    traces["TRACE_SAMPLE_INTERVAL"] = interval
    for name, values in heads.items():
        traces[name] = values
    data = np.asarray(samples, dtype=np.float32)
    traces["samples"] = float32_to_ibm(data) if fmt == 1 else data
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        with open(path, "wb") as fh:
            fh.write(head.tobytes())
            fh.write(traces.tobytes())
    except OSError as exc:
        raise DataError(f"cannot write SEG-Y {path}: {exc}") from exc
    return path


# ----------------------------------------------------------------------------
# reading
# ----------------------------------------------------------------------------
@dataclass(frozen=True)
class SegYReadOptions:
    """Reading options (ertm/segy.py:64-90)."""

    component: str = "p"
    group_by: str = "source"  # source | field_record | file
    receiver_depth: float = 0.0
    source_depth: float = 0.0
    receiver_depth_from: str = "constant"  # constant | group_y
    coordinate_scalar: int | None = None
    shot_limit: int = 0
    sort_by_offset: bool = True
    f0: float = 20.0
    delay: float = 0.0

    def validate(self) -> None:
        if self.group_by not in {"source", "field_record", "file"}:
            raise DataError(f"group_by must be source/field_record/file, got {self.group_by!r}")
        if self.receiver_depth_from not in {"constant", "group_y"}:
            raise DataError("receiver_depth_from must be constant/group_y, "
                            f"got {self.receiver_depth_from!r}")
        if self.shot_limit < 0:
            raise DataError("shot_limit must not be negative")


def expand_segy_paths(path: str | Path) -> list[Path]:
    """File, directory (its .sgy/.segy files) or glob pattern -> sorted file
    list (ertm/segy.py:158-184)."""
    text = str(path)
    if any(ch in text for ch in "*?["):
        files = [p for p in sorted(Path(m) for m in _glob.glob(text)) if p.is_file()]
        if not files:
            raise DataError(f"no SEG-Y file matches {text}")
        return files
    p = Path(text)
    if p.is_dir():
        files = sorted(q for q in p.iterdir() if q.is_file() and q.suffix.lower() in {".sgy", ".segy"})
        if not files:
            raise DataError(f"no SEG-Y file in directory {p}")
        return files
    if not p.is_file():
        raise DataError(f"SEG-Y file not found: {p}")
    return [p]


def summarize_segy(path: str | Path) -> dict[str, Any]:
    """Header-only QC summary (ertm/segy.py:187-234): files, traces, shots
    (distinct source x per file), dt, ns, coordinate and offset ranges."""
    files = expand_segy_paths(path)
    n_traces = n_shots = 0
    dts: set[float] = set()
    nss: set[int] = set()
    sx_all, gx_all, off_all = [], [], []
    for fp in files:
        f = _open(fp, headers_only=True)
        h = f.headers
        sx = _apply_scalar(h["SourceX"], h["SourceGroupScalar"])
        gx = _apply_scalar(h["GroupX"], h["SourceGroupScalar"])
        n_traces += h.size
        n_shots += np.unique(sx).size
        dts.add(f.interval * 1e-6)
        nss.add(f.ns)
        sx_all.append(sx)
        gx_all.append(gx)
        off_all.append(np.abs(h["offset"].astype(np.float64)))
    sx, gx, off = (np.concatenate(v) for v in (sx_all, gx_all, off_all))
    return {
        "files": [str(p) for p in files],
        "n_files": len(files),
        "n_traces": int(n_traces),
        "n_shots": int(n_shots),
        "dt": sorted(dts)[0] if len(dts) == 1 else sorted(dts),
        "ns": sorted(nss)[0] if len(nss) == 1 else sorted(nss),
        "source_x_range": [float(sx.min()), float(sx.max())],
        "receiver_x_range": [float(gx.min()), float(gx.max())],
        "offset_range": [float(off.min()), float(off.max())],
    }


def _read_single(fp: Path, opts: SegYReadOptions) -> list[ShotRecord]:
    f = _open(fp)
    h = f.headers
    scal = opts.coordinate_scalar if opts.coordinate_scalar else h["SourceGroupScalar"]
    sx = _apply_scalar(h["SourceX"], scal)
    gx = _apply_scalar(h["GroupX"], scal)
    gy = _apply_scalar(h["GroupY"], scal)
    data = _float_samples(f)
    dt = f.interval * 1e-6
    if opts.group_by == "file":
        keys = np.zeros(h.size, dtype=np.int64)
    elif opts.group_by == "field_record":
        keys = h["FieldRecord"].astype(np.int64)
    else:
        keys = np.unique(sx, return_inverse=True)[1]
    shots = []
    for key in np.unique(keys):
        idx = np.nonzero(keys == key)[0]
        if opts.sort_by_offset:
            idx = idx[np.argsort(gx[idx], kind="stable")]
        rz = gy[idx] if opts.receiver_depth_from == "group_y" else np.full(idx.size,
                                                                            float(opts.receiver_depth))
        traces = data[:, idx].astype(np.float32)
        rec = ShotRecord(
            data=traces, dt=dt, t=np.arange(traces.shape[0], dtype=np.float64) * dt,
            receivers=ReceiverSpec(x=gx[idx], z=rz, component=opts.component),
            source=SourceSpec(x=float(sx[idx[0]]), z=float(opts.source_depth), f0=float(opts.f0),
                              kind="ricker", amplitude=1.0, delay=float(opts.delay)),
            meta={"input": "segy", "file": str(fp), "component": opts.component,
                  "group_by": opts.group_by, "group_key": str(key), "n_traces": int(idx.size),
                  "placeholder": False},
            components={opts.component: traces})
        rec.validate()
        shots.append(rec)
    if opts.shot_limit:
        shots = shots[: int(opts.shot_limit)]
    logger.info("SEG-Y %s: %d shots, dt=%.3e s", fp, len(shots), dt)
    return shots


def read_segy_shots(path: str | Path, options: SegYReadOptions | None = None) -> list[ShotRecord]:
    """Shots of a SEG-Y file, directory or glob (ertm/segy.py:237-252)."""
    opts = options or SegYReadOptions()
    opts.validate()
    shots: list[ShotRecord] = []
    for fp in expand_segy_paths(path):
        shots.extend(_read_single(fp, opts))
    if opts.shot_limit:
        shots = shots[: int(opts.shot_limit)]
    return shots


def read_segy_components(component_files: dict[str, str | Path],
                         options: SegYReadOptions | None = None) -> list[ShotRecord]:
    """One file set per component, merged shot by shot (ertm/segy.py:326-376)."""
    if not component_files:
        raise DataError("component_files must not be empty")
    base = options or SegYReadOptions()
    per = {name: read_segy_shots(p, SegYReadOptions(**{**base.__dict__, "component": name}))
           for name, p in component_files.items()}
    counts = {name: len(v) for name, v in per.items()}
    if len(set(counts.values())) != 1:
        raise DataError(f"components hold different shot counts: {counts}")
    merged = []
    for i in range(next(iter(counts.values()))):
        first = next(iter(per.values()))[i]
        comps = {name: v[i].data for name, v in per.items()}
        merged.append(ShotRecord(data=comps.get("p", first.data), dt=first.dt, t=first.t,
                                 receivers=first.receivers, source=first.source,
                                 meta={**first.meta, "components": list(comps)},
                                 components=comps))
    return merged


# ----------------------------------------------------------------------------
# writing
# ----------------------------------------------------------------------------
def _shot_headers(shot: ShotRecord, field_record: int) -> dict[str, np.ndarray]:
    rx = np.asarray(shot.receivers.x, dtype=np.float64)
    rz = np.asarray(shot.receivers.z, dtype=np.float64)
    n = rx.size
    return {
        "SourceX": np.full(n, int(round(float(shot.source.x)))),
        "SourceY": np.full(n, int(round(float(shot.source.z)))),
        "GroupX": np.rint(rx).astype(np.int64),
        "GroupY": np.rint(rz).astype(np.int64),
        "SourceGroupScalar": np.ones(n, dtype=np.int64),
        "offset": np.rint(np.abs(rx - float(shot.source.x))).astype(np.int64),
        "FieldRecord": np.full(n, int(field_record)),
    }


def write_segy_shot(shot: ShotRecord, path: str | Path, *, fmt: int = 5,
                    component: str | None = None) -> Path:
    """One shot, trace = receiver, sample = time (ertm/segy.py:382-434)."""
    shot.validate()
    data = np.asarray(shot.component(component) if component else shot.data, dtype=np.float32)
    ns = data.shape[0]
    heads = _shot_headers(shot, 1)
    heads["TraceNumber"] = np.arange(1, data.shape[1] + 1)
    out = _write(Path(path), f"ERTM shot {shot.source.x:.1f}/{shot.source.z:.1f}", ns,
                 int(round(shot.dt * 1e6)), int(fmt), heads, data.T)
    logger.info("shot written to SEG-Y: %s (%d traces)", out, data.shape[1])
    return out


def write_segy_shots(shots: Sequence[ShotRecord], path: str | Path, *, fmt: int = 5) -> Path:
    """Several shots in one file, FieldRecord = shot number (ertm/segy.py:437-492)."""
    if not shots:
        raise DataError("writing SEG-Y needs at least one shot")
    for s in shots:
        s.validate()
    ns = int(shots[0].data.shape[0])
    if any(int(s.data.shape[0]) != ns for s in shots):
        raise DataError("all shots written to one file need the same record length")
    parts = [_shot_headers(s, k + 1) for k, s in enumerate(shots)]
    heads = {name: np.concatenate([p[name] for p in parts]) for name in parts[0]}
    total = sum(int(s.data.shape[1]) for s in shots)
    heads["TraceNumber"] = np.arange(1, total + 1)
    data = np.concatenate([np.asarray(s.data, dtype=np.float32) for s in shots], axis=1)
    out = _write(Path(path), f"ERTM shots x{len(shots)}", ns, int(round(shots[0].dt * 1e6)),
                 int(fmt), heads, data.T)
    logger.info("SEG-Y written: %s (%d shots, %d traces)", out, len(shots), total)
    return out


def write_segy_image(image: ImageResult, path: str | Path, *, fmt: int = 5) -> Path:
    """Depth image as SEG-Y: trace = x, sample = depth; the interval field
    holds dz in millimetres (ertm/segy.py:495-535)."""
    image.validate()
    data = np.asarray(image.image, dtype=np.float32)     # (nz, nx)
    nz, nx = data.shape
    dz_mm = max(int(round(float(image.grid.dz) * 1000)), 1)
    x = np.rint(np.array([image.grid.x_at(i) for i in range(nx)])).astype(np.int64)
    heads = {"SourceX": x, "GroupX": x, "SourceGroupScalar": np.ones(nx, dtype=np.int64),
             "CDP": np.arange(1, nx + 1)}
    out = _write(Path(path), f"ERTM image {image.mode.upper()} (depth domain)", nz, dz_mm,
                 int(fmt), heads, data.T)
    logger.info("image written to SEG-Y: %s (%d traces x %d samples)", out, nx, nz)
    return out


def read_segy_image(path: str | Path) -> tuple[np.ndarray, float, np.ndarray]:
    """(image[nz, nx], dz, x) of a ``write_segy_image`` file (ertm/segy.py:538-554)."""
    fp = Path(path)
    if not fp.is_file():
        raise DataError(f"SEG-Y file not found: {fp}")
    f = _open(fp)
    data = _float_samples(f).astype(np.float32)
    return data.reshape(f.ns, -1), f.interval / 1000.0, f.headers["GroupX"].astype(np.float64)
