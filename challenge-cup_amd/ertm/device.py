"""GPU-resident state of the engine: one ``Core`` per propagator, receiver and
imaging tables, and ``Batch`` handles over the C-ABI.

Memory layout (all allocated as torch tensors, which are used only as HBM
allocations; every computation is a HIP kernel of libertm_hip.so):

* two ping-pong state buffers per propagator, each holding the field slabs
  (padded grid, rows padded to 64 floats = 256 B) followed by the CPML memory
  strips;
* medium arrays on the padded grid: (kappa, 1/rho) acoustic, (lambda, mu,
  1/rho) elastic, (K_R, 1/rho, Q class) + per-Q tables viscoacoustic -- the
  face averages are recomputed in the kernels;
* CPML profiles per padded row / column.

Set-up arithmetic (edge padding, moduli, CPML profiles, fractional receiver
indices and weights) is done on the host in the reference's own precision
and operation order (ertm/propagation.py:165-187, :251-262, :331-369), so
the kernels see exactly the reference's float32 inputs.
"""

from __future__ import annotations

import contextlib
import os
import ctypes as C
import hashlib
import threading
import weakref

import numpy as np

from . import _lib as L
from . import numerics as num
from .exceptions import SolverError

FIELDS = {
    "acoustic": ("p", "vx", "vz"),
    "elastic": ("vx", "vz", "sxx", "szz", "sxz"),
    "viscoacoustic": ("p", "vx", "vz", "eps"),
}
# CPML memory names per engine, in psix / psiz order (the reference's names)
PSI_X = {
    "acoustic": ("vx_dp_dx", "p_dvx_dx"),
    "viscoacoustic": ("vx_dp_dx", "p_dvx_dx"),
    "elastic": ("vx_dsxx", "vz_dsxz", "sxx_dvx", "sxz_dvz"),
}
PSI_Z = {
    "acoustic": ("vz_dp_dz", "p_dvz_dz"),
    "viscoacoustic": ("vz_dp_dz", "p_dvz_dz"),
    "elastic": ("vx_dsxz", "vz_dszz", "szz_dvz", "sxz_dvx"),
}
OFFSETS = {"vx": (-0.5, 0.0), "vz": (0.0, -0.5), "sxz": (-0.5, -0.5)}
ROW_ALIGN = 64


def torch_mod():
    import torch

    return torch


def _rup(x: int, m: int) -> int:
    return (int(x) + m - 1) // m * m


def edge_pad(a: np.ndarray, pad: int, dtype=np.float32) -> np.ndarray:
    """ertm/propagation.py:165-166."""
    return np.pad(np.asarray(a).astype(dtype, copy=False), pad, mode="edge")


def to_device(a: np.ndarray, dtype=None):
    torch = torch_mod()
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda", non_blocking=False)


def padded_rows(a: np.ndarray, ld: int):
    """Upload an (nzp, nxp) host array into a (nzp, ld) device array."""
    torch = torch_mod()
    nzp, nxp = a.shape
    out = torch.zeros((nzp, ld), dtype=torch.from_numpy(np.zeros(1, a.dtype)).dtype, device="cuda")
    out[:, :nxp] = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    return out


try:
    import xxhash as _xxhash
except ImportError:  # pragma: no cover - xxhash ships with this image
    _xxhash = None

# Device copies of medium arrays shared by every propagator built on the same
# values (the shots of one survey share one model): one HBM copy, small
# enough to stay in the Infinity Cache while a batched launch reads it for
# every shot.
_MEDIA: "weakref.WeakValueDictionary" = weakref.WeakValueDictionary()

# Per-thread scope of model objects the caller has validated (see
# ``trusted_models``), with the digests computed inside it.
_TRUST = threading.local()


@contextlib.contextmanager
def trusted_models(models):
    """Within the block, propagators built on these model objects (validated
    by the caller and not modified meanwhile) skip re-validation, and each
    array's digest is computed once.  The cache holds a reference to every
    array it keys, so an id cannot be reused while the scope lasts."""
    prev = getattr(_TRUST, "scope", None)
    _TRUST.scope = {"ids": {id(m) for m in models}, "models": list(models), "digests": {}}
    try:
        yield
    finally:
        _TRUST.scope = prev


def is_trusted(model) -> bool:
    scope = getattr(_TRUST, "scope", None)
    return scope is not None and id(model) in scope["ids"]


def _digest(a: np.ndarray) -> bytes:
    scope = getattr(_TRUST, "scope", None)
    if scope is not None:
        hit = scope["digests"].get(id(a))
        if hit is not None and hit[0] is a:
            return hit[1]
    buf = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
    if _xxhash is not None:
        d = _xxhash.xxh3_128_digest(buf)
    else:
        d = hashlib.blake2b(buf, digest_size=16).digest()
    if scope is not None:
        scope["digests"][id(a)] = (a, d)
    return d


def _cache_put(key, t):
    """Enter ``t`` in the shared cache with an event recorded on the current
    stream after the work that filled it (``padded_rows`` fills the tensor
    asynchronously on this stream)."""
    torch = torch_mod()
    ev = torch.cuda.Event()
    ev.record(torch.cuda.current_stream())
    t._ertm_ready = ev
    _MEDIA[key] = t
    return t


def _cache_get(key):
    """A cached tensor, with the current stream made to wait until its upload
    has finished: a propagator built on another stream (a serving worker)
    must not read a medium that is still being written."""
    out = _MEDIA.get(key)
    if out is not None:
        torch_mod().cuda.current_stream().wait_event(out._ertm_ready)
    return out


def shared_padded_rows(a: np.ndarray, ld: int, key=None):
    """``padded_rows`` through a cache keyed by the array's bytes (or by
    ``key``, a digest of what the array was computed from), shape and the row
    pitch; the cached tensor lives while some propagator holds it."""
    key = (_digest(a) if key is None else key, a.shape, a.dtype.str, int(ld))
    out = _cache_get(key)
    if out is None:
        out = _cache_put(key, padded_rows(a, ld))
    return out


def pack_host(items) -> tuple[np.ndarray, list[int]]:
    """((array, dtype) pairs) -> one byte buffer holding each array, cast to
    its dtype, from a 256-byte-aligned offset; returns (buffer, offsets)."""
    arrs = [np.ascontiguousarray(np.asarray(a, dtype=dt)).reshape(-1) for a, dt in items]
    offs, total = [], 0
    for a in arrs:
        offs.append(total)
        total += _rup(max(a.nbytes, 1), 256)
    host = np.zeros(total, dtype=np.uint8)
    for a, o in zip(arrs, offs):
        host[o:o + a.nbytes] = a.view(np.uint8)
    return host, offs


def shared_vector(a: np.ndarray):
    """``to_device`` through the cache of ``shared_padded_rows``: the CPML
    profiles are the same for every propagator on one grid, spacing and time
    step, and kernels only read them."""
    key = ("vector", _digest(a), a.shape, a.dtype.str)
    out = _cache_get(key)
    if out is None:
        out = _cache_put(key, to_device(a))
    return out


def media_digest(*parts) -> tuple:
    """Key of a medium: the engine, padding and the bytes of the model arrays
    it is computed from."""
    return tuple((_digest(p), p.shape, p.dtype.str) if isinstance(p, np.ndarray) else p
                 for p in parts)


_HOST_MEDIA: dict = {}
_HOST_MEDIA_MAX = 4

# Media of at most this many distinct value tuples (layered models, a few
# anomalies) are also kept as one byte per cell indexing a table; the
# single-launch elastic step then reads 1 B of medium per cell instead of
# 12 B.  ERTM_MEDIUM_PALETTE=0 turns this off (the results are the same).
PALETTE_MAX = 255
MEDIUM_PALETTE = os.environ.get("ERTM_MEDIUM_PALETTE", "1") != "0"


def medium_palette(arrays: list) -> tuple | None:
    """(index, table, n) of same-shape float32 media, or None when palettes
    are off or the media hold more than ``PALETTE_MAX`` distinct tuples.

    index: uint8 of the arrays' shape, 1 + the tuple's position in the
    table; table: float32 (len(arrays) * 256,), array a's value of index i
    at a * 256 + i (index 0 is all zeros: padding and rows outside the
    grid); n: the number of distinct tuples.  Values are compared bit for
    bit, so the table holds exactly the arrays' float32 values.  Only the
    tuples at the starts of runs of equal tuples along the rows are sorted,
    which keeps layered media cheap (a 2000 x 6000 model: 0.03 s)."""
    if not MEDIUM_PALETTE or not arrays or len(arrays) > 3:
        return None
    shape = np.shape(arrays[0])
    bits = [np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).reshape(-1)
            for a in arrays]
    n = bits[0].size
    if n == 0 or any(b.size != n for b in bits):
        return None
    change = np.zeros(n, dtype=bool)
    change[0] = True
    for b in bits:
        change[1:] |= b[1:] != b[:-1]
    starts = np.flatnonzero(change)
    keys = np.stack([b[starts] for b in bits], axis=1)
    if len(starts) > 65536:
        # many runs: give up early when a sample already has too many tuples
        probe = keys[np.linspace(0, len(starts) - 1, 65536).astype(np.int64)]
        if len(np.unique(probe, axis=0)) > PALETTE_MAX:
            return None
    uniq, inv = np.unique(keys, axis=0, return_inverse=True)
    if len(uniq) > PALETTE_MAX:
        return None
    run = np.cumsum(change) - 1
    index = (np.asarray(inv).reshape(-1)[run] + 1).astype(np.uint8).reshape(shape)
    table = np.zeros((len(arrays), 256), dtype=np.uint32)
    table[:, 1:len(uniq) + 1] = uniq.T
    return index, table.view(np.float32).reshape(-1), int(len(uniq))


def host_media(key: tuple, build) -> dict:
    """Host medium arrays computed once per key (a few recent models kept):
    every shot of a survey builds its propagators on the same model."""
    out = _HOST_MEDIA.get(key)
    if out is None:
        out = build()
        if len(_HOST_MEDIA) >= _HOST_MEDIA_MAX:
            _HOST_MEDIA.pop(next(iter(_HOST_MEDIA)))
        _HOST_MEDIA[key] = out
    return out


class Core:
    """One propagator's device state and its C descriptor."""

    def __init__(self, engine: str, nz: int, nx: int, dx: float, dz: float, x0: float, z0: float,
                 order: int, dt: float, npml: int, boundary: str, vmax: float, media: dict,
                 nmech: int = 0, media_key: tuple | None = None):
        torch = torch_mod()
        L.lib()
        self.engine = engine
        self.nz, self.nx = int(nz), int(nx)
        self.dx, self.dz, self.x0, self.z0 = float(dx), float(dz), float(x0), float(z0)
        self.order = int(order)
        self.pad = num.ghost_width(order)
        self.nzp, self.nxp = self.nz + 2 * self.pad, self.nx + 2 * self.pad
        self.ld = _rup(self.nxp, ROW_ALIGN)
        self.dt = float(dt)
        self.boundary = boundary
        self.nmech = int(nmech)
        enabled = boundary != "none"
        free = boundary == "free"
        ax, bx, wx = num.cpml_profile(self.nxp, self.dx, npml, self.dt, vmax, 25.0, True, True,
                                      enabled)
        az, bz, wz = num.cpml_profile(self.nzp, self.dz, npml, self.dt, vmax, 25.0, not free, True,
                                      enabled)
        self.wx = wx if enabled else 0
        self.wzt = wz if (enabled and not free) else 0
        self.wzb = wz if enabled else 0
        self.prof = {"ax": ax, "bx": bx, "az": az, "bz": bz}
        self.d_prof = {k: shared_vector(v) for k, v in self.prof.items()}
        self.field_names = FIELDS[engine]
        self.nfields = len(self.field_names) + (self.nmech if engine == "viscoacoustic" else 0)
        self.fstride = _rup(self.nzp * self.ld, 64)
        off = self.nfields * self.fstride
        self.psix, self.psiz = [-1] * 4, [-1] * 4
        nterms = len(PSI_X[engine])
        self.xsize = self.nzp * 2 * self.wx
        self.zsize = (self.wzt + self.wzb) * self.ld
        for k in range(nterms):
            if self.wx > 0:
                self.psix[k] = off
                off += _rup(self.xsize, 64)
        for k in range(nterms):
            if self.zsize > 0:
                self.psiz[k] = off
                off += _rup(self.zsize, 64)
        self.state_floats = off
        self.state = torch.zeros((2, self.state_floats), dtype=torch.float32, device="cuda")
        self.cur = 0
        self.media = {k: (shared_padded_rows(v, self.ld, None if media_key is None
                                             else (media_key, k))
                          if isinstance(v, np.ndarray) and v.ndim == 2
                          else (to_device(v) if isinstance(v, np.ndarray) else v))
                      for k, v in media.items()}
        self.src = (-1, -1)
        self.mechanism = L.SRC_EXPLOSIVE
        self._desc = None

    # ----------------------------------------------------------- descriptor
    def desc(self) -> L.ErtmProp:
        d = L.ErtmProp()
        d.engine = L.ENGINE_CODE[self.engine]
        d.nzp, d.nxp, d.ld = self.nzp, self.nxp, self.ld
        d.pad, d.nz, d.nx, d.order = self.pad, self.nz, self.nx, self.order
        d.free_surface = int(self.boundary == "free" and self.engine != "viscoacoustic")
        d.nmech = self.nmech
        d.wx, d.wzt, d.wzb = self.wx, self.wzt, self.wzb
        d.nfields = self.nfields
        d.dx, d.dz, d.dt = self.dx, self.dz, self.dt  # rounded to float32 by ctypes
        for k, c in enumerate(num.coef_f32(self.order)):
            d.coef[k] = c
        d.dt64 = self.dt
        d.rdx = 1.0 / float(np.float32(self.dx))
        d.rdz = 1.0 / float(np.float32(self.dz))
        d.fstride = self.fstride
        for k in range(4):
            d.psix[k] = self.psix[k]
            d.psiz[k] = self.psiz[k]
        d.state_floats = self.state_floats
        m = self.media
        d.kappa = L.ptr(m.get("kappa"))
        d.irho = L.ptr(m.get("irho"))
        d.lam = L.ptr(m.get("lam"))
        d.mu = L.ptr(m.get("mu"))
        d.qidx = L.ptr(m.get("qidx"))
        d.q_a = L.ptr(m.get("q_a"))
        d.q_delta = L.ptr(m.get("q_delta"))
        d.q_tau = L.ptr(m.get("q_tau"))
        p = self.d_prof
        d.ax, d.bx, d.az, d.bz = L.ptr(p["ax"]), L.ptr(p["bx"]), L.ptr(p["az"]), L.ptr(p["bz"])
        d.state[0] = L.ptr(self.state[0])
        d.state[1] = L.ptr(self.state[1])
        d.src_iz, d.src_ix = self.src
        d.mechanism = self.mechanism
        d.midx = L.ptr(m.get("midx"))
        d.mtab = L.ptr(m.get("mtab"))
        d.npal = int(m.get("npal", 0)) if m.get("midx") is not None else 0
        return d

    # ------------------------------------------------------------- fields
    def slab(self, k: int, cur: int | None = None):
        """(nzp, nxp) device view of field slab k of state[cur]."""
        c = self.cur if cur is None else cur
        s = self.state[c, k * self.fstride : k * self.fstride + self.nzp * self.ld]
        return s.view(self.nzp, self.ld)[:, : self.nxp]

    def field_index(self, name: str) -> int:
        return self.field_names.index(name)

    def psi_views(self, cur: int | None = None) -> dict:
        c = self.cur if cur is None else cur
        out = {}
        for k, name in enumerate(PSI_X[self.engine]):
            if self.psix[k] >= 0:
                o = self.psix[k]
                out[name] = self.state[c, o : o + self.xsize].view(self.nzp, 2 * self.wx)
        for k, name in enumerate(PSI_Z[self.engine]):
            if self.psiz[k] >= 0:
                o = self.psiz[k]
                out[name] = self.state[c, o : o + self.zsize].view(self.wzt + self.wzb,
                                                                    self.ld)[:, : self.nxp]
        return out

    def psi_bytes(self) -> int:
        return 4 * (len(PSI_X[self.engine]) * self.xsize
                    + len(PSI_Z[self.engine]) * (self.wzt + self.wzb) * self.nxp)

    def expand_psi(self, name: str, compact: np.ndarray) -> np.ndarray:
        """Full padded-grid array of one CPML memory (zeros outside the strips)."""
        full = np.zeros((self.nzp, self.nxp), np.float32)
        if name in PSI_X[self.engine]:
            w = self.wx
            full[:, :w] = compact[:, :w]
            full[:, self.nxp - w :] = compact[:, w:]
        else:
            full[: self.wzt, :] = compact[: self.wzt]
            if self.wzb:
                full[self.nzp - self.wzb :, :] = compact[self.wzt :]
        return full

    # ----------------------------------------------------------- geometry
    def node_index(self, x: float, z: float) -> tuple[int, int]:
        """Nearest padded node (round half to even, clipped), ertm/propagation.py:246-249."""
        ix = int(round((float(x) - self.x0) / self.dx))
        iz = int(round((float(z) - self.z0) / self.dz))
        ix = min(max(ix, 0), self.nx - 1)
        iz = min(max(iz, 0), self.nz - 1)
        return iz + self.pad, ix + self.pad

    def fractional_index(self, x, z, comp: str):
        """ertm/propagation.py:251-262 with the per-component half-cell offsets."""
        ox, oz = OFFSETS.get(comp, (0.0, 0.0))
        xf = (np.asarray(x, dtype=np.float64) - self.x0) / self.dx + self.pad + ox
        zf = (np.asarray(z, dtype=np.float64) - self.z0) / self.dz + self.pad + oz
        return zf, xf

    def corners(self, x, z, comp: str):
        zf, xf = self.fractional_index(x, z, comp)
        z0 = np.floor(zf).astype(np.int64)
        x0 = np.floor(xf).astype(np.int64)
        tz, tx = zf - z0, xf - x0
        ok = (z0 >= 0) & (z0 < self.nzp - 1) & (x0 >= 0) & (x0 < self.nxp - 1)
        return np.clip(z0, 0, self.nzp - 2), np.clip(x0, 0, self.nxp - 2), tz, tx, ok


# ---------------------------------------------------------------------------
# receiver tables
# ---------------------------------------------------------------------------
def sample_field_code(engine: str, comp: str) -> tuple[int, str]:
    """(slab or special code, offset name) for sampling ``comp``."""
    if engine == "elastic":
        if comp == "p":
            return -2, "sxx"
        if comp in FIELDS["elastic"]:
            return FIELDS["elastic"].index(comp), comp
    elif engine == "acoustic":
        if comp == "sxz":
            return -1, comp
        if comp in FIELDS["acoustic"]:
            return FIELDS["acoustic"].index(comp), comp
    else:
        if comp in ("p", "vx", "vz"):
            return ("p", "vx", "vz").index(comp), comp
    names = {"elastic": "vx, vz, sxx, szz, sxz, p", "acoustic": "p, vx, vz",
             "viscoacoustic": "p, vx, vz"}[engine]
    raise SolverError(f"{engine} propagator cannot sample component {comp!r} (available: {names})")


def inject_targets(engine: str, comp: str) -> list[tuple[int, str, bool]]:
    """[(slab, offset name, scaled_by_1/K_R)] that ``inject(comp)`` writes."""
    if engine == "elastic":
        if comp == "p":  # full value into both normal stresses, ertm/propagation.py:907-910
            return [(2, "sxx", False), (3, "sxx", False)]
        if comp in FIELDS["elastic"]:
            return [(FIELDS["elastic"].index(comp), comp, False)]
    elif engine == "acoustic":
        if comp in FIELDS["acoustic"]:
            return [(FIELDS["acoustic"].index(comp), comp, False)]
    else:
        if comp == "p":  # pressure also enters the strain, ertm/propagation.py:770-774
            return [(0, "p", False), (3, "p", True)]
        if comp in ("vx", "vz"):
            return [(("p", "vx", "vz").index(comp), comp, False)]
    raise SolverError(f"{engine} propagator cannot inject component {comp!r}")


class RecvTables:
    """Device tables behind one ErtmRecv (sampling and/or injection)."""

    def __init__(self):
        self.s = L.ErtmRecv()
        self.keep: list = []

    def _dev(self, a, dtype):
        t = to_device(np.asarray(a, dtype=dtype))
        self.keep.append(t)
        return L.ptr(t)

    def _dev_pack(self, items) -> list:
        """Several small host arrays ((array, dtype) pairs) in one upload:
        pointers into one device buffer, each array on a 256-byte boundary.
        One copy instead of one per array, for the tables of many shots."""
        host, offs = pack_host(items)
        t = to_device(host)
        self.keep.append(t)
        base = L.ptr(t)
        return [base + o for o in offs]

    def set_records(self, records: list):
        """records[k]: (nt, nrec) float32 device tensors, one per component."""
        self.s.ncomp = len(records)
        for k, r in enumerate(records):
            self.s.records[k] = L.ptr(r)
        self.keep.extend(records)

    def build_sampling(self, core: Core, rx, rz, comps):
        i0, j0, fld, rec, col, val, tz, tx = [], [], [], [], [], [], [], []
        nrec = len(rx)
        for ci, comp in enumerate(comps):
            code, offname = sample_field_code(core.engine, comp)
            zc, xc, tzz, txx, ok = core.corners(rx, rz, offname)
            i0.append(zc)
            j0.append(xc)
            fld.append(np.full(nrec, code))
            rec.append(np.full(nrec, ci))
            col.append(np.arange(nrec))
            val.append(ok.astype(np.int32))
            tz.append(tzz)
            tx.append(txx)
        s = self.s
        s.nrec = nrec
        s.nsamp = nrec * len(comps)
        cat = np.concatenate
        (s.s_i0, s.s_j0, s.s_field, s.s_rec, s.s_col, s.s_valid, s.s_tz,
         s.s_tx) = self._dev_pack([(cat(i0), np.int32), (cat(j0), np.int32),
                                   (cat(fld), np.int32), (cat(rec), np.int32),
                                   (cat(col), np.int32), (cat(val), np.int32),
                                   (cat(tz), np.float64), (cat(tx), np.float64)])

    def build_injection(self, core: Core, rx, rz, comps, k_relaxed: np.ndarray | None = None):
        """Ordered scatter tables reproducing np.add.at (ertm/propagation.py:331-369,
        :472-494, :916-942): entries in (component, target, corner pass,
        receiver) order, grouped per target cell with that order kept."""
        nrec = len(rx)
        mode = 1 if core.engine == "viscoacoustic" else 0
        keys, comp_i, rec_i, wa_l, wb_l = [], [], [], [], []
        plane = core.fstride
        for ci, comp in enumerate(comps):
            for slab, offname, scaled in inject_targets(core.engine, comp):
                zc, xc, tz, tx, ok = core.corners(rx, rz, offname)
                zc, xc, tz, tx = zc[ok], xc[ok], tz[ok], tx[ok]
                ridx = np.arange(nrec)[ok]
                if mode == 1:
                    if scaled:
                        sc = 1.0 / np.maximum(np.abs(k_relaxed), 1e-20)  # float32 ops
                        fac = np.asarray(sc, np.float32)[zc, xc].astype(np.float64)
                    else:
                        fac = np.float64(np.float32(1.0))
                for dz_, dx_ in ((0, 0), (0, 1), (1, 0), (1, 1)):
                    wz = tz if dz_ else 1 - tz
                    wx = tx if dx_ else 1 - tx
                    off = (zc + dz_) * core.ld + (xc + dx_)
                    keys.append(slab * plane + off)
                    comp_i.append(np.full(ridx.size, ci))
                    rec_i.append(ridx)
                    if mode == 0:
                        wa_l.append(wz)
                        wb_l.append(wx)
                    else:
                        wa_l.append(wz * wx * fac)
                        wb_l.append(np.zeros(ridx.size))
        s = self.s
        s.inject_mode = mode
        s.nrec = nrec
        if not keys:
            s.ncells = 0
            return
        key = np.concatenate(keys)
        order = np.argsort(key, kind="stable")
        key = key[order]
        cells, start = np.unique(key, return_index=True)
        ptrs = np.concatenate([start, [key.size]]).astype(np.int32)
        s.ncells = int(cells.size)
        s.nentries = int(key.size)
        (s.cell_ptr, s.cell_field, s.cell_off, s.e_comp, s.e_rec, s.e_wa,
         s.e_wb) = self._dev_pack([(ptrs, np.int32), (cells // plane, np.int32),
                                   (cells % plane, np.int64),
                                   (np.concatenate(comp_i)[order], np.int32),
                                   (np.concatenate(rec_i)[order], np.int32),
                                   (np.concatenate(wa_l)[order], np.float64),
                                   (np.concatenate(wb_l)[order], np.float64)])


class ImageBuffers:
    """Accumulators and replay slots behind one ErtmImage."""

    def __init__(self, nz: int, nx: int, separation: str, swave: str, rwaves, start_step: int,
                 illum: int, excitation: bool = False, rbuf=None, nslots: int = 0,
                 ring: int = 0, ring_ld: int = 0, rring=None):
        torch = torch_mod()
        self.s = L.ErtmImage()
        s = self.s
        s.separation = L.SEP_DIVCURL if separation == "divcurl" else L.SEP_STRESS
        s.swave = L.WAVE_S if swave == "S" else L.WAVE_P
        s.nrw = len(rwaves)
        for k, w in enumerate(rwaves):
            s.rwave[k] = L.WAVE_S if w == "S" else L.WAVE_P
        s.start_step = int(start_step)
        s.illum = int(illum)  # ILLUM_SRC | ILLUM_REC bits
        z = lambda dt: torch.zeros((nz, nx), dtype=dt, device="cuda")  # noqa: E731
        # one allocation, num[0..nrw) | ill | rill[0..nrw): the fused step
        # kernel reaches them all through one buffer descriptor
        nrw = len(rwaves)
        self.acc = torch.zeros((2 * nrw + 1, nz, nx), dtype=torch.float64, device="cuda")
        self.num = [self.acc[k] for k in range(nrw)]
        self.ill = self.acc[nrw]
        self.rill = [self.acc[nrw + 1 + k] for k in range(nrw)]
        for k in range(len(rwaves)):
            s.num[k] = L.ptr(self.num[k])
            s.rill[k] = L.ptr(self.rill[k])
        s.ill = L.ptr(self.ill)
        self.rbuf = rbuf
        s.rbuf = L.ptr(rbuf)
        s.nslots = int(nslots)
        self.rring = None
        if ring > 0 and nrw > 0:
            # rows on the propagator's pitch (ring_ld >= nx + 2 pad), cell (i, j)
            # at column j + pad
            if ring_ld < nx:
                raise ValueError("ring_ld must be the propagator's row pitch")
            shape = (int(ring), nrw, nz, int(ring_ld))
            if rring is None:
                rring = torch.empty(shape, dtype=torch.float32, device="cuda")
            elif tuple(rring.shape) != shape or rring.dtype != torch.float32:
                raise ValueError(f"ring buffer of shape {tuple(rring.shape)}, expected {shape}")
            self.rring = rring
            s.rring = L.ptr(self.rring)
            s.ring_len = int(ring)
        if excitation:
            self.peak_step = torch.full((nz, nx), -1, dtype=torch.int32, device="cuda")
            self.peak_amp = z(torch.float64)
            self.running = z(torch.float64)
            s.peak_step = L.ptr(self.peak_step)
            s.peak_amp = L.ptr(self.peak_amp)
            s.running = L.ptr(self.running)


# ---------------------------------------------------------------------------
# batches
# ---------------------------------------------------------------------------
class Batch:
    """A set of propagators advanced together (one ertm_batch_t)."""

    def __init__(self, cores: list, recvs: list | None = None, imgs: list | None = None,
                 variant: tuple[int, int] | None = None, graph_steps: int = 0,
                 two_step: tuple[int, int] | None = (0, 0)):
        lib = L.lib()
        n = len(cores)
        self.cores = list(cores)
        self.recvs = recvs
        self.imgs = imgs
        self._props = (L.ErtmProp * n)(*[c.desc() for c in cores])
        self._recv = (L.ErtmRecv * n)(*[r.s for r in recvs]) if recvs else None
        self._img = (L.ErtmImage * n)(*[i.s for i in imgs]) if imgs else None
        h = C.c_void_p()
        L.check(lib.ertm_batch_create(self._props, self._recv, self._img, n, C.byref(h)),
                "ertm_batch_create")
        self.h = h
        self.n = n
        self.graph_steps = int(graph_steps)  # ertm_run: chunks replayed as one graph
        from . import tuning

        c0 = cores[0]
        same = all((c.engine, c.order, c.nmech) == (c0.engine, c0.order, c0.nmech)
                   for c in cores)
        fusable = same and tuning.fused_supported(c0.engine, c0.order, c0.nmech)
        if variant is None:
            variant = (tuning.lookup(c0.engine, c0.order, max(c.nzp for c in cores),
                                     max(c.nxp for c in cores), n, c0.nmech)
                       if fusable else (0, 64))
        elif int(variant[0]) != 0 and not fusable:
            variant = (0, 64)  # no single-launch kernel for this engine/order
        self.set_variant(*variant)
        self.two_step = (0, 0)
        if two_step is None:
            two_step = tuning.two_step_lookup(c0.engine, c0.order, max(c.nzp for c in cores),
                                              max(c.nxp for c in cores), n)
        if int(two_step[0]) != 0 and same and c0.engine == "elastic" and c0.order == 8:
            self.set_two_step(*two_step)

    def set_two_step(self, variant: int, zc: int) -> None:
        """Two time steps per launch in the forward and replay passes of
        ``run`` (ertm_batch_set_two_step); (0, 0) = off."""
        L.check(L.lib().ertm_batch_set_two_step(self.h, int(variant), int(zc)),
                "ertm_batch_set_two_step")
        self.two_step = (int(variant), int(zc))

    def set_ring_flush(self, period: int) -> None:
        """Fold the receiver-field ring every ``period`` backward steps on a
        stream of the library's own, beside the next steps
        (ertm_batch_set_ring_flush); 0 = fold on the launching stream when
        the ring is full."""
        L.check(L.lib().ertm_batch_set_ring_flush(self.h, int(period)),
                "ertm_batch_set_ring_flush")

    def set_variant(self, variant: int, zc: int = 64) -> None:
        L.check(L.lib().ertm_batch_set_variant(self.h, int(variant), int(zc)),
                "ertm_batch_set_variant")
        self.variant = (int(variant), int(zc))

    def update_images(self, imgs: list):
        self.imgs = imgs
        self._img = (L.ErtmImage * self.n)(*[i.s for i in imgs])
        L.check(L.lib().ertm_batch_update_image(self.h, self._img, L.stream_handle()),
                "ertm_batch_update_image")

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            L.lib().ertm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    # single-call wrappers ------------------------------------------------
    def step(self, cur: int, amps) -> None:
        a = (C.c_float * self.n)(*[float(v) for v in amps])
        L.check(L.lib().ertm_step(self.h, cur, a, L.stream_handle()), "ertm_step")

    def sample(self, cur: int, t) -> None:
        tt = (C.c_int32 * self.n)(*[int(v) for v in t])
        L.check(L.lib().ertm_sample(self.h, cur, tt, L.stream_handle()), "ertm_sample")

    def inject(self, cur: int, t) -> None:
        tt = (C.c_int32 * self.n)(*[int(v) for v in t])
        L.check(L.lib().ertm_inject(self.h, cur, tt, L.stream_handle()), "ertm_inject")

    def image(self, cur: int, t, op: int) -> None:
        tt = (C.c_int32 * self.n)(*[int(v) for v in t])
        L.check(L.lib().ertm_image(self.h, cur, tt, op, L.stream_handle()), "ertm_image")

    def stabilize(self, cur: int, step_index: int, cap: float, interval: int, ctl) -> None:
        L.check(L.lib().ertm_stabilize(self.h, cur, int(step_index), float(cap), int(interval),
                                       C.c_void_p(L.ptr(ctl)), L.stream_handle()),
                "ertm_stabilize")

    def run_lanes(self, segs, nseg: int, curs: list, *, max_ticks: int = 0,
                  stop_at_lane_end: bool = False) -> tuple[int, int, int]:
        """One ertm_run_lanes call over ``nseg`` ErtmSeg records at ``segs``
        (a ctypes pointer; their ``done`` counters advance in place);
        ``curs`` holds each propagator's parity and is updated.  The host
        wavelets the segments point to must stay alive for the call.
        Returns (ticks, lane steps, step-launch groups)."""
        if len(curs) != self.n:
            raise ValueError("one parity per propagator expected")
        cur = (C.c_int32 * self.n)(*[int(c) for c in curs])
        stats = (C.c_int64 * 3)()
        L.check(L.lib().ertm_run_lanes(self.h, segs, int(nseg), cur, int(max_ticks),
                                       int(bool(stop_at_lane_end)), stats, L.stream_handle()),
                "ertm_run_lanes")
        curs[:] = list(cur)
        return int(stats[0]), int(stats[1]), int(stats[2])

    def _ptr_array(self, tensors, keep):
        arr = (C.c_void_p * self.n)(*[L.ptr(a) for a in tensors])
        keep.append(arr)
        return C.cast(arr, C.POINTER(C.c_void_p))

    def run(self, phase: int, t_begin: int, t_end: int, cur0: int, *, wavelets=None,
            do_sample=False, op_after=-1, ck_interval=0, ck_origin=0, save=None, save_stride=0,
            restore=None, restore_stride=0, restore_slot=0, stabilize=False, step_index0=0,
            lp_interval=0, cap=0.0, ctl=None) -> int:
        """One ertm_run call; returns the parity of the final state."""
        r = L.ErtmRun()
        r.phase, r.t_begin, r.t_end, r.cur0 = phase, t_begin, t_end, cur0
        r.do_sample, r.op_after = int(do_sample), int(op_after)
        r.ck_interval, r.ck_origin = int(ck_interval), int(ck_origin)
        keep = []
        if save is not None:
            r.save_arenas = self._ptr_array(save, keep)
            r.save_stride = int(save_stride)
        if restore is not None:
            r.restore_arenas = self._ptr_array(restore, keep)
            r.restore_stride = int(restore_stride)
            r.restore_slot = int(restore_slot)
        if wavelets is not None:
            ws = [np.ascontiguousarray(w, dtype=np.float32) for w in wavelets]
            keep.extend(ws)
            arr = (C.c_void_p * self.n)(*[w.ctypes.data for w in ws])
            keep.append(arr)
            r.wavelets = C.cast(arr, C.POINTER(C.c_void_p))
            r.nw = min(len(w) for w in ws)
        r.stabilize = int(stabilize)
        r.step_index0 = int(step_index0)
        r.lp_interval = int(lp_interval)
        r.graph_steps = self.graph_steps
        r.cap = float(cap)
        r.ctl = L.ptr(ctl)
        cur = C.c_int32(0)
        L.check(L.lib().ertm_run(self.h, C.byref(r), C.byref(cur), L.stream_handle()), "ertm_run")
        return int(cur.value)


# ---------------------------------------------------------------------------
# utilities
# ---------------------------------------------------------------------------
def derivative(array: np.ndarray, axis: int, order: int, h: float, plus: bool) -> np.ndarray:
    torch = torch_mod()
    lib = L.lib()
    is64 = np.result_type(array.dtype, np.float32) == np.float64
    dt = np.float64 if is64 else np.float32
    x = to_device(np.asarray(array, dtype=dt))
    out = torch.empty_like(x)
    n0, n1 = array.shape
    L.check(lib.ertm_derivative(C.c_void_p(L.ptr(x)), C.c_void_p(L.ptr(out)), n0, n1, axis,
                                int(plus), order, float(h), int(is64), L.stream_handle()),
            "ertm_derivative")
    return out.cpu().numpy()


def absmax(t) -> float:
    torch = torch_mod()
    lib = L.lib()
    out = torch.zeros(1, dtype=torch.float32, device="cuda")
    L.check(lib.ertm_absmax(C.c_void_p(L.ptr(t)), t.numel(), C.c_void_p(L.ptr(out)),
                            L.stream_handle()), "ertm_absmax")
    return float(out.item())


def postprocess(num, ill, rill, *, dx: float, dz: float, eps: float, div_ill: bool,
                div_rill: bool, laplacian: bool, lowpass: bool, fraction: float, npml: int,
                mask_top: bool):
    """Normalise, Laplacian, wavenumber low-pass and PML mask of one image on
    the device (ertm_postprocess); returns a (nz, nx) float32 device tensor."""
    torch = torch_mod()
    nz, nx = num.shape
    work = torch.empty(4 * nz * nx, dtype=torch.float64, device="cuda")
    out = torch.empty((nz, nx), dtype=torch.float32, device="cuda")
    p = L.ErtmPost()
    p.nz, p.nx = int(nz), int(nx)
    p.div_ill, p.div_rill = int(bool(div_ill)), int(bool(div_rill))
    p.laplacian, p.lowpass = int(bool(laplacian)), int(bool(lowpass))
    p.npml, p.mask_top = int(npml), int(bool(mask_top))
    p.dx, p.dz, p.eps, p.fraction = float(dx), float(dz), float(eps), float(fraction)
    p.num = L.ptr(num)
    p.ill = L.ptr(ill) if div_ill else None
    p.rill = L.ptr(rill) if div_rill else None
    p.work = L.ptr(work)
    p.out = L.ptr(out)
    L.check(L.lib().ertm_postprocess(C.byref(p), L.stream_handle()), "ertm_postprocess")
    return out


class StepTimer:
    """HIP-event timing of every step-kernel launch while active."""

    def __enter__(self):
        L.check(L.lib().ertm_step_timing(1), "ertm_step_timing")
        return self

    def __exit__(self, *exc):
        L.check(L.lib().ertm_step_timing(0), "ertm_step_timing")
        return False

    @staticmethod
    def read() -> dict:
        """The step launches without imaging: summed ms, launches, cell-updates,
        fused (1) / two-launch (0) / both (2) / none (-1)."""
        ms = C.c_double(0.0)
        n = C.c_int64(0)
        cells = C.c_int64(0)
        fused = C.c_int32(-1)
        L.check(L.lib().ertm_step_timing_read(C.byref(ms), C.byref(n), C.byref(cells),
                                              C.byref(fused)), "ertm_step_timing_read")
        return {"total_ms": float(ms.value), "launches": int(n.value),
                "cell_updates": int(cells.value), "fused": int(fused.value)}

    @staticmethod
    def read_kinds() -> dict:
        """Every timed launch kind: {kind: {"ms", "launches", "cells"}}."""
        k = len(L.TIMING_KINDS)
        ms = (C.c_double * k)()
        n = (C.c_int64 * k)()
        cells = (C.c_int64 * k)()
        L.check(L.lib().ertm_timing_read(ms, n, cells), "ertm_timing_read")
        return {name: {"ms": float(ms[i]), "launches": int(n[i]), "cells": int(cells[i])}
                for i, name in enumerate(L.TIMING_KINDS)}

