"""Oracle propagators: acoustic, elastic P-SV and viscoacoustic GSLS engines.

Test infrastructure only.  Each class reproduces the reference engine's
per-step arithmetic in the same order (ertm/propagation.py), operating on
plain arrays so that the checker shares no code with the product.

Inputs are physical-grid arrays (nz, nx); ``dtype`` is "float32" (reference
order, bit-exact target) or "float64" (the reference's float64 mode, where the
CPML memory variables stay float32 exactly as ertm/numerics.py:299-304 does).
"""

from __future__ import annotations

import numpy as np

from . import numerics as onum

# receiver half-cell offsets (x, z), ertm/propagation.py:340-344 and :377-381
OFFSETS = {"vx": (-0.5, 0.0), "vz": (0.0, -0.5), "sxz": (-0.5, -0.5)}


def _edge_pad(a, pad, dt):
    """ertm/propagation.py:165-166."""
    return np.pad(np.asarray(a).astype(dt, copy=False), pad, mode="edge")


def _avg_x(a, dt):
    """ertm/propagation.py:169-174: mean onto x faces, last column copied."""
    out = np.empty_like(a)
    out[:, :-1] = a.dtype.type(0.5) * (a[:, :-1] + a[:, 1:])
    out[:, -1] = a[:, -1]
    return out.astype(dt, copy=False)


def _avg_z(a, dt):
    """ertm/propagation.py:177-182: mean onto z faces, last row copied."""
    out = np.empty_like(a)
    out[:-1, :] = a.dtype.type(0.5) * (a[:-1, :] + a[1:, :])
    out[-1, :] = a[-1, :]
    return out.astype(dt, copy=False)


class Grid:
    """Geometry of one propagator (ertm/types.py:20-82 + propagation.py:196-215)."""

    def __init__(self, nz, nx, dx, dz, x0=0.0, z0=0.0, order=8):
        self.nz, self.nx = int(nz), int(nx)
        self.dx, self.dz = float(dx), float(dz)
        self.x0, self.z0 = float(x0), float(z0)
        self.order = int(order)
        self.pad = onum.ghost(order)
        self.shape = (self.nz + 2 * self.pad, self.nx + 2 * self.pad)

    def node(self, x, z):
        """Nearest padded node, Python round (half to even), clipped.

        ertm/types.py:73-82 + ertm/propagation.py:246-249.
        """
        ix = int(round((float(x) - self.x0) / self.dx))
        iz = int(round((float(z) - self.z0) / self.dz))
        ix = min(max(ix, 0), self.nx - 1)
        iz = min(max(iz, 0), self.nz - 1)
        return iz + self.pad, ix + self.pad

    def frac(self, x, z, comp):
        """Fractional padded indices, ertm/propagation.py:251-262."""
        ox, oz = OFFSETS.get(comp, (0.0, 0.0))
        xf = (np.asarray(x, np.float64) - self.x0) / self.dx + self.pad + ox
        zf = (np.asarray(z, np.float64) - self.z0) / self.dz + self.pad + oz
        return zf, xf

    def crop(self, f):
        p = self.pad
        return f[p : p + self.nz, p : p + self.nx]


class _Base:
    """Shared machinery: CPML memories, source, receivers, state I/O."""

    fields: tuple = ()

    def __init__(self, grid: Grid, vp, dt, npml=20, boundary="cpml", dtype="float32"):
        self.g = grid
        self.dt = float(dt)
        self.dtype = np.dtype(dtype)
        self.boundary = boundary
        self.npml = int(npml)
        self.vmax = float(np.max(vp))
        self.step_index = 0
        self.src = None
        self.mechanism = "explosive"
        self.rx = self.rz = None
        self._make_pml()

    def _make_pml(self):
        """ertm/propagation.py:218-234 (f0 fixed at 25 Hz for alpha_max)."""
        nzp, nxp = self.g.shape
        en = self.boundary != "none"
        free = self.boundary == "free"
        self.ax, self.bx, self.wx = onum.cpml_profile(
            nxp, self.g.dx, self.npml, self.dt, self.vmax, 25.0, True, True, en)
        self.az, self.bz, self.wz = onum.cpml_profile(
            nzp, self.g.dz, self.npml, self.dt, self.vmax, 25.0, not free, True, en)
        self.psi = {}

    # -- derivative with CPML memory, ertm/numerics.py:221-227 and :283-308
    def d(self, name, f, axis, plus, spacing):
        raw = onum.derivative(f, axis, self.g.order, spacing, plus)
        mem = self.psi.get(name)
        if mem is None:
            mem = np.zeros(raw.shape, np.float32)
            self.psi[name] = mem
        if axis == 1:
            a, b = self.ax[None, :], self.bx[None, :]
        else:
            a, b = self.az[:, None], self.bz[:, None]
        mem *= b
        mem += a * raw
        return raw + mem

    def set_source(self, x, z, mechanism="explosive"):
        self.src = self.g.node(x, z)
        self.mechanism = mechanism

    def set_receivers(self, x, z):
        self.rx = np.asarray(x, np.float64).reshape(-1)
        self.rz = np.asarray(z, np.float64).reshape(-1)

    def _sample(self, field, comp):
        zf, xf = self.g.frac(self.rx, self.rz, comp)
        return onum.bilinear(field, zf, xf).astype(self.dtype, copy=False)

    def _corners(self, comp):
        zf, xf = self.g.frac(self.rx, self.rz, comp)
        z0 = np.floor(zf).astype(np.int64)
        x0 = np.floor(xf).astype(np.int64)
        tz, tx = zf - z0, xf - x0
        nz, nx = self.g.shape
        ok = (z0 >= 0) & (z0 < nz - 1) & (x0 >= 0) & (x0 < nx - 1)
        return np.clip(z0, 0, nz - 2)[ok], np.clip(x0, 0, nx - 2)[ok], tz[ok], tx[ok], ok

    def _scatter(self, field, values, comp):
        """Transposed bilinear scatter with ordered accumulation.

        ertm/propagation.py:472-494 (acoustic), :916-942 (elastic): weights
        applied as (v*wz)*wx in float64, four np.add.at passes 00,01,10,11.
        """
        zc, xc, tz, tx, ok = self._corners(comp)
        v = np.asarray(values, self.dtype)[ok]
        np.add.at(field, (zc, xc), v * (1 - tz) * (1 - tx))
        np.add.at(field, (zc, xc + 1), v * (1 - tz) * tx)
        np.add.at(field, (zc + 1, xc), v * tz * (1 - tx))
        np.add.at(field, (zc + 1, xc + 1), v * tz * tx)

    # -- state, ertm/propagation.py:283-311
    def state(self):
        st = {n: getattr(self, n).copy() for n in self.fields}
        for k, v in self.psi.items():
            st["pml:" + k] = v.copy()
        st["_step"] = np.array(self.step_index, np.int64)
        return st

    def load_state(self, st):
        for n in self.fields:
            getattr(self, n)[...] = st[n]
        for k, mem in self.psi.items():
            if ("pml:" + k) in st:
                mem[...] = st["pml:" + k]
        self.step_index = int(np.asarray(st["_step"]).item())

    def state_bytes(self):
        return int(sum(getattr(self, n).nbytes for n in self.fields)
                   + sum(v.nbytes for v in self.psi.values()))


class Acoustic(_Base):
    """ertm/propagation.py:397-506."""

    fields = ("p", "vx", "vz")

    def __init__(self, grid, vp, rho, dt, npml=20, boundary="cpml", dtype="float32"):
        super().__init__(grid, vp, dt, npml, boundary, dtype)
        t = self.dtype
        r = _edge_pad(rho, grid.pad, t)
        v = _edge_pad(vp, grid.pad, t)
        self.kappa = (r * v * v).astype(t)
        ir = (1.0 / r).astype(t)
        self.irx = _avg_x(ir, t)
        self.irz = _avg_z(ir, t)
        self.p = np.zeros(grid.shape, t)
        self.vx = np.zeros(grid.shape, t)
        self.vz = np.zeros(grid.shape, t)

    def step(self, amp=0.0):
        """ertm/propagation.py:430-454."""
        t = self.dtype.type
        dt, g = t(self.dt), self.g
        self.vx += dt * self.irx * self.d("vx_dp_dx", self.p, 1, True, g.dx)
        self.vz += dt * self.irz * self.d("vz_dp_dz", self.p, 0, True, g.dz)
        div = self.d("p_dvx_dx", self.vx, 1, False, g.dx)
        div += self.d("p_dvz_dz", self.vz, 0, False, g.dz)
        self.p += dt * self.kappa * div
        if amp != 0.0 and self.src is not None:
            self.p[self.src] += t(amp)
        if self.boundary == "free":
            self.p[g.pad, :] = 0.0
        self.step_index += 1

    def sample(self, comps=("p",)):
        """ertm/propagation.py:456-470."""
        out = {}
        for c in comps:
            if c == "sxz":
                out[c] = np.zeros(self.rx.size, self.dtype)
            else:
                out[c] = self._sample({"p": self.p, "vx": self.vx, "vz": self.vz}[c], c)
        return out

    def inject(self, comp, values):
        self._scatter({"p": self.p, "vx": self.vx, "vz": self.vz}[comp], values, comp)

    def image_fields(self, sep="stress"):
        return {"P": self.g.crop(self.p)}


class Elastic(_Base):
    """ertm/propagation.py:804-983."""

    fields = ("vx", "vz", "sxx", "szz", "sxz")

    def __init__(self, grid, vp, vs, rho, dt, npml=20, boundary="cpml", dtype="float32"):
        super().__init__(grid, vp, dt, npml, boundary, dtype)
        t = self.dtype
        r = _edge_pad(rho, grid.pad, t)
        v = _edge_pad(vp, grid.pad, t)
        s = _edge_pad(vs, grid.pad, t)
        self.mu = (r * s * s).astype(t)
        self.lam = (r * (v * v - 2.0 * s * s)).astype(t)
        self.lam2mu = (self.lam + 2.0 * self.mu).astype(t)
        self.muxz = _avg_z(_avg_x(self.mu, t), t)
        ir = (1.0 / r).astype(t)
        self.irx = _avg_x(ir, t)
        self.irz = _avg_z(ir, t)
        for n in self.fields:
            setattr(self, n, np.zeros(grid.shape, t))

    def step(self, amp=0.0):
        """ertm/propagation.py:840-883."""
        t = self.dtype.type
        dt, g = t(self.dt), self.g
        d = self.d
        self.vx += dt * self.irx * (d("vx_dsxx", self.sxx, 1, True, g.dx)
                                    + d("vx_dsxz", self.sxz, 0, False, g.dz))
        self.vz += dt * self.irz * (d("vz_dsxz", self.sxz, 1, False, g.dx)
                                    + d("vz_dszz", self.szz, 0, True, g.dz))
        ex = d("sxx_dvx", self.vx, 1, False, g.dx)
        ez = d("szz_dvz", self.vz, 0, False, g.dz)
        self.sxx += dt * (self.lam2mu * ex + self.lam * ez)
        self.szz += dt * (self.lam * ex + self.lam2mu * ez)
        self.sxz += dt * self.muxz * (d("sxz_dvx", self.vx, 0, True, g.dz)
                                      + d("sxz_dvz", self.vz, 1, True, g.dx))
        if amp != 0.0 and self.src is not None:
            a = t(amp)
            if self.mechanism == "shear":
                self.sxz[self.src] += a
            elif self.mechanism == "force_z":
                self.vz[self.src] += a
            else:
                self.sxx[self.src] += a
                self.szz[self.src] += a
        if self.boundary == "free":
            self.szz[g.pad, :] = 0.0
            self.sxz[g.pad, :] = 0.0
        self.step_index += 1

    def pressure(self):
        return self.dtype.type(0.5) * (self.sxx + self.szz)

    def sample(self, comps=("vx", "vz")):
        """ertm/propagation.py:885-898 ("p" uses the sxx offsets)."""
        out = {}
        for c in comps:
            if c == "p":
                out[c] = self._sample(self.pressure(), "sxx")
            else:
                out[c] = self._sample(getattr(self, c), c)
        return out

    def inject(self, comp, values):
        """ertm/propagation.py:900-914: "p" goes in full into sxx and szz."""
        if comp == "p":
            self._scatter(self.sxx, values, "sxx")
            self._scatter(self.szz, values, "sxx")
        else:
            self._scatter(getattr(self, comp), values, comp)

    def image_fields(self, sep="divcurl"):
        """ertm/propagation.py:945-977 (divcurl derivatives carry no CPML)."""
        g = self.g
        if sep == "stress":
            return {"P": g.crop(self.pressure()), "S": g.crop(self.sxz)}
        div = onum.derivative(self.vx, 1, g.order, g.dx, False) + onum.derivative(
            self.vz, 0, g.order, g.dz, False)
        curl = onum.derivative(self.vx, 0, g.order, g.dz, False) - onum.derivative(
            self.vz, 1, g.order, g.dx, False)
        return {"P": g.crop(div), "S": g.crop(curl)}


class Visco(_Base):
    """ertm/propagation.py:512-798 (GSLS memory-variable viscoacoustic engine).

    ``relax`` supplies the per-unique-Q tables (see ``gsls.relaxation``):
    unique Q values, tau (U, L), delta (U, L) already sign-flipped and scaled
    when compensating, and modulus_ref (U,).
    """

    fields = ("p", "vx", "vz", "eps", "r")

    def __init__(self, grid, vp, rho, qp, dt, relax, npml=20, boundary="cpml",
                 dtype="float32", compensation=False, amplitude_cap=50.0,
                 lowpass_interval=20):
        super().__init__(grid, vp, dt, npml, boundary, dtype)
        t = self.dtype
        r = _edge_pad(rho, grid.pad, t)
        v = _edge_pad(vp, grid.pad, t)
        q = _edge_pad(qp, grid.pad, t)
        kr0 = (r * v * v).astype(t)
        ir = (1.0 / r).astype(t)
        self.irx = _avg_x(ir, t)
        self.irz = _avg_z(ir, t)
        vals = np.clip(np.asarray(q, np.float64).ravel(), 2.5, 1.0e5)
        uniq = relax["unique"]
        idx = np.searchsorted(uniq, vals).reshape(grid.shape)
        tau_n = np.moveaxis(relax["tau"][idx], -1, 0)
        dl_n = np.moveaxis(relax["delta"][idx], -1, 0)
        self.kr = (kr0 / relax["modulus_ref"][idx]).astype(t)
        self.a = np.exp(-self.dt / tau_n).astype(t)
        self.b = (self.kr[None, :, :] * dl_n * (1.0 - self.a) * tau_n / self.dt).astype(t)
        self.L = tau_n.shape[0]
        self.compensation = bool(compensation)
        self.cap = float(amplitude_cap)
        self.interval = int(lowpass_interval)
        self.ref_amp = 0.0
        self.clipped = 0
        self.lowpass_steps = 0
        z = lambda: np.zeros(grid.shape, t)  # noqa: E731
        self.p, self.vx, self.vz, self.eps = z(), z(), z(), z()
        self.r = np.zeros((self.L,) + grid.shape, t)

    def step(self, amp=0.0):
        """ertm/propagation.py:686-723."""
        t = self.dtype.type
        dt, g = t(self.dt), self.g
        self.vx += dt * self.irx * self.d("vx_dp_dx", self.p, 1, True, g.dx)
        self.vz += dt * self.irz * self.d("vz_dp_dz", self.p, 0, True, g.dz)
        div = self.d("p_dvx_dx", self.vx, 1, False, g.dx)
        div += self.d("p_dvz_dz", self.vz, 0, False, g.dz)
        inc = (dt * div).astype(self.dtype, copy=False)
        self.eps += inc
        self.r *= self.a
        self.r += self.b * inc[None, :, :]
        tot = self.r.sum(axis=0)
        np.add(self.kr * self.eps, tot, out=self.p)
        if amp != 0.0 and self.src is not None:
            self.eps[self.src] += t(amp / float(self.kr[self.src]))
            self.p[self.src] += t(amp)
        if self.compensation:
            self._stabilize()
        self.step_index += 1

    def _stabilize(self):
        """ertm/propagation.py:726-749: amplitude cap + periodic k-space low-pass."""
        peak = float(np.max(np.abs(self.p)))
        if self.step_index < 30:
            self.ref_amp = max(self.ref_amp, peak)
            return
        limit = self.cap * max(self.ref_amp, 1e-30)
        if peak > limit:
            s = self.dtype.type(limit / peak)
            for n in ("p", "eps", "vx", "vz", "r"):
                getattr(self, n)[...] *= s
            self.clipped += 1
        if self.interval > 1 and self.step_index % self.interval == 0:
            self.lowpass_steps += 1
            tt = self.step_index * self.dt
            ratio = max(0.9 - 0.4 * min(tt / 1.0, 1.0), 0.35)
            self.p[:] = onum.lowpass(self.p, self.g.dx, self.g.dz, ratio)
            self.eps[:] = onum.lowpass(self.eps, self.g.dx, self.g.dz, ratio)

    def sample(self, comps=("p",)):
        return {c: self._sample({"p": self.p, "vx": self.vx, "vz": self.vz}[c], c)
                for c in comps}

    def _scatter_scaled(self, field, values, comp, scale=None):
        """ertm/propagation.py:331-369: weight ((wz*wx)*factor), factor taken
        at the (z0, x0) corner for all four corners."""
        zc, xc, tz, tx, ok = self._corners(comp)
        v = np.asarray(values, self.dtype)[ok]
        fac = np.float32(1.0) if scale is None else np.asarray(scale, self.dtype)[zc, xc]
        np.add.at(field, (zc, xc), v * ((1 - tz) * (1 - tx) * fac))
        np.add.at(field, (zc, xc + 1), v * ((1 - tz) * tx * fac))
        np.add.at(field, (zc + 1, xc), v * (tz * (1 - tx) * fac))
        np.add.at(field, (zc + 1, xc + 1), v * (tz * tx * fac))

    def inject(self, comp, values):
        """ertm/propagation.py:764-774: pressure also enters the strain."""
        self._scatter_scaled({"p": self.p, "vx": self.vx, "vz": self.vz}[comp], values, comp)
        if comp == "p":
            sc = 1.0 / np.maximum(np.abs(self.kr), 1e-20)
            self._scatter_scaled(self.eps, values, comp, sc)

    def image_fields(self, sep="stress"):
        return {"P": self.g.crop(self.p)}
