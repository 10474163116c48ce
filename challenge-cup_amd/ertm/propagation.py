"""Forward and adjoint wavefield extrapolation on the GPU (ertm/propagation.py).

The public surface is the reference's: ``PropagationSpec``, ``check_stability``,
``QOptions``, the three propagator classes with the reference method set
(``set_source``, ``set_receivers``, ``step``, ``sample``, ``inject``,
``image_fields``, ``state``, ``load_state``, ``state_bytes``, ``reset``,
field attributes), ``build_propagator``, ``source_wavelet``, ``run_forward``
and ``forward_shot``.

Underneath, every propagator is a ``device.Core``: its fields, CPML strips
and medium live in HBM and every update is a HIP kernel of libertm_hip.so.
Field attributes (``prop.p`` ...) and ``state()`` return host copies.  The
engine computes in float32.  A float64 spec (``SolverConfig.dtype``, which
the reference honours with float64 wavefields, ertm/propagation.py:52, :207)
is refused with ``SolverError`` rather than computed at a narrower precision
than asked for.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any

import numpy as np

from . import _lib as L
from . import device as dev
from . import numerics as num
from .config import SolverConfig
from .exceptions import SolverError
from .types import ElasticModel, ReceiverSpec, ShotRecord, SourceSpec
from .viscoelastic import relaxation_tables


@dataclass(frozen=True)
class PropagationSpec:
    """Discretisation of one extrapolation (ertm/propagation.py:44-95)."""

    dt: float
    nt: int
    order: int = 8
    npml: int = 20
    dtype: str = "float32"
    boundary: str = "cpml"

    def __post_init__(self) -> None:
        if self.dt <= 0:
            raise ValueError(f"dt must be positive, got {self.dt}")
        if self.nt < 1:
            raise ValueError(f"nt must be at least 1, got {self.nt}")
        if self.order not in num.SUPPORTED_ORDERS:
            raise ValueError(f"order must be one of {num.SUPPORTED_ORDERS}, got {self.order}")
        if self.npml < 1:
            raise ValueError(f"npml must be >= 1, got {self.npml}")
        if self.boundary not in {"cpml", "free", "none"}:
            raise ValueError(f"boundary must be cpml/free/none, got {self.boundary!r}")

    @property
    def numpy_dtype(self) -> np.dtype:
        return np.dtype(self.dtype)

    @property
    def time(self) -> np.ndarray:
        return np.arange(self.nt, dtype=np.float64) * self.dt

    @classmethod
    def from_config(cls, config: SolverConfig, model: ElasticModel | None = None, *,
                    npml: int | None = None) -> "PropagationSpec":
        """Adaptive npml = min(20, max(8, min(nx, nz)//8)) unless given (:84-87)."""
        cells = 20 if model is None else int(min(20, max(8, min(model.grid.nx, model.grid.nz) // 8)))
        return cls(dt=float(config.dt), nt=int(config.nt), order=int(config.order),
                   npml=int(npml if npml is not None else cells), dtype=str(config.dtype),
                   boundary=str(config.boundary))


def stability_report(model: ElasticModel, spec: PropagationSpec) -> dict[str, float]:
    """ertm/propagation.py:98-110."""
    vmax = float(np.max(model.vp))
    vmin = float(np.min(model.vp))
    limit = num.cfl_dt(vmax, model.grid.dx, model.grid.dz, spec.order)
    return {"vmax": vmax, "vmin": vmin, "dt": spec.dt, "dt_max": limit,
            "cfl_ratio": float(spec.dt / limit) if limit > 0 else float("inf"),
            "ppw_vmin": num.points_per_wavelength(vmin, 1.0 / (4.0 * spec.dt), model.grid.dx)}


def check_stability(model: ElasticModel, spec: PropagationSpec, *, safety: float = 0.55) -> None:
    """Reject dt above safety * the CFL bound (ertm/propagation.py:113-127)."""
    vmax = float(np.max(model.vp))
    limit = num.cfl_dt(vmax, model.grid.dx, model.grid.dz, spec.order, safety=safety)
    if spec.dt > limit:
        hint = num.cfl_dt(vmax, model.grid.dx, model.grid.dz, spec.order, safety=0.5)
        raise SolverError(
            f"time step dt={spec.dt:.3e} s exceeds the CFL limit {limit:.3e} s "
            f"(vmax={vmax:.0f} m/s, dx={model.grid.dx}, dz={model.grid.dz}, order={spec.order}); "
            f"resample the data to dt <= {hint:.3e} s first")


def check_hard_stability(model: ElasticModel, spec: PropagationSpec) -> None:
    """The scheme's true bound (safety 1): used when the reference guard is waived."""
    check_stability(model, spec, safety=1.0)


@dataclass(frozen=True)
class QOptions:
    """Viscoacoustic GSLS options (ertm/propagation.py:130-159)."""

    enabled: bool = False
    f_ref: float = 30.0
    n_mechanisms: int = 3
    band_ratio: float = 1.5
    compensation: bool = False
    alpha_c: float = 1.0
    amplitude_cap: float = 50.0
    lowpass_interval: int = 20
    f_max: float | None = None
    gain_budget_db: float = 40.0

    def validate(self) -> None:
        if self.f_ref <= 0:
            raise ValueError("q.f_ref must be positive")
        if int(self.n_mechanisms) < 1:
            raise ValueError("q.n_mechanisms must be at least 1")
        if int(self.n_mechanisms) > L.MAX_MECH:
            raise ValueError(f"q.n_mechanisms must be at most {L.MAX_MECH}")
        if self.band_ratio <= 1.0:
            raise ValueError("q.band_ratio must be > 1")
        if not 0.0 <= self.alpha_c <= 1.0:
            raise ValueError("q.alpha_c must lie in [0, 1]")
        if self.amplitude_cap < 1.0:
            raise ValueError("q.amplitude_cap must be >= 1")
        if self.f_max is not None and self.f_max <= 0:
            raise ValueError("q.f_max must be positive or None")
        if not 0.0 < self.gain_budget_db <= 200.0:
            raise ValueError("q.gain_budget_db must lie in (0, 200]")


class CPMLState:
    """CPML bookkeeping visible to callers (stands where CPML2D sits).

    Memories exist only in the absorbing strips; like the reference they
    count as allocated from the first step on (``memory_bytes`` is 0 before),
    which is what the checkpoint budget rule of ertm/imaging.py:121-124 sees.
    """

    def __init__(self, core: dev.Core):
        self._core = core
        self.touched = False
        self.x = _Profile(core.prof["ax"], core.prof["bx"], core.wx)
        self.z = _Profile(core.prof["az"], core.prof["bz"], max(core.wzt, core.wzb))

    def memory_bytes(self) -> int:
        return self._core.psi_bytes() if self.touched else 0

    @property
    def _memory(self) -> dict:
        if not self.touched:
            return {}
        return {k: self._core.expand_psi(k, v.cpu().numpy())
                for k, v in self._core.psi_views().items()}


@dataclass
class _Profile:
    a: np.ndarray
    b: np.ndarray
    npml: int


def check_compute_dtype(spec) -> None:
    """The GPU engine's kernels compute in float32 only; a float64 request is
    an error, not a silent downgrade."""
    if np.dtype(spec.dtype) != np.float32:
        raise SolverError(
            f"dtype={spec.dtype!r}: the GPU engine computes in float32 only; a float64 run "
            "needs the reference's NumPy engine")


class Propagator2D:
    """Common machinery of the GPU propagators (ertm/propagation.py:190-391)."""

    name = "propagator-2d"
    components: tuple[str, ...] = ()
    _engine = ""

    def __init__(self, model: ElasticModel, spec: PropagationSpec) -> None:
        check_compute_dtype(spec)
        if not dev.is_trusted(model):  # batches validate their models once
            model.validate()
        self.model = model
        self.spec = spec
        self.grid = model.grid
        self.dt = float(spec.dt)
        self.nt = int(spec.nt)
        self.order = int(spec.order)
        self.coeffs = num.staggered_coefficients(self.order)
        self.pad = num.ghost_width(self.order)
        self.dtype = np.dtype(np.float32)
        self.compute_dtype = "float32"
        self.shape = (self.grid.nz + 2 * self.pad, self.grid.nx + 2 * self.pad)
        self.dx, self.dz = float(self.grid.dx), float(self.grid.dz)
        self.x0, self.z0 = float(self.grid.x0), float(self.grid.z0)
        self.step_index = 0
        self._source = None
        self._source_index: tuple[int, int] | None = None
        self._mechanism = "explosive"
        self._receivers: ReceiverSpec | None = None
        self._batches: dict = {}

    # ------------------------------------------------------------ set-up
    def _make_core(self, media: dict, nmech: int = 0, media_key: tuple | None = None) -> None:
        vmax = float(np.max(self.model.vp))
        self._core = dev.Core(self._engine, self.grid.nz, self.grid.nx, self.dx, self.dz, self.x0,
                              self.z0, self.order, self.dt, self.spec.npml, self.spec.boundary,
                              vmax, media, nmech, media_key)
        self.pml = CPMLState(self._core)

    def _staged(self, array: np.ndarray) -> np.ndarray:
        return dev.edge_pad(array, self.pad, np.float32)

    def _invalidate(self) -> None:
        for b in self._batches.values():
            b.close()
        self._batches.clear()

    def _batch(self, key="step", recv=None, img=None) -> dev.Batch:
        b = self._batches.get(key)
        if b is None:
            b = dev.Batch([self._core], [recv] if recv is not None else None,
                          [img] if img is not None else None)
            self._batches[key] = b
        return b

    # ------------------------------------------------------------ geometry
    def node_index(self, x: float, z: float) -> tuple[int, int]:
        return self._core.node_index(x, z)

    def fractional_index(self, x, z, *, x_offset: float = 0.0, z_offset: float = 0.0):
        xf = (np.asarray(x, dtype=np.float64) - self.x0) / self.dx + self.pad + x_offset
        zf = (np.asarray(z, dtype=np.float64) - self.z0) / self.dz + self.pad + z_offset
        return zf, xf

    def set_source(self, source: SourceSpec) -> None:
        self._source_index = self.node_index(source.x, source.z)
        self._source = source
        self._mechanism = str(getattr(source, "mechanism", "explosive"))
        self._core.src = self._source_index
        self._core.mechanism = L.MECH_CODE.get(self._mechanism, L.SRC_EXPLOSIVE)
        self._invalidate()

    def set_receivers(self, receivers: ReceiverSpec) -> None:
        self._receivers = receivers
        self._invalidate()

    # ------------------------------------------------------------ fields
    def _field_names(self) -> tuple[str, ...]:
        raise NotImplementedError

    def _crop(self, field: np.ndarray) -> np.ndarray:
        lo = self.pad
        return field[lo : lo + self.grid.nz, lo : lo + self.grid.nx]

    def _host(self, name: str) -> np.ndarray:
        c = self._core
        if name == "r":
            return np.stack([c.slab(4 + l).cpu().numpy() for l in range(c.nmech)])
        return c.slab(c.field_index(name)).cpu().numpy()

    def __getattr__(self, name: str):
        # field attributes (p, vx, vz, sxx, ...) are host copies of device slabs
        core = self.__dict__.get("_core")
        if core is not None and (name in core.field_names or (name == "r" and core.nmech)):
            return self._host(name)
        raise AttributeError(name)

    def set_field(self, name: str, values: np.ndarray) -> None:
        """Overwrite one field of the current state from a host array."""
        import torch

        c = self._core
        arr = torch.from_numpy(np.ascontiguousarray(values, dtype=np.float32)).to("cuda")
        if name == "r":
            for l in range(c.nmech):
                c.slab(4 + l).copy_(arr[l])
        else:
            c.slab(c.field_index(name)).copy_(arr)

    # ------------------------------------------------------------ state I/O
    def state(self) -> dict[str, np.ndarray]:
        """Host copy of fields + CPML memories + step counter (ertm/propagation.py:283-289)."""
        st = {name: self._host(name) for name in self._field_names()}
        if self.pml.touched:
            for k, v in self._core.psi_views().items():
                st[f"pml:{k}"] = v.cpu().numpy()
        st["_step"] = np.array(self.step_index, dtype=np.int64)
        return st

    def load_state(self, state: dict[str, np.ndarray]) -> None:
        """ertm/propagation.py:291-299: CPML memories are restored only once
        they exist (after a first step); before that they restart from zero."""
        for name in self._field_names():
            self.set_field(name, state[name])
        views = self._core.psi_views()
        for k, v in views.items():
            if not self.pml.touched:
                v.zero_()
            elif f"pml:{k}" in state:
                import torch

                v.copy_(torch.from_numpy(np.ascontiguousarray(state[f"pml:{k}"])).to("cuda"))
        self.step_index = int(np.asarray(state["_step"]).item())

    def state_bytes(self) -> int:
        """Fields + CPML memory bytes (ertm/propagation.py:301-306)."""
        c = self._core
        return int(c.nfields * c.nzp * c.nxp * 4 + self.pml.memory_bytes())

    def reset(self) -> None:
        self._core.state.zero_()
        self._core.cur = 0
        self.pml.touched = False
        self.step_index = 0

    # ------------------------------------------------------------ stepping
    def step(self, source_amplitude: float = 0.0) -> None:
        """Advance one step; ``source_amplitude`` is the wavelet value now."""
        c = self._core
        amp = float(source_amplitude) if self._source_index is not None else 0.0
        self._batch().step(c.cur, [amp])
        c.cur ^= 1
        self.pml.touched = True
        self._after_step()
        self.step_index += 1

    def _after_step(self) -> None:
        pass

    # ------------------------------------------------------------ receivers
    def _default_components(self) -> tuple[str, ...]:
        return ("p",)

    def sample(self, components: tuple[str, ...] | None = None) -> dict[str, np.ndarray]:
        """Bilinear samples at the receivers, one (nrec,) float32 array per component."""
        import torch

        if self._receivers is None:
            return {}
        names = tuple(components or self._default_components())
        key = ("sample", names)
        b = self._batches.get(key)
        if b is None:
            recs = [torch.zeros((1, self._receivers.nrec), dtype=torch.float32, device="cuda")
                    for _ in names]
            tab = dev.RecvTables()
            tab.build_sampling(self._core, self._receivers.x, self._receivers.z, names)
            tab.set_records(recs)
            b = self._batch(key, recv=tab)
            b.records = recs
        b.sample(self._core.cur, [0])
        return {n: r[0].cpu().numpy() for n, r in zip(names, b.records)}

    def inject(self, component: str, values: np.ndarray) -> None:
        """Transposed bilinear scatter of ``values`` (ordered, like np.add.at)."""
        import torch

        if self._receivers is None:
            raise SolverError("receivers must be set before injecting data")
        key = ("inject", component)
        b = self._batches.get(key)
        if b is None:
            rec = torch.zeros((1, self._receivers.nrec), dtype=torch.float32, device="cuda")
            tab = dev.RecvTables()
            tab.build_injection(self._core, self._receivers.x, self._receivers.z, (component,),
                                self._k_relaxed_host())
            tab.set_records([rec])
            b = self._batch(key, recv=tab)
            b.records = [rec]
        vals = np.asarray(values, dtype=np.float32).reshape(1, -1)
        b.records[0].copy_(torch.from_numpy(vals).to("cuda"))
        b.inject(self._core.cur, [0])

    def _k_relaxed_host(self):
        return None

    # ------------------------------------------------------------ imaging
    def image_fields(self, separation: str = "stress") -> dict[str, np.ndarray]:
        return {"P": self._image_field("P", separation)}

    def _image_field(self, wave: str, separation: str) -> np.ndarray:
        import torch

        key = ("imgfield", wave, separation)
        b = self._batches.get(key)
        if b is None:
            rbuf = torch.zeros((1, self.grid.nz, self.grid.nx), dtype=torch.float32, device="cuda")
            img = dev.ImageBuffers(self.grid.nz, self.grid.nx, separation, wave, (), 0, False,
                                   rbuf=rbuf, nslots=1)
            b = self._batch(key, img=img)
            b.rbuf = rbuf
        b.image(self._core.cur, [0], L.IMG_EXTRACT)
        return b.rbuf[0].cpu().numpy()

    @property
    def max_amplitude(self) -> float:
        return dev.absmax(self._core.slab(0).contiguous())

    # ------------------------------------------------------------ hot path
    def run_steps(self, wavelet: np.ndarray, components: tuple[str, ...] | None = None,
                  record_every: int = 1) -> dict[str, np.ndarray]:
        """nt steps with sampling after each, one C-side loop (no host sync per step)."""
        import torch

        nt = int(len(wavelet))
        names = tuple(components or self._default_components())
        nrec = self._receivers.nrec if self._receivers is not None else 0
        recs = [torch.zeros((nt, max(nrec, 1)), dtype=torch.float32, device="cuda")
                for _ in names]
        if nrec and names:
            tab = dev.RecvTables()
            tab.build_sampling(self._core, self._receivers.x, self._receivers.z, names)
            tab.set_records(recs)
            b = dev.Batch([self._core], [tab])
        else:
            b = dev.Batch([self._core])
        w = np.asarray(wavelet, dtype=np.float32)
        if self._source_index is None:
            w = np.zeros_like(w)
        kw = self._run_kwargs()
        self._core.cur = b.run(0, 0, nt, self._core.cur, wavelets=[w], do_sample=bool(nrec),
                               step_index0=self.step_index, **kw)
        self.pml.touched = self.pml.touched or nt > 0
        self.step_index += nt
        self._after_run(kw)
        b.close()
        out = {n: r.cpu().numpy()[:, :nrec] for n, r in zip(names, recs)}
        if record_every > 1:
            for n in names:
                keep = (np.arange(nt) % record_every) == 0
                out[n][~keep] = 0.0
        return out

    def _run_kwargs(self) -> dict:
        return {}

    def _after_run(self, kw: dict) -> None:
        pass


class AcousticPropagator(Propagator2D):
    """Velocity-pressure acoustic engine (ertm/propagation.py:397-506)."""

    name = "acoustic-velocity-pressure"
    components = ("p", "vx", "vz")
    _engine = "acoustic"

    def __init__(self, model: ElasticModel, spec: PropagationSpec) -> None:
        super().__init__(model, spec)

        def build():
            rho = self._staged(model.rho)
            vp = self._staged(model.vp)
            kappa = (rho * vp * vp).astype(np.float32)
            irho = (1.0 / rho).astype(np.float32)
            out = {"kappa": kappa, "irho": irho}
            pal = dev.medium_palette([irho, kappa])
            if pal is not None:
                out["midx"], out["mtab"], out["npal"] = pal
            return out

        key = dev.media_digest("acoustic", self.pad, model.vp, model.rho)
        self._make_core(dev.host_media(key, build), media_key=key)

    def _field_names(self) -> tuple[str, ...]:
        return ("p", "vx", "vz")


class ViscoacousticPropagator(Propagator2D):
    """GSLS memory-variable viscoacoustic engine (ertm/propagation.py:512-798).

    ``q_options.compensation`` flips delta_l -> -alpha_l delta_l (Q-inverse
    backward propagation) and enables the stabiliser (amplitude cap + periodic
    k-space low-pass), all on the GPU.
    """

    name = "viscoacoustic-gsls"
    components = ("p", "vx", "vz")
    _engine = "viscoacoustic"

    def __init__(self, model: ElasticModel, spec: PropagationSpec, *,
                 q_options: QOptions | None = None) -> None:
        super().__init__(model, spec)
        self.q_options = q_options or QOptions(enabled=True)
        self.q_options.validate()
        qo = self.q_options
        rho = self._staged(model.rho)
        vp = self._staged(model.vp)
        qp = self._staged(model.qp)
        k0 = (rho * vp * vp).astype(np.float32)
        irho = (1.0 / rho).astype(np.float32)
        n = int(qo.n_mechanisms)
        self.n_mechanisms = n
        tab = relaxation_tables(qp, qo.f_ref, n, qo.band_ratio, self.dt, self.nt,
                                compensation=qo.compensation, alpha_c=qo.alpha_c, f_max=qo.f_max,
                                gain_budget_db=qo.gain_budget_db)
        idx = tab["index"]
        self.k_relaxed = (k0 / tab["modulus_ref"][idx]).astype(np.float32)
        a_tab = np.exp(-self.dt / tab["tau"]).astype(np.float32)
        self.q_fit_error = float(tab["fit_error"])
        if qo.compensation:
            self.alpha_mechanisms = tuple(float(v) for v in np.mean(tab["alpha"], axis=0))
        else:
            self.alpha_mechanisms = tuple(float(qo.alpha_c) for _ in range(n))
        self.f_peak_mechanisms = tuple(float(v) for v in np.mean(tab["f_peak"], axis=0))
        media = {"kappa": self.k_relaxed, "irho": irho, "qidx": idx.astype(np.int32),
                 "q_a": a_tab.ravel(), "q_delta": tab["delta"].ravel().astype(np.float64),
                 "q_tau": tab["tau"].ravel().astype(np.float64)}
        self._tables = tab
        self._make_core(media, nmech=n)
        self._ctl = None
        self._clipped = 0
        self._lowpass_steps = 0

    def _field_names(self) -> tuple[str, ...]:
        return ("p", "vx", "vz", "eps", "r")

    def _k_relaxed_host(self):
        return self.k_relaxed

    def _ctl_buf(self):
        import torch

        if self._ctl is None:
            self._ctl = torch.zeros(4, dtype=torch.float64, device="cuda")
        return self._ctl

    def _after_step(self) -> None:
        if self.q_options.compensation:
            qo = self.q_options
            self._batch().stabilize(self._core.cur, self.step_index, qo.amplitude_cap,
                                    qo.lowpass_interval, self._ctl_buf())

    def _run_kwargs(self) -> dict:
        if not self.q_options.compensation:
            return {}
        qo = self.q_options
        return dict(stabilize=True, cap=qo.amplitude_cap, lp_interval=qo.lowpass_interval,
                    ctl=self._ctl_buf())

    def reset(self) -> None:
        super().reset()
        if self._ctl is not None:
            self._ctl.zero_()

    @property
    def quality_summary(self) -> dict[str, Any]:
        ctl = self._ctl.cpu().numpy() if self._ctl is not None else np.zeros(4)
        qo = self.q_options
        return {
            "mechanisms": int(self.n_mechanisms),
            "q_fit_error": float(self.q_fit_error),
            "compensation": bool(qo.compensation),
            "alpha_c": float(qo.alpha_c),
            "alpha_mechanisms": list(self.alpha_mechanisms),
            "f_peak_mechanisms_hz": [round(v, 1) for v in self.f_peak_mechanisms],
            "f_max": None if qo.f_max is None else float(qo.f_max),
            "gain_budget_db": float(qo.gain_budget_db),
            "clipped": int(ctl[1]),
            "lowpass_steps": int(ctl[2]),
        }


class ElasticPropagator(Propagator2D):
    """Velocity-stress P-SV engine (ertm/propagation.py:804-983)."""

    name = "elastic-velocity-stress"
    components = ("vx", "vz", "sxx", "szz", "sxz", "p")
    _engine = "elastic"

    def __init__(self, model: ElasticModel, spec: PropagationSpec) -> None:
        super().__init__(model, spec)

        def build():
            rho = self._staged(model.rho)
            vp = self._staged(model.vp)
            vs = self._staged(model.vs)
            mu = (rho * vs * vs).astype(np.float32)
            lam = (rho * (vp * vp - 2.0 * vs * vs)).astype(np.float32)
            irho = (1.0 / rho).astype(np.float32)
            out = {"lam": lam, "mu": mu, "irho": irho}
            pal = dev.medium_palette([irho, mu, lam])
            if pal is not None:
                out["midx"], out["mtab"], out["npal"] = pal
            return out

        key = dev.media_digest("elastic", self.pad, model.vp, model.vs, model.rho)
        self._make_core(dev.host_media(key, build), media_key=key)

    def _field_names(self) -> tuple[str, ...]:
        return ("vx", "vz", "sxx", "szz", "sxz")

    def _default_components(self) -> tuple[str, ...]:
        return ("vx", "vz")

    def image_fields(self, separation: str = "stress") -> dict[str, np.ndarray]:
        mode = str(separation).lower()
        if mode not in ("stress", "divcurl"):
            raise SolverError(f"unknown P/S separation {separation!r} (stress/divcurl)")
        return {"P": self._image_field("P", mode), "S": self._image_field("S", mode)}

    def pressure_field(self) -> np.ndarray:
        return np.float32(0.5) * (self.sxx + self.szz)

    def shear_field(self) -> np.ndarray:
        return self.sxz


def build_propagator(model: ElasticModel, spec: PropagationSpec, *, engine: str = "acoustic",
                     q_options: QOptions | None = None) -> Propagator2D:
    """ertm/propagation.py:986-1001."""
    engine = str(engine).lower()
    if engine == "acoustic":
        return AcousticPropagator(model, spec)
    if engine == "viscoacoustic":
        return ViscoacousticPropagator(model, spec, q_options=q_options)
    if engine == "elastic":
        return ElasticPropagator(model, spec)
    raise SolverError(f"unknown engine {engine!r} (acoustic/viscoacoustic/elastic)")


def source_wavelet(source: SourceSpec, dt: float, nt: int) -> np.ndarray:
    """Wavelet samples on the time grid (ertm/propagation.py:1004-1015)."""
    t = np.arange(nt, dtype=np.float64) * dt
    if source.kind == "ricker":
        wave = num.ricker(t, source.f0, source.delay)
    elif source.kind == "gaussian":
        wave = num.gaussian_pulse(t, source.f0, source.delay)
    elif source.kind == "custom":
        raise SolverError("a custom source needs an explicit wavelet array")
    else:
        raise SolverError(f"unsupported source kind: {source.kind}")
    return (float(source.amplitude) * wave).astype(np.float32)


def run_forward(model: ElasticModel, source: SourceSpec, receivers: ReceiverSpec,
                spec: PropagationSpec, *, engine: str = "acoustic",
                components: tuple[str, ...] | None = None, wavelet: np.ndarray | None = None,
                propagator: Propagator2D | None = None, record_every: int = 1,
                q_options: QOptions | None = None) -> dict[str, np.ndarray]:
    """Forward modelling, {component: (nt, nrec) float32} (ertm/propagation.py:1018-1052)."""
    prop = propagator or build_propagator(model, spec, engine=engine, q_options=q_options)
    prop.set_source(source)
    prop.set_receivers(receivers)
    if wavelet is None:
        wavelet = source_wavelet(source, spec.dt, spec.nt)
    names = components or prop.components[:1]
    return prop.run_steps(np.asarray(wavelet, dtype=np.float32), tuple(names), record_every)


def batch_run_kwargs(props: list[Propagator2D]) -> dict:
    """ertm_run options of several propagators run together: the viscoacoustic
    stabiliser needs one control record (4 doubles) per propagator in one
    contiguous buffer."""
    kws = [p._run_kwargs() for p in props]
    if not kws[0]:
        return {}
    import torch

    kw = dict(kws[0])
    if len(props) > 1:
        kw["ctl"] = torch.cat([k["ctl"] for k in kws])
    return kw


def batch_after_run(props: list[Propagator2D], kw: dict) -> None:
    """Hand the batch's control records back to each propagator."""
    if len(props) > 1 and "ctl" in kw:
        for k, p in enumerate(props):
            p._run_kwargs()["ctl"].copy_(kw["ctl"][4 * k: 4 * k + 4])


def run_steps_batch(props: list[Propagator2D], wavelets: list[np.ndarray],
                    components: tuple[str, ...]) -> list[dict[str, np.ndarray]]:
    """``run_steps`` of several propagators advanced together (one launch per
    step for all of them); each result equals its propagator's ``run_steps``."""
    import torch

    nts = {int(len(w)) for w in wavelets}
    if len(nts) != 1 or len(props) != len(wavelets):
        raise SolverError("batched propagators need one wavelet each, all of one length")
    nt = nts.pop()
    steps = {p.step_index for p in props}
    curs = {p._core.cur for p in props}
    if len(steps) != 1 or len(curs) != 1:
        raise SolverError("batched propagators must be at the same step")
    names = tuple(components)
    tabs, recs_all, ws = [], [], []
    for p, w in zip(props, wavelets):
        nrec = p._receivers.nrec if p._receivers is not None else 0
        if not nrec or not names:
            raise SolverError("batched forward modelling needs receivers and components")
        recs = [torch.zeros((nt, nrec), dtype=torch.float32, device="cuda") for _ in names]
        tab = dev.RecvTables()
        tab.build_sampling(p._core, p._receivers.x, p._receivers.z, names)
        tab.set_records(recs)
        tabs.append(tab)
        recs_all.append(recs)
        w = np.asarray(w, dtype=np.float32)
        ws.append(w if p._source_index is not None else np.zeros_like(w))
    b = dev.Batch([p._core for p in props], tabs)
    kw = batch_run_kwargs(props)
    cur = b.run(0, 0, nt, curs.pop(), wavelets=ws, do_sample=True, step_index0=steps.pop(), **kw)
    b.close()
    batch_after_run(props, kw)
    out = []
    for p, recs in zip(props, recs_all):
        p._core.cur = cur
        p.pml.touched = p.pml.touched or nt > 0
        p.step_index += nt
        out.append({n: r.cpu().numpy() for n, r in zip(names, recs)})
    return out


def forward_shots(model: ElasticModel | list[ElasticModel], sources: list[SourceSpec],
                  receivers: ReceiverSpec | list[ReceiverSpec], spec: PropagationSpec, *,
                  engine: str = "acoustic", components: tuple[str, ...] | None = None,
                  q_options: QOptions | None = None) -> list[ShotRecord]:
    """``forward_shot`` for many sources, advanced together; every record
    equals the one ``forward_shot`` returns for that source alone."""
    n = len(sources)
    models = model if isinstance(model, list) else [model] * n
    recvs = receivers if isinstance(receivers, list) else [receivers] * n
    names = components or (("p",) if engine in ("acoustic", "viscoacoustic") else ("vx", "vz", "sxz"))
    props = []
    for m, s, r in zip(models, sources, recvs):
        p = build_propagator(m, spec, engine=engine, q_options=q_options)
        p.set_source(s)
        p.set_receivers(r)
        props.append(p)
    data = run_steps_batch(props, [source_wavelet(s, spec.dt, spec.nt) for s in sources],
                           tuple(names))
    return [_shot_record(d, names, spec, engine, s, r) for d, s, r in zip(data, sources, recvs)]


def _shot_record(data, names, spec, engine, source, receivers) -> ShotRecord:
    first = names[0]
    t = np.arange(data[first].shape[0], dtype=np.float64) * spec.dt
    meta: dict[str, Any] = {
        "engine": engine, "solver": f"{engine}-velocity-stress", "dt": spec.dt, "nt": int(t.size),
        "order": spec.order, "npml": spec.npml, "components": list(names), "placeholder": False,
        "method": "finite-difference two-way wave equation", "device": "hip-gfx950",
        "compute_dtype": "float32",
    }
    shot = ShotRecord(data=data[first].astype(np.float32), dt=float(spec.dt), t=t,
                      receivers=receivers, source=source, meta=meta,
                      components={n: data[n].astype(np.float32) for n in names})
    shot.validate()
    return shot


def forward_shot(model: ElasticModel, source: SourceSpec, receivers: ReceiverSpec,
                 spec: PropagationSpec, *, engine: str = "acoustic",
                 components: tuple[str, ...] | None = None,
                 q_options: QOptions | None = None) -> ShotRecord:
    """Forward shot as a ShotRecord (ertm/propagation.py:1055-1093)."""
    names = components or (("p",) if engine in ("acoustic", "viscoacoustic") else ("vx", "vz", "sxz"))
    data = run_forward(model, source, receivers, spec, engine=engine, components=names,
                       q_options=q_options)
    return _shot_record(data, names, spec, engine, source, receivers)
