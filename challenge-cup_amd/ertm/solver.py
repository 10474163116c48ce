"""Forward-modelling facade (ertm/solver.py:34-154).

``FiniteDifferenceSolver`` turns a ``SolverConfig`` into a ``PropagationSpec``
(with the configured npml, unlike migration which uses the adaptive one --
the reference's behaviour), checks stability and runs ``forward_shot`` on the
GPU engine.  It satisfies the ``ForwardSolver`` protocol.
"""

from __future__ import annotations

import numpy as np

from .config import GpuConfig, QConfig, SolverConfig
from .exceptions import SolverError
from .propagation import (
    PropagationSpec,
    QOptions,
    check_hard_stability,
    check_stability,
    forward_shot,
    forward_shots,
    source_wavelet,
    stability_report,
)
from .types import ElasticModel, ReceiverSpec, ShotRecord, SourceSpec


class FiniteDifferenceSolver:
    """Velocity-stress finite-difference forward solver on the GPU."""

    name = "fd-two-way"

    def __init__(self, config: SolverConfig, q_config: QConfig | None = None,
                 gpu_config: GpuConfig | None = None) -> None:
        self.config = config
        self.config.validate()
        self.q_config = q_config or QConfig()
        self.gpu_config = gpu_config or GpuConfig()

    def q_options(self, *, compensation: bool = False) -> QOptions | None:
        """ertm/solver.py:44-59."""
        if self.config.engine != "viscoacoustic":
            return None
        q = self.q_config
        return QOptions(enabled=True, f_ref=float(q.f_ref), n_mechanisms=int(q.n_mechanisms),
                        band_ratio=float(q.band_ratio), compensation=bool(compensation),
                        alpha_c=float(q.alpha_c), amplitude_cap=float(q.amplitude_cap),
                        lowpass_interval=int(q.lowpass_interval),
                        f_max=None if q.f_max is None else float(q.f_max),
                        gain_budget_db=float(q.gain_budget_db))

    def spec(self, model: ElasticModel) -> PropagationSpec:
        """Configured npml, CFL-checked (ertm/solver.py:62-65)."""
        spec = PropagationSpec.from_config(self.config, model, npml=self.config.npml)
        if self.gpu_config.strict_cfl:
            check_stability(model, spec)
        else:
            check_hard_stability(model, spec)
        return spec

    def components(self, n_components: int = 1) -> tuple[str, ...]:
        """ertm/solver.py:67-79."""
        if self.config.engine in ("acoustic", "viscoacoustic"):
            return ("p",)
        if int(n_components) >= 4:
            return ("p", "vx", "vz", "sxz")
        return ("vz",)

    def report(self, model: ElasticModel) -> dict[str, float]:
        return stability_report(model, self.spec(model))

    def forward(self, model: ElasticModel, source: SourceSpec, receivers: ReceiverSpec, *,
                components: tuple[str, ...] | None = None,
                wavelet: np.ndarray | None = None) -> ShotRecord:
        """Model one shot (ertm/solver.py:85-131).  Like the reference, a
        ``wavelet`` argument is accepted and not used: the source's own
        wavelet drives the shot."""
        try:
            model.validate()
        except ValueError as exc:
            raise SolverError(f"invalid forward model: {exc}") from exc
        spec = self.spec(model)
        names = components or self.components()
        record = forward_shot(model, source, receivers, spec, engine=self.config.engine,
                              components=names, q_options=self.q_options())
        record.meta.update({
            "placeholder": False, "solver": self.name, "engine": self.config.engine,
            "order": self.config.order, "boundary": self.config.boundary,
            "npml": self.config.npml, "dt": spec.dt, "nt": spec.nt, "components": list(names),
        })
        return record

    def forward_batch(self, model: ElasticModel | list[ElasticModel], sources: list[SourceSpec],
                      receivers: ReceiverSpec | list[ReceiverSpec], *,
                      components: tuple[str, ...] | None = None) -> list[ShotRecord]:
        """``forward`` for many sources, advanced together with one launch per
        step; every record equals what ``forward`` returns for that source."""
        models = model if isinstance(model, list) else [model] * len(sources)
        if len(models) != len(sources):
            raise SolverError("one model per source expected")
        specs = []
        for m in models:
            try:
                m.validate()
            except ValueError as exc:
                raise SolverError(f"invalid forward model: {exc}") from exc
            specs.append(self.spec(m))
        if len({(s.nt, s.npml) for s in specs}) != 1:
            raise SolverError("batched forward modelling needs one nt and npml")
        names = components or self.components()
        records = forward_shots(models, list(sources), receivers, specs[0],
                                engine=self.config.engine, components=names,
                                q_options=self.q_options())
        for rec, spec in zip(records, specs):
            rec.meta.update({
                "placeholder": False, "solver": self.name, "engine": self.config.engine,
                "order": self.config.order, "boundary": self.config.boundary,
                "npml": self.config.npml, "dt": spec.dt, "nt": spec.nt,
                "components": list(names),
            })
        return records

    def wavelet(self, source: SourceSpec, model: ElasticModel) -> np.ndarray:
        spec = self.spec(model)
        return source_wavelet(source, spec.dt, spec.nt)

    def adjoint(self, record: ShotRecord, model: ElasticModel) -> np.ndarray:
        raise SolverError("adjoint propagation runs inside RTMImager.migrate")


def build_solver(config: SolverConfig, q_config: QConfig | None = None,
                 gpu_config: GpuConfig | None = None) -> FiniteDifferenceSolver:
    """ertm/solver.py:149-154."""
    return FiniteDifferenceSolver(config, q_config, gpu_config)
