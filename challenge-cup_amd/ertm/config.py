"""Hot-path configuration sections (ertm/config.py:283-436) plus the engine's
own ``GpuConfig``.

``SolverConfig``, ``ImagingConfig`` and ``QConfig`` keep the reference's field
names, defaults and validation rules so that a reference configuration object
(or its JSON section) drives this engine unchanged.  ``GpuConfig`` holds what
only the MI355X engine knows about: CFL strictness, memory budgets, fused
multi-mode imaging and batching.
"""

from __future__ import annotations

from dataclasses import dataclass, fields

from .exceptions import ConfigError

_ORDERS = {2, 4, 6, 8}
_BOUNDARIES = {"cpml", "free", "none"}
_MODES = {"pp", "ps", "sp", "ss"}
_DTYPES = {"float32", "float64"}
_ENGINES = {"acoustic", "elastic", "viscoacoustic"}
_CONDITIONS = {"zero_lag", "normalized", "excitation"}
_SEPARATIONS = {"divcurl", "stress"}


def _from_mapping(cls, data: dict):
    """Build a section from a dict, ignoring unknown keys (like the reference loader)."""
    names = {f.name for f in fields(cls)}
    return cls(**{k: v for k, v in dict(data).items() if k in names})


@dataclass(frozen=True)
class SolverConfig:
    """ertm/config.py:283-307."""

    nt: int = 1200
    dt: float = 5.0e-4
    order: int = 8
    boundary: str = "cpml"
    dtype: str = "float32"
    engine: str = "acoustic"
    npml: int = 20

    def validate(self) -> None:
        if self.nt < 1:
            raise ConfigError("solver.nt must be at least 1")
        if self.dt <= 0:
            raise ConfigError("solver.dt must be positive")
        if self.order not in _ORDERS:
            raise ConfigError(f"solver.order must be one of {sorted(_ORDERS)}")
        if self.boundary not in _BOUNDARIES:
            raise ConfigError(f"solver.boundary must be one of {sorted(_BOUNDARIES)}")
        if self.dtype not in _DTYPES:
            raise ConfigError(f"solver.dtype must be one of {sorted(_DTYPES)}")
        if self.engine not in _ENGINES:
            raise ConfigError(f"solver.engine must be one of {sorted(_ENGINES)}")
        if self.npml < 3:
            raise ConfigError(f"solver.npml must be at least 3, got {self.npml}")

    @classmethod
    def from_dict(cls, data: dict) -> "SolverConfig":
        return _from_mapping(cls, data)


@dataclass(frozen=True)
class ImagingConfig:
    """ertm/config.py:310-351."""

    mode: str = "pp"
    condition: str = "normalized"
    separation: str = "divcurl"
    normalize: bool = True
    illumination: bool = True
    receiver_normalization: bool = False
    laplacian: bool = True
    lowpass_fraction: float = 0.85
    checkpoint_mb: float = 256.0
    eps: float = 1.0e-10
    stack_normalize: bool = True
    start_periods: float = 1.5

    def validate(self) -> None:
        if self.mode not in _MODES:
            raise ConfigError(f"imaging.mode must be one of {sorted(_MODES)}")
        if self.condition not in _CONDITIONS:
            raise ConfigError(f"imaging.condition must be one of {sorted(_CONDITIONS)}")
        if self.separation not in _SEPARATIONS:
            raise ConfigError(f"imaging.separation must be one of {sorted(_SEPARATIONS)}")
        if not 0.0 < self.lowpass_fraction <= 1.0:
            raise ConfigError("imaging.lowpass_fraction must lie in (0, 1]")
        if self.checkpoint_mb <= 0:
            raise ConfigError("imaging.checkpoint_mb must be positive")
        if self.eps <= 0:
            raise ConfigError("imaging.eps must be positive")
        if self.start_periods < 0:
            raise ConfigError("imaging.start_periods must not be negative")

    @classmethod
    def from_dict(cls, data: dict) -> "ImagingConfig":
        return _from_mapping(cls, data)


@dataclass(frozen=True)
class QConfig:
    """ertm/config.py:390-436."""

    enabled: bool = False
    model: str = "filter"
    alpha_c: float = 0.8
    f_ref: float = 30.0
    gain_cap: float = 30.0
    apply_attenuation: bool = False
    attenuation_strength: float = 1.0
    n_mechanisms: int = 3
    band_ratio: float = 1.5
    amplitude_cap: float = 50.0
    lowpass_interval: int = 20
    f_max: float | None = None
    gain_budget_db: float = 40.0
    compare: bool = False

    def validate(self) -> None:
        if not 0.0 <= self.alpha_c <= 1.0:
            raise ConfigError("q.alpha_c must lie in [0, 1]")
        if self.f_ref <= 0:
            raise ConfigError("q.f_ref must be positive")
        if self.gain_cap < 1.0:
            raise ConfigError("q.gain_cap must be >= 1")
        if not 0.0 <= self.attenuation_strength <= 1.0:
            raise ConfigError("q.attenuation_strength must lie in [0, 1]")
        if self.model not in {"filter", "gsls"}:
            raise ConfigError(f"q.model must be filter or gsls, got {self.model!r}")
        if self.n_mechanisms < 1:
            raise ConfigError("q.n_mechanisms must be at least 1")
        if self.band_ratio <= 1.0:
            raise ConfigError("q.band_ratio must be > 1")
        if self.amplitude_cap < 1.0:
            raise ConfigError("q.amplitude_cap must be >= 1")
        if self.lowpass_interval < 1:
            raise ConfigError("q.lowpass_interval must be at least 1")
        if self.f_max is not None and self.f_max <= 0:
            raise ConfigError("q.f_max must be positive or null")
        if not 0.0 < self.gain_budget_db <= 200.0:
            raise ConfigError("q.gain_budget_db must lie in (0, 200]")

    @classmethod
    def from_dict(cls, data: dict) -> "QConfig":
        return _from_mapping(cls, data)


@dataclass(frozen=True)
class GpuConfig:
    """Engine options with no counterpart in the reference.

    strict_cfl     apply the reference's CFL guard (check_stability); False
                   lets through time steps that are stable but above the
                   guard's 0.55 safety factor.
    device_mem_gb  HBM the engine may use for checkpoints and replay slots of
                   one shot before falling back to nested (segment) replay.
    replay_slots   maximum image-field slots held per replay segment
                   (0 = one whole checkpoint block).
    replay_budget_gb  HBM for the replay slots of the shots migrated together
                   (capped at nine tenths of the HBM left free by the
                   checkpoints, receiver rings, backward states and
                   accumulators of the batch, counting what
                   torch's allocator holds unused); over it a block is
                   replayed in segments from sub-checkpoints.
    max_batch      most shots advanced together by ``migrate_batch``.
    kernel_variant (variant, zc) of the step kernels (see
                   ertm_batch_set_variant); None = the tuner's choice.  Every
                   variant gives the same results.
    image_ring     backward steps whose receiver fields a fused batch folds
                   into the image accumulators in one pass (0 = fold every
                   step inside the step kernel, or in a pass of its own after
                   the viscoacoustic stabiliser).  Same results; 32 is ~1 %
                   faster than 16 on 64 elastic shots.
    device_postprocess  normalise, Laplacian-filter, low-pass and mask the
                   images on the device (float64; the FFT rounds differently
                   from NumPy's, so images can differ from the host path in
                   the last float32 bit) instead of on the host.
    graph_steps    capture the time loops in chunks of this many steps as one
                   hipGraph each and replay it (even; 0 = launch every step
                   from the host).  Same results; 7-12 % faster migrations of
                   small grids (launch-bound steps), neutral on large ones.
                   For one host thread driving the GPU: a capture can be
                   invalidated by work other host threads issue meanwhile,
                   which the serving loop's workers do.
    two_step       (2, zc) runs the forward and replay passes of elastic
                   migrations at order 8 two steps per launch (128-lane
                   blocks, zc rows per chunk; ertm_batch_set_two_step);
                   (0, 0) = one step per launch; None = the tuner's choice
                   (one step: two per launch measured slower on MI355X).
                   Same results.
    device_resample  resample records whose dt differs from the solver's on
                   the device (ertm_resample_record: weights from the host,
                   the tap sums on the GPU, fused with the length fit) when
                   they are float32 with two or more traces and no
                   preprocess callable is given; same results as the host
                   path (whose einsum sums the taps in the same order).
    keep_tail      batched migrations keep the source fields of the last
                   steps in the replay slots during the checkpointing pass
                   instead of replaying them, when the source wavelet is
                   bitwise constant from that block's checkpoint on (replay
                   step t then does what forward step t + 1 did: same state,
                   same amplitude) and the slots hold more than the last
                   block.  Excitation imaging replays the blocks from the
                   first such block start in one run carried on to nt, so its
                   checkpointing pass ends at that block's checkpoint.  Same
                   results.
    overlap_flush  fold the receiver-field ring into the image accumulators
                   on a stream of its own while the next backward steps run
                   (the ring then holds 2 x image_ring steps; eager loops
                   only, not the compensated viscoacoustic pass, whose folds
                   follow its stabiliser).  Same results; off by default:
                   on 64 elastic shots the backward steps ran about as much
                   slower beside the folds as the folds had taken between
                   them.
    tail_budget_gb HBM the kept tail may fill: None = the replay budget (a
                   batch migration's default); 0 = no more than the
                   whole-block replay slots the migration holds anyway (the
                   serving loop's default: its requests share the HBM, and a
                   request's tail slots would cost the batching of others).
    """

    strict_cfl: bool = True
    device_mem_gb: float = 200.0
    replay_slots: int = 0
    replay_budget_gb: float = 16.0
    max_batch: int = 64
    kernel_variant: tuple | None = None
    image_ring: int = 32
    device_postprocess: bool = False
    graph_steps: int = 0
    two_step: tuple | None = None
    device_resample: bool = True
    keep_tail: bool = True
    overlap_flush: bool = False
    tail_budget_gb: float | None = None

    def validate(self) -> None:
        if self.device_mem_gb <= 0:
            raise ConfigError("gpu.device_mem_gb must be positive")
        if self.replay_slots < 0:
            raise ConfigError("gpu.replay_slots must not be negative")
        if self.replay_budget_gb <= 0:
            raise ConfigError("gpu.replay_budget_gb must be positive")
        if self.tail_budget_gb is not None and self.tail_budget_gb < 0:
            raise ConfigError("gpu.tail_budget_gb must be non-negative or null")
        if self.max_batch < 1:
            raise ConfigError("gpu.max_batch must be at least 1")
        if self.image_ring < 0:
            raise ConfigError("gpu.image_ring must not be negative")
        if self.graph_steps < 0 or self.graph_steps % 2:
            raise ConfigError("gpu.graph_steps must be a non-negative even number")
