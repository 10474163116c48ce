"""MI355X-native reverse-time-migration engine with the ``ertm`` API.

Drop-in for the reference package's hot path: ``ertm.solver`` (engine
dispatch), ``ertm.propagation`` (forward / adjoint extrapolation) and
``ertm.imaging`` (checkpoints, replay, imaging conditions).  Every time step,
receiver operation and imaging-condition accumulation runs as a hand-written
HIP kernel for gfx950 (libertm_hip.so, called through ctypes); PyTorch-ROCm
only provides the HBM allocations and the stream.  There is no CPU fallback.
"""

from __future__ import annotations

__version__ = "0.3.0+mi355x"

from .config import GpuConfig, ImagingConfig, QConfig, SolverConfig  # noqa: F401
from .exceptions import (  # noqa: F401
    ConfigError,
    DataError,
    ERTMError,
    ImagingError,
    ModelError,
    PipelineError,
    SolverError,
)
from .types import ElasticModel, GridSpec, ImageResult, ReceiverSpec, ShotRecord, SourceSpec  # noqa: F401
