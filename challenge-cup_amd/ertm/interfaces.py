"""Plugin protocols of the hot path (ertm/interfaces.py:40-82).

``FiniteDifferenceSolver`` and ``RTMImager`` satisfy these structurally, so a
caller that accepts any ``ForwardSolver`` / ``Imager`` (the reference's
pipeline does, ertm/pipeline.py:167-189) can take this engine's objects.
"""

from __future__ import annotations

from typing import Any, Protocol, runtime_checkable

import numpy as np

from .types import ElasticModel, ImageResult, ReceiverSpec, ShotRecord, SourceSpec


@runtime_checkable
class ForwardSolver(Protocol):
    name: str

    def forward(self, model: ElasticModel, source: SourceSpec, receivers: ReceiverSpec, *,
                components: tuple[str, ...] | None = ...) -> ShotRecord:
        ...


@runtime_checkable
class Imager(Protocol):
    name: str

    def migrate(self, shot: ShotRecord, model: ElasticModel, *, wavelet: np.ndarray | None = ...,
                preprocess: Any = ...) -> ImageResult:
        ...

    def image_condition(self, source_wavefield: np.ndarray,
                        receiver_wavefield: np.ndarray) -> np.ndarray:
        ...
