"""Stacking of per-shot images (ertm/pipeline.py:133-161).

Only the stacking step of the reference's pipeline module lives here; the
orchestration around it (configuration, acquisition effects, file output) is
not part of this package.

``stack_images`` is the reference's operation: each image is scaled by the
inverse of its largest absolute value (float64), the scaled images are summed
in list order, the sum is divided by the number of images and rounded to
float32; a single image is returned as it is.

``StackAccumulator`` computes the same result incrementally for images that
arrive in any order (a serving loop finishing shots out of order): an image
is folded into the float64 sum as soon as every image before it in the stack
order has been folded, so the additions happen in list order and the result
is bit-identical to ``stack_images`` over the ordered list.  Images that
arrive early wait in a reorder buffer.
"""

from __future__ import annotations

import threading
from typing import Any

import numpy as np

from .exceptions import PipelineError
from .types import ImageResult


def _scaled(image: np.ndarray, normalize: bool) -> np.ndarray:
    data = np.asarray(image, dtype=np.float64)
    if normalize:
        scale = float(np.max(np.abs(data)))
        if scale > 0:
            data = data / scale
    return data


def _stacked_meta(first: ImageResult, n: int, normalize: bool) -> dict[str, Any]:
    meta = dict(first.meta)
    meta.update({"stacked": True, "n_shots": n, "stack_normalize": bool(normalize)})
    return meta


def stack_images(images: list[ImageResult], *, normalize: bool = True) -> ImageResult:
    """Mean of the max-abs normalised images (ertm/pipeline.py:133-161)."""
    if not images:
        raise PipelineError("nothing to stack")
    if len(images) == 1:
        return images[0]
    total = np.zeros(np.asarray(images[0].image).shape, dtype=np.float64)
    for im in images:
        total += _scaled(im.image, normalize)
    total /= len(images)
    return ImageResult(image=total.astype(np.float32), grid=images[0].grid, mode=images[0].mode,
                       meta=_stacked_meta(images[0], len(images), normalize))


class StackAccumulator:
    """``stack_images`` over ``count`` images that arrive in any order.

    ``add(index, image)`` takes the image at position ``index`` of the stack
    order; ``done`` turns true with the last one and ``result()`` then
    returns what ``stack_images`` returns for the ordered list.  Thread-safe.
    """

    def __init__(self, count: int, *, normalize: bool = True) -> None:
        if count < 1:
            raise PipelineError(f"a stack needs at least one image, got {count}")
        self.count = int(count)
        self.normalize = bool(normalize)
        self._lock = threading.Lock()
        self._waiting: dict[int, ImageResult] = {}
        self._next = 0
        self._total: np.ndarray | None = None
        self._first: ImageResult | None = None

    @property
    def done(self) -> bool:
        return self._next == self.count

    def add(self, index: int, image: ImageResult) -> bool:
        """Fold image ``index`` in; returns ``done``."""
        index = int(index)
        with self._lock:
            if not 0 <= index < self.count:
                raise PipelineError(f"stack index {index} outside [0, {self.count})")
            if index < self._next or index in self._waiting:
                raise PipelineError(f"stack index {index} given twice")
            self._waiting[index] = image
            while self._next in self._waiting:
                im = self._waiting.pop(self._next)
                if self._next == 0:
                    self._first = im
                if self.count > 1:
                    if self._total is None:
                        self._total = np.zeros(np.asarray(im.image).shape, dtype=np.float64)
                    self._total += _scaled(im.image, self.normalize)
                self._next += 1
            return self.done

    def result(self) -> ImageResult:
        with self._lock:
            if self._next != self.count:
                raise PipelineError(f"stack incomplete: {self._next} of {self.count} images")
            first = self._first
            if self.count == 1:
                return first
            mean = self._total / self.count
            return ImageResult(image=mean.astype(np.float32), grid=first.grid, mode=first.mode,
                               meta=_stacked_meta(first, self.count, self.normalize))
