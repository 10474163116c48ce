"""Stacking (ertm/pipeline.py:133-161): ``stack_images`` against the
oracle and closed forms, and the out-of-order accumulator against
``stack_images`` over the ordered list (bit-exact)."""

from __future__ import annotations

import numpy as np
import pytest

from ertm.exceptions import PipelineError
from ertm.pipeline import StackAccumulator, stack_images
from ertm.types import GridSpec, ImageResult
from oracle import rtm as ortm


def _images(n, seed=0, shape=(13, 17)):
    rng = np.random.default_rng(seed)
    g = GridSpec(nx=shape[1], nz=shape[0], dx=5.0, dz=5.0)
    out = []
    for i in range(n):
        a = (rng.normal(size=shape) * 10.0 ** rng.uniform(-3, 3)).astype(np.float32)
        out.append(ImageResult(image=a, grid=g, mode="pp", meta={"shot": i}))
    out.append(ImageResult(image=np.zeros(shape, np.float32), grid=g, mode="pp"))  # scale 0
    return out


def test_oracle_stack_closed_forms():
    a = np.array([[1.0, -4.0], [2.0, 0.5]], np.float32)
    b = np.array([[3.0, 0.0], [-6.0, 1.5]], np.float32)
    # normalised mean: (a/4 + b/6) / 2
    want = ((a.astype(np.float64) / 4.0 + b.astype(np.float64) / 6.0) / 2.0).astype(np.float32)
    assert np.array_equal(ortm.stack([a, b]), want)
    assert np.array_equal(ortm.stack([a, b], normalize=False),
                          ((a.astype(np.float64) + b) / 2.0).astype(np.float32))
    assert ortm.stack([a]) is a or np.array_equal(ortm.stack([a]), a)
    # every image at its own scale stacks to the same normalised image
    assert np.array_equal(ortm.stack([a, 3.0 * a, 0.25 * a]), (a / np.float32(4.0)))


@pytest.mark.parametrize("normalize", [True, False])
def test_stack_images_matches_oracle(normalize):
    ims = _images(7)
    got = stack_images(ims, normalize=normalize)
    want = ortm.stack([im.image for im in ims], normalize=normalize)
    assert got.image.dtype == np.float32 and np.array_equal(got.image, want)
    assert got.meta["stacked"] and got.meta["n_shots"] == 8 and got.meta["shot"] == 0
    assert stack_images(ims[:1]) is ims[0]
    with pytest.raises(PipelineError):
        stack_images([])


@pytest.mark.parametrize("count", [1, 2, 9])
def test_accumulator_out_of_order_is_bit_exact(count):
    ims = _images(count - 1, seed=count) if count > 1 else _images(0)[:1]
    ims = ims[:count]
    order = np.random.default_rng(3).permutation(count)
    acc = StackAccumulator(count)
    for k, i in enumerate(order):
        assert acc.add(int(i), ims[i]) == (k == count - 1)
    got = acc.result()
    want = stack_images(ims)
    assert np.array_equal(got.image, want.image) and got.meta == want.meta


def test_accumulator_rejects_bad_indices():
    ims = _images(2)
    acc = StackAccumulator(3)
    acc.add(1, ims[1])
    with pytest.raises(PipelineError):
        acc.add(1, ims[1])
    with pytest.raises(PipelineError):
        acc.add(3, ims[0])
    with pytest.raises(PipelineError):
        acc.result()
