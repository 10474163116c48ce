"""The shared device-media cache hands a tensor to another stream only after
the upload that filled it: the builder records an event on its stream and
every cache hit makes the taking stream wait on that event.  Checked on the
CPU with stand-ins for torch's streams and events."""

from __future__ import annotations

import types

import numpy as np


class _Event:
    def __init__(self):
        self.recorded_on = None

    def record(self, stream):
        self.recorded_on = stream


class _Stream:
    def __init__(self, name):
        self.name = name
        self.waited = []

    def wait_event(self, ev):
        self.waited.append(ev)


class _Tensor:  # weak-referenceable and takes attributes, like a torch tensor
    pass


def _fake_torch(current):
    cuda = types.SimpleNamespace(Event=_Event, current_stream=lambda: current[0])
    return types.SimpleNamespace(cuda=cuda)


def test_cache_hit_waits_for_the_builders_upload(monkeypatch):
    from ertm import device as dev

    current = [_Stream("builder")]
    uploads = []

    def fake_padded_rows(a, ld):
        t = _Tensor()
        uploads.append((t, current[0]))
        return t

    monkeypatch.setattr(dev, "torch_mod", lambda: _fake_torch(current))
    monkeypatch.setattr(dev, "padded_rows", fake_padded_rows)
    monkeypatch.setattr(dev, "to_device", lambda a: fake_padded_rows(a, 0))
    monkeypatch.setattr(dev, "_MEDIA", dev.weakref.WeakValueDictionary())

    a = np.arange(12, dtype=np.float32).reshape(3, 4)
    t1 = dev.shared_padded_rows(a, 64)
    assert len(uploads) == 1
    ev = t1._ertm_ready
    assert ev.recorded_on is current[0]           # recorded after the upload, on its stream
    assert current[0].waited == []                # the builder itself needs no wait

    worker = _Stream("worker")
    current[0] = worker
    t2 = dev.shared_padded_rows(a.copy(), 64)     # same bytes: a hit
    assert t2 is t1 and len(uploads) == 1
    assert worker.waited == [ev]                  # the worker waits for the upload

    v = np.linspace(0.0, 1.0, 5).astype(np.float32)
    current[0] = _Stream("other")
    p1 = dev.shared_vector(v)
    current[0] = second = _Stream("second")
    p2 = dev.shared_vector(v.copy())
    assert p2 is p1 and second.waited == [p1._ertm_ready]
