"""The kept tail of a migration's checkpointing pass (host logic, no GPU).

Replaying a block restarts from the state saved after step s0 and steps it
with w[s0], w[s0 + 1], ... (ertm/imaging.py:140-162); the forward pass
stepped that state with w[s0 + 1], ...  Once every wavelet is bitwise
constant from s0 on, replay step t is forward step t + 1, so the forward
pass can keep those source fields instead of a replay recomputing them.
The first tests check that property on the oracle, field for field, and
that it fails where the wavelet still varies (so the comparison can fail);
the rest check how ``keep_tail_start`` places the tail.
"""
from __future__ import annotations

import numpy as np
import pytest

from ertm.imaging import constant_block_start, keep_tail_start
from oracle import rtm as ortm


def _model(nz, nx, seed):
    rng = np.random.default_rng(seed)
    vp = rng.uniform(1800.0, 2600.0, (nz, nx)).astype(np.float32)
    vs = (vp / rng.uniform(1.7, 2.0, (nz, nx))).astype(np.float32)
    rho = rng.uniform(1900.0, 2400.0, (nz, nx)).astype(np.float32)
    qp = np.full((nz, nx), 80.0, np.float32)
    return ortm.Model(nz, nx, 10.0, 10.0, vp, vs, rho, qp)


def _wavelet(nt, n_vary, tail_value, seed):
    w = np.random.default_rng(seed).standard_normal(nt).astype(np.float32)
    w[n_vary:] = np.float32(tail_value)
    return w


def _forward_fields(prop, w, sep):
    """Image fields after forward steps 0 .. nt, step nt taking w[nt - 1]."""
    out = []
    for a in list(w) + [w[-1]]:
        prop.step(float(a))
        out.append({k: np.array(v, copy=True) for k, v in prop.image_fields(sep).items()})
    return out


def _same(a, b):
    return a.keys() == b.keys() and all(np.array_equal(a[k].view(np.uint32),
                                                       b[k].view(np.uint32)) for k in a)


@pytest.mark.parametrize("engine,sep,boundary,tail_value", [
    ("elastic", "divcurl", "free", -0.0),
    ("elastic", "stress", "cpml", 0.375),
    ("acoustic", "stress", "cpml", 0.0),
])
def test_replayed_fields_are_the_next_forward_fields_once_the_wavelet_is_constant(
        engine, sep, boundary, tail_value):
    nz, nx, dt, nt, n_vary = 28, 34, 1.0e-3, 40, 13
    m = _model(nz, nx, 3)
    w = _wavelet(nt, n_vary, tail_value, 5)

    def prop():
        p = ortm.build(engine, m, dt, nt, order=8, npml=8, boundary=boundary)
        p.set_source(150.0, 60.0)
        return p

    # the reference's 1 MB budget floor: blocks of 2 steps (elastic), 1 (acoustic)
    st = ortm.record(prop(), w, budget_mb=1.0)
    assert st.interval == (2 if engine == "elastic" else 1)
    fwd = _forward_fields(prop(), w, sep)
    replayer = prop()
    for b, (s0, _) in enumerate(st.states):
        s0_, s1 = ortm.block_bounds(st, b, nt)
        fields = ortm.replay(replayer, w, st, b, sep)
        shifted = all(_same(fields[i], fwd[s0 + i + 1]) for i in range(s1 - s0))
        if s0 >= n_vary:
            assert shifted, f"block {b} at {s0}"
            # the same index of the forward pass is a different field
            assert not _same(fields[0], fwd[s0])
        else:
            # w still varies inside or after the block start: not the same
            assert not shifted, f"block {b} at {s0}"


@pytest.mark.parametrize("engine,sep,boundary", [("elastic", "divcurl", "free"),
                                                 ("acoustic", "stress", "cpml")])
def test_a_replay_carried_on_past_its_block_replays_the_next_blocks(engine, sep, boundary):
    """The excitation maps replay the blocks in order; from the first block
    at which the wavelet is constant, one replay run on to nt gives every
    later block's fields, so their checkpoints are not needed."""
    nz, nx, dt, nt, n_vary = 26, 30, 1.0e-3, 36, 11
    m = _model(nz, nx, 7)
    w = _wavelet(nt, n_vary, -0.0, 9)

    def prop():
        p = ortm.build(engine, m, dt, nt, order=8, npml=8, boundary=boundary)
        p.set_source(140.0, 50.0)
        return p

    st = ortm.record(prop(), w, budget_mb=1.0)
    starts = [s0 for s0, _ in st.states]
    first = constant_block_start([w], nt, st.interval)
    assert first == -(-n_vary // st.interval) * st.interval
    replayer = prop()
    per_block = {}
    for b, s0 in enumerate(starts):
        for i, f in enumerate(ortm.replay(replayer, w, st, b, sep)):
            per_block[s0 + i] = f
    for b0 in (starts.index(first), starts.index(first) - 1):
        # one run from block b0's checkpoint to nt
        replayer.load_state(st.states[b0][1])
        chained = []
        for t in range(starts[b0], nt):
            replayer.step(float(w[t]))
            chained.append({k: np.array(v, copy=True)
                            for k, v in replayer.image_fields(sep).items()})
        same = all(_same(f, per_block[starts[b0] + i]) for i, f in enumerate(chained))
        assert same == (starts[b0] >= first), (b0, starts[b0], first)


def test_constant_block_start():
    nt, interval = 100, 10
    assert constant_block_start([_wavelet(nt, 31, 0.0, 1)], nt, interval) == 40
    assert constant_block_start([_wavelet(nt, 30, 0.0, 1)], nt, interval) == 30
    assert constant_block_start([_wavelet(nt, 31, 0.0, 1), _wavelet(nt, 95, 1.5, 2)], nt,
                                interval) is None
    assert constant_block_start([np.zeros(nt, np.float32)], nt, interval) == 0


def _ricker(nt, f0, delay, dt):
    return ortm.wavelet("ricker", f0, delay, 1.0, dt, nt)


def test_tail_of_the_64_shot_workload():
    """Ricker 30 Hz, delay 40 ms, dt 0.4 ms: bitwise -0.0 from step 378."""
    nt, interval = 4000, 237
    w = _ricker(nt, 30.0, 0.04, 4.0e-4)
    u = w.view(np.uint32)
    assert np.all(u[378:] == 0x80000000) and u[377] != 0x80000000
    # room for 1000 slots: the tail starts at nt - 1000 (mid-block)
    assert keep_tail_start([w], nt, interval, 1000) == 3000
    # a tail no longer than the last block (3792 .. 3999) gains nothing
    assert keep_tail_start([w], nt, interval, 208) is None
    assert keep_tail_start([w], nt, interval, 209) == 3791
    # unlimited room: the first block start at or after step 378
    assert keep_tail_start([w], nt, interval, 10 ** 6) == 474


def test_tail_waits_for_every_wavelet_and_a_whole_block():
    nt, interval = 100, 10
    a = _wavelet(nt, 31, 0.0, 1)
    b = _wavelet(nt, 57, 0.0, 2)
    assert keep_tail_start([a], nt, interval, 10 ** 6) == 40
    assert keep_tail_start([a, b], nt, interval, 10 ** 6) == 60
    # block start exactly where the wavelet becomes constant
    c = _wavelet(nt, 60, 0.0, 3)
    assert keep_tail_start([c], nt, interval, 10 ** 6) == 60
    # the tail begins mid-block only inside an all-constant block
    assert keep_tail_start([a], nt, interval, 45) == 55


def test_no_tail_without_a_constant_end():
    nt = 300
    w = np.sin(np.arange(nt, dtype=np.float32) * 0.01).astype(np.float32)
    assert keep_tail_start([w], nt, 20, 10 ** 6) is None
    # +0.0 and -0.0 are different amplitudes for the step
    z = np.zeros(nt, np.float32)
    z[-1] = -0.0
    assert keep_tail_start([z], nt, 20, 10 ** 6) is None
    # one block only: nothing before the last block to keep
    assert keep_tail_start([np.zeros(nt, np.float32)], nt, nt, 10 ** 6) is None
    # a constant wavelet from step 0 keeps all but the first block's start
    assert keep_tail_start([np.zeros(nt, np.float32)], nt, 20, 10 ** 6) == 0
