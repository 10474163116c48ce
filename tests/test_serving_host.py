"""Serving-loop host logic without a GPU: admission order, HBM budget,
same-key batching, and the request price."""

from __future__ import annotations

import numpy as np
import pytest

from ertm.config import GpuConfig, ImagingConfig, SolverConfig
from ertm.serving import Admission, ShotRequest, request_bytes
from ertm.types import ElasticModel, GridSpec, ReceiverSpec, ShotRecord, SourceSpec


def _item(name, nbytes, key, lane=None):
    return (name, None, nbytes, key, 0.0, None, lane)


def _names(got):
    return [i[0] for i in got[2]]


def test_admission_is_first_come_first_served_within_budget():
    adm = Admission(budget=100, streams=4, max_group=8)
    for name, nb, key in [("a", 60, 1), ("b", 50, 2), ("c", 30, 1), ("d", 10, 3)]:
        adm.add(_item(name, nb, key))
    g1 = adm.take()
    # "a" starts with its same-key companion "c" (60 + 30 <= 100)
    assert g1[0] == "group" and _names(g1) == ["a", "c"]
    # "b" (50) does not fit next to 90 reserved; "d" must not overtake it
    assert adm.take() is None
    adm.release(g1[1], g1[2])
    g2 = adm.take()
    assert _names(g2) == ["b"]
    g3 = adm.take()
    assert _names(g3) == ["d"]
    assert adm.peak <= 100


def test_admission_respects_streams_and_group_cap():
    adm = Admission(budget=10 ** 9, streams=2, max_group=2)
    for k in range(5):
        adm.add(_item(k, 1, "same"))
    g1, g2 = adm.take(), adm.take()
    assert _names(g1) == [0, 1] and _names(g2) == [2, 3] and g1[1] != g2[1]
    assert adm.take() is None          # no idle stream
    adm.release(g1[1], g1[2])
    assert _names(adm.take()) == [4]


def test_lanes_start_on_idle_workers_then_join_the_fewest_of_their_engine():
    adm = Admission(budget=100, streams=3, max_group=2)
    for name, nb, lane in [("e0", 10, "elastic"), ("a0", 10, "acoustic"), ("e1", 10, "elastic"),
                           ("e2", 10, "elastic"), ("e3", 10, "elastic"), ("g", 10, None),
                           ("e4", 10, "elastic")]:
        adm.add(_item(name, nb, "k" + name, lane))
    got = [adm.take() for _ in range(5)]
    assert [(g[0], g[1], _names(g)) for g in got] == [
        ("lane", 0, ["e0"]), ("lane", 1, ["a0"]), ("lane", 2, ["e1"]), ("lane", 0, ["e2"]),
        ("lane", 2, ["e3"])]
    # no idle worker for the group request, and the lane behind it waits too
    assert adm.take() is None and adm.kind == ["elastic", "acoustic", "elastic"]
    adm.release(1, got[1][2])          # the acoustic worker is idle again
    g = adm.take()
    assert g[0] == "group" and g[1] == 1 and _names(g) == ["g"]
    # elastic workers are full (max_group 2): e4 waits until a lane leaves
    assert adm.take() is None
    adm.release(2, got[2][2])
    e4 = adm.take()
    assert (e4[0], e4[1], _names(e4)) == ("lane", 2, ["e4"]) and adm.count == [2, 1, 2]
    assert adm.reserved == 50 and adm.peak <= 100


def _request(nz, nx, nt, engine="elastic", checkpoint_mb=256.0):
    g = GridSpec(nx=nx, nz=nz, dx=5.0, dz=5.0)
    vp = np.full((nz, nx), 3000.0, np.float32)
    model = ElasticModel(g, vp=vp, vs=vp / 1.8, rho=np.full_like(vp, 2200.0),
                         qp=np.full_like(vp, 100.0), qs=np.full_like(vp, 50.0))
    rx = np.arange(0, nx, 2) * 5.0
    rec = ReceiverSpec(x=rx, z=np.full(rx.size, 10.0))
    comps = ("vx", "vz") if engine == "elastic" else ("p",)
    data = {c: np.zeros((nt, rx.size), np.float32) for c in comps}
    shot = ShotRecord(data=data[comps[0]], dt=4e-4, t=np.arange(nt) * 4e-4, receivers=rec,
                      source=SourceSpec(x=100.0, z=10.0), components=data)
    return ShotRequest(shot=shot, model=model,
                       solver=SolverConfig(nt=nt, dt=4e-4, engine=engine, boundary="free"),
                       imaging=ImagingConfig(checkpoint_mb=checkpoint_mb))


def test_request_price_covers_the_large_parts_and_grows_with_the_job():
    small = request_bytes(_request(256, 256, 1500))
    big = request_bytes(_request(600, 1200, 4000))
    assert big > small > 0
    r = _request(600, 1200, 4000)
    nzp, ld = 610, 1216
    state = 5 * nzp * ld * 4
    # two propagators x two ping-pong states, the checkpoint arena (18 states
    # at the 256 MB rule) and a whole block of replay slots (interval 223)
    floor = 4 * state + 18 * state + 223 * 600 * 1200 * 4
    assert request_bytes(r) >= floor
    # a replay-slot cap shrinks the price
    assert request_bytes(r, GpuConfig(replay_slots=8)) < request_bytes(r)
    # acoustic requests are cheaper than elastic ones of the same shape
    assert request_bytes(_request(600, 1200, 4000, "acoustic")) < request_bytes(r)


def test_one_engine_does_not_block_the_other():
    """An acoustic request with no worker to go to lets later elastic ones
    start, and the busy worker with the fewest lanes drains so that the
    acoustic one gets a worker."""
    adm = Admission(budget=100, streams=2, max_group=8)
    for name, lane in [("e0", "elastic"), ("e1", "elastic"), ("a0", "acoustic"),
                       ("e2", "elastic"), ("e3", "elastic")]:
        adm.add(_item(name, 10, "k" + name, lane))
    got = [adm.take() for _ in range(3)]
    assert [(g[1], _names(g)) for g in got] == [(0, ["e0"]), (1, ["e1"]), (1, ["e2"])]
    # a0 waits; worker 0 (fewest lanes) drains: e3 joins worker 1 instead
    assert adm.draining == {0}
    e3 = adm.take()
    assert (e3[1], _names(e3)) == (1, ["e3"]) and adm.take() is None
    adm.release(0, got[0][2])
    a0 = adm.take()
    assert (a0[0], a0[1], _names(a0)) == ("lane", 0, ["a0"]) and adm.kind == ["acoustic", "elastic"]
    assert not adm.draining and not adm.pending


def test_serving_admission_counts_kept_tails_only_inside_the_block_slots():
    """The default SourceSpec's Ricker wavelet is bitwise constant long
    before nt = 4000, so a batch migration keeps a tail in up to the replay
    budget; the serving loop keeps tails only inside the whole-block slots,
    so a request's HBM price is the one without a tail."""
    from ertm.serving import serving_gpu_config

    r = _request(600, 1200, 4000)
    plain = request_bytes(r, GpuConfig(keep_tail=False))
    assert request_bytes(r, GpuConfig(tail_budget_gb=0.0)) == plain
    assert request_bytes(r) > plain + 8 * 2 ** 30
    # a tail budget between the two prices the tail it allows
    assert plain < request_bytes(r, GpuConfig(tail_budget_gb=4.0)) < request_bytes(r)
    assert serving_gpu_config(None).tail_budget_gb == 0.0
    assert serving_gpu_config(GpuConfig(tail_budget_gb=2.0)).tail_budget_gb == 2.0
    from ertm.exceptions import ConfigError

    try:
        GpuConfig(tail_budget_gb=-1.0).validate()
    except ConfigError:
        pass
    else:
        raise AssertionError("a negative tail budget was accepted")


def test_taking_a_lane_behind_a_waiting_request_does_not_compare_requests():
    """A lane taken from behind a waiting request leaves the queue by
    identity: the queued items hold ShotRequest dataclasses, whose equality
    compares their records' NumPy arrays (an error for arrays of different
    shapes).  Same schedule as the draining test above, with real requests
    of different grids."""
    reqs = {name: _request(64 + 4 * k, 80 + 8 * k, 300 + k,
                           "acoustic" if name.startswith("a") else "elastic")
            for k, name in enumerate(["e0", "e1", "a0", "e2", "e3"])}
    with pytest.raises(ValueError):
        bool(reqs["a0"] == reqs["e2"])   # what removing e2 with list.remove evaluated
    adm = Admission(budget=100, streams=2, max_group=8)
    for name, req in reqs.items():
        adm.add((req, None, 10, "k" + name, 0.0, None, req.solver.engine))
    got = [adm.take() for _ in range(3)]
    # e2 is taken from behind the waiting a0
    assert [(g[1], g[2][0][0]) for g in got] == [(0, reqs["e0"]), (1, reqs["e1"]),
                                                 (1, reqs["e2"])]
    assert [x[0] for x in adm.pending] == [reqs["a0"], reqs["e3"]]
