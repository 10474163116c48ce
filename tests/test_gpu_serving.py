"""Serving loop on the GPU: every image equals the one its shot gets alone,
requests sharing a launch key run batched, a failing request fails alone, and
admission keeps the reserved HBM within the budget."""

from __future__ import annotations

import dataclasses
import sys

import numpy as np
import pytest

from serving_cases import make_request

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def _engine():
    from ertm import _lib

    _lib.lib()


def _alone(req):
    from ertm.imaging import RTMImager

    return RTMImager(req.imaging, req.solver, req.q).migrate(req.shot, req.model)


@pytest.mark.parametrize("lanes", [True, False])
def test_results_identical_alone_and_in_the_mix(lanes):
    _say("serving mix vs alone, lanes", lanes)
    from ertm.serving import ServingLoop

    reqs = [make_request(100 + k, rid=k) for k in range(5)]
    # three shots that share a launch key (same engine, grid, nt and model)
    base = make_request(7, engine="elastic", nz=56, nx=80, nt=420, rid="s0")
    for k in range(3):
        src = dataclasses.replace(base.shot.source, x=base.shot.source.x + 20.0 * k)
        reqs.append(dataclasses.replace(base, shot=dataclasses.replace(base.shot, source=src),
                                        rid=f"s{k}"))
    with ServingLoop(hbm_budget_gb=16.0, streams=3, lanes=lanes) as loop:
        futs = [loop.submit(r) for r in reqs]
        res = [f.result(timeout=600) for f in futs]
        stats = loop.stats()
    _say("stats", stats)
    assert all(r.ok for r in res), [r.error for r in res if not r.ok]
    assert max(r.group_size for r in res) >= 2
    for req, r in zip(reqs, res):
        alone = _alone(req)
        assert np.abs(alone.image).max() > 0, req.rid
        assert np.array_equal(r.images[req.imaging.mode].image, alone.image), req.rid
    assert stats["completed"] == len(reqs) and stats["p99_s"] >= stats["p50_s"] > 0


@pytest.mark.parametrize("lanes", [True, False])
def test_a_failing_request_fails_alone(lanes):
    _say("serving isolation, lanes", lanes)
    from ertm.serving import ServingLoop

    good = make_request(11, engine="elastic", nz=48, nx=64, nt=420, rid="good")
    # a finite record so large that the wavefield overflows: it passes
    # validation and breaks at run time, inside a launch group with "good"
    huge = {k: np.full_like(v, 3e38) for k, v in good.shot.components.items()}
    first = next(iter(huge))
    bad = dataclasses.replace(good, rid="bad", shot=dataclasses.replace(
        good.shot, data=huge[first], components=huge))
    invalid = dataclasses.replace(good, rid="invalid",
                                  solver=dataclasses.replace(good.solver, dt=1.0))
    blocker = make_request(12, nz=64, nx=96, nt=460, rid="blocker")
    with ServingLoop(hbm_budget_gb=8.0, streams=1, lanes=lanes) as loop:
        # one stream: "good" and "bad" wait behind "blocker" and start together
        # (as a launch group, or as lanes of one worker)
        futs = {r.rid: loop.submit(r) for r in (blocker, good, bad, invalid)}
        res = {k: f.result(timeout=600) for k, f in futs.items()}
    assert res["blocker"].ok and res["good"].ok
    if not lanes:
        assert res["bad"].rerun_alone and res["good"].rerun_alone  # their group broke
    else:
        # lanes keep their own buffers: the bad one fails, the others run on
        assert res["good"].group_size > 1 and not res["good"].rerun_alone
    assert not res["bad"].ok and not res["invalid"].ok
    assert res["invalid"].error.startswith("rejected")
    assert np.array_equal(res["good"].images["pp"].image, _alone(good).image)


def test_admission_keeps_reserved_hbm_within_budget():
    _say("serving admission")
    from ertm.serving import ServingLoop, request_bytes

    reqs = [make_request(200 + k, rid=k) for k in range(6)]
    need = max(request_bytes(r) for r in reqs)
    budget_gb = 1.5 * need / 1024 ** 3          # at most one request at a time
    with ServingLoop(hbm_budget_gb=budget_gb, streams=4) as loop:
        futs = [loop.submit(r) for r in reqs]
        big = dataclasses.replace(make_request(300, nz=400, nx=400, nt=400), rid="too big")
        rej = loop.submit(big).result(timeout=60)
        res = [f.result(timeout=600) for f in futs]
        stats = loop.stats()
    assert all(r.ok for r in res)
    assert not rej.ok and "HBM" in rej.error
    assert stats["peak_reserved_gib"] <= budget_gb + 1e-9
    # at every moment the requests holding HBM together fit the budget
    for r in res:
        t = r.start
        live = sum(q.hbm_bytes for q in res if q.start <= t < q.released)
        assert live <= budget_gb * 1024 ** 3


def test_stack_as_shots_finish_equals_stack_images():
    """Shots of one stack finish in whatever order the loop runs them; the
    stacked image equals stack_images over the shots alone, in stack order
    (ertm/pipeline.py:133-161)."""
    _say("serving stack")
    from ertm.pipeline import stack_images
    from ertm.serving import ServingLoop

    base = make_request(11, engine="elastic", nz=52, nx=76, nt=400, rid="b")
    reqs = []
    for k in range(4):
        src = dataclasses.replace(base.shot.source, x=base.shot.source.x + 30.0 * k)
        reqs.append(dataclasses.replace(base, shot=dataclasses.replace(base.shot, source=src),
                                        rid=f"st{k}", stack="line-1", stack_index=k))
    others = [make_request(200 + k, rid=f"o{k}") for k in range(2)]
    with ServingLoop(hbm_budget_gb=16.0, streams=2, max_group=2) as loop:
        stack = loop.open_stack("line-1", len(reqs))
        futs = [loop.submit(r) for r in (reqs[3], others[0], reqs[1], reqs[0], others[1], reqs[2])]
        res = [f.result(timeout=600) for f in futs]
        stacked = stack.result(timeout=600)
        bad = loop.submit(dataclasses.replace(reqs[0], stack="nope")).result(timeout=60)
    assert all(r.ok for r in res), [r.error for r in res if not r.ok]
    assert not bad.ok and "stack not open" in bad.error
    mode = base.imaging.mode
    want = stack_images([_alone(r) for r in reqs])
    assert np.array_equal(stacked[mode].image, want.image)
    assert stacked[mode].meta["n_shots"] == 4
