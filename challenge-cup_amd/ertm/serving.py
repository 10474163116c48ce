"""Serving loop for shot-migration requests on one GPU.

The reference migrates shots one call at a time (ertm/imaging.py:205-349,
stacked by ertm/pipeline.py:133-161); it has no request loop.  This module
runs a stream of independent migration requests on one device:

* **admission by HBM**: every request is priced (``request_bytes``: the
  forward and backward propagators' states and media, the checkpoint arena,
  replay slots and sub-checkpoints, float64 accumulators, the record on the
  device) and admitted only while the reserved total stays within the
  budget.  Admission is first-come first-served: the oldest waiting request
  is admitted first, and a request larger than the whole budget is rejected;
* **continuous batching**: acoustic and elastic zero_lag/normalized requests
  run as lanes (``ertm.lanes``): each worker advances up to ``max_group``
  lanes of one engine with shared launches per time step, whatever each
  request's nt, dt, grid, checkpoint interval or phase; an admitted request
  starts on an idle worker or, when none is idle, joins the worker with the
  fewest lanes of its engine within ``lane_ticks`` steps, and a finished one
  leaves at once.  Other requests
  run in launch groups: waiting requests that can share the oldest one's
  launches (same engine, configs, nt and checkpoint interval -- see
  ``RTMImager.migrate_batch``) and that also fit join it;
* **concurrency**: several workers run at once, each driven by its own host
  thread on its own HIP stream; HBM freed by a finished request admits the
  next requests at once.  Wavefields, checkpoints and partial images stay on
  the device for a request's whole life;
* **isolation**: a request that fails validation fails alone; if a worker's
  shared run fails, its requests are re-run one by one, so a bad request
  fails only itself.  A shot's image does not depend on what else is running
  (each image equals the one ``migrate`` gives for that shot alone);
* **stacking as shots finish**: requests that name a stack (``stack``,
  ``stack_index``) are folded into that stack's float64 sum as they finish
  (``open_stack``), in stack order through a reorder buffer, so the stacked
  image equals ``stack_images`` over the shots in stack order
  (ertm/pipeline.py:133-161).

Results carry arrival, start and finish times and how many shots advanced
with them per step; ``stats`` reports shots/s, the p50/p99 latency, the mean
of that group size and the lanes sharing each step launch.
"""

from __future__ import annotations

import dataclasses
import math
import threading
import time
from collections import deque
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import numerics as num
from ._lib import EngineUnavailable
from .config import GpuConfig, ImagingConfig, QConfig, SolverConfig
from .exceptions import ERTMError, ImagingError, PipelineError
from .imaging import RTMImager, checkpoint_interval, slot_plan
from .lanes import Lane, LaneRunner, lane_slot_room, lanes_supported
from .pipeline import StackAccumulator
from .propagation import PropagationSpec, source_wavelet
from .types import ElasticModel, ImageResult, ShotRecord

_GIB = 1024.0 ** 3


@dataclass
class ShotRequest:
    """One shot to migrate: record, model and the configs of one migrate call."""

    shot: ShotRecord
    model: ElasticModel
    solver: SolverConfig
    imaging: ImagingConfig
    q: QConfig | None = None
    modes: tuple[str, ...] | None = None
    rid: Any = None
    stack: Any = None       # key of the stack this shot belongs to (``open_stack``)
    stack_index: int = 0    # its position in that stack's order


@dataclass
class ShotResult:
    rid: Any
    images: dict[str, ImageResult] | None
    error: str | None
    arrival: float
    start: float = math.nan
    finish: float = math.nan
    group_size: float = 0      # shots advanced with it per step (mean over its steps)
    hbm_bytes: int = 0
    rerun_alone: bool = False   # its launch group failed; it was re-run by itself
    released: float = math.nan  # when its HBM reservation was returned

    def __post_init__(self) -> None:
        if math.isnan(self.released):
            self.released = self.finish

    @property
    def latency(self) -> float:
        return self.finish - self.arrival

    @property
    def ok(self) -> bool:
        return self.error is None


def _round_up(x: int, m: int) -> int:
    return (int(x) + m - 1) // m * m


def request_bytes(req: ShotRequest, gpu: GpuConfig | None = None) -> int:
    """HBM one request holds while it runs (an upper estimate)."""
    gpu = gpu or GpuConfig()
    engine = req.solver.engine
    nz, nx = req.model.shape
    spec = PropagationSpec.from_config(req.solver, req.model)
    pad = num.ghost_width(int(spec.order))
    nzp, nxp = nz + 2 * pad, nx + 2 * pad
    ld = _round_up(nxp, 64)
    nmech = int(req.q.n_mechanisms) if (req.q is not None and engine == "viscoacoustic") else 0
    nf = {"acoustic": 3, "elastic": 5}.get(engine, 4 + nmech)
    w = int(spec.npml) + pad
    nterms = 4 if engine == "elastic" else 2
    psi = nterms * (nzp * 2 * w + 2 * w * ld) + 4 * nterms * 64
    state = nf * _round_up(nzp * ld, 64) + psi                    # floats, one buffer
    media = (3 if engine == "elastic" else 3 + nmech) * nzp * ld
    if engine == "elastic":
        media += nzp * ld // 4 + 768                              # medium palette
    props = 2 * (2 * state + media)                               # forward + backward
    nt = int(spec.nt)
    interval = checkpoint_interval(nt, nf * nzp * nxp * 4, float(req.imaging.checkpoint_mb))
    arena = math.ceil(nt / interval) * state
    cells = nz * nx
    # a whole block, or the kept tail of the checkpointing pass
    wave = source_wavelet(req.shot.source, spec.dt, nt)
    slots, _ = slot_plan([wave], nt, interval, lane_slot_room(gpu, cells), gpu, cells)
    sub = math.ceil(interval / slots) * state if slots < interval else 0
    nmodes = len(req.modes or (req.imaging.mode,))
    accum = (2 * nmodes + 1 + 3) * cells * 2                      # float64 -> 2 floats
    # receiver-field ring of the backward pass
    ring = int(gpu.image_ring) * nmodes * nz * ld
    nrec = int(req.shot.receivers.nrec)
    ncomp = max(1, len(req.shot.components))
    record = nt * nrec * ncomp + 16 * nrec * ncomp
    floats = props + arena + slots * cells + sub + accum + ring + record
    return int(floats * 4 * 1.05) + (8 << 20)


def serving_gpu_config(gpu: GpuConfig | None) -> GpuConfig:
    """The engine options a serving loop runs with: ``gpu`` (default
    ``GpuConfig()``) with kept tails only inside the whole-block replay
    slots a request holds anyway unless ``tail_budget_gb`` is set.  The
    requests share the HBM, and slots one request's tail took would cost
    the batching of others."""
    gpu = gpu or GpuConfig()
    if gpu.tail_budget_gb is None:
        gpu = dataclasses.replace(gpu, tail_budget_gb=0.0)
    return gpu


def _key(req: ShotRequest, imager_key) -> tuple:
    def tup(c):
        return None if c is None else dataclasses.astuple(c)
    return (tup(req.solver), tup(req.imaging), tup(req.q),
            tuple(req.modes or (req.imaging.mode,)), imager_key)


class Admission:
    """Admission by bytes onto the loop's workers.

    Items are (request, job, bytes, key, arrival, future, lane) tuples; only
    bytes, key and lane are read here.  ``lane`` is the engine name of a
    request that runs as a lane (``ertm.lanes``), None for one that runs in a
    launch group.  A worker is idle, runs one launch group, or runs lanes of
    one engine (up to ``max_group``).  ``take`` returns the next assignment:
    ("group", worker, [items]) -- an item and its same-key companions on an
    idle worker -- or ("lane", worker, [item]) onto an idle worker, else the
    worker whose lanes of that engine are fewest; or None.

    Order: the bytes are first-come first-served (nothing overtakes an item
    that waits for HBM).  An item that fits but has no worker to go to (every
    worker busy with the other engine, or no idle one for a launch group)
    lets later items of other kinds start, so one engine never blocks the
    other; and so that it does not wait for ever, the busy worker with the
    fewest lanes is marked draining: it takes no new lanes and goes idle
    when its own finish.
    """

    def __init__(self, budget: int, streams: int, max_group: int) -> None:
        self.budget = int(budget)
        self.max_group = int(max_group)
        self.pending: deque = deque()
        self.reserved = 0
        self.peak = 0
        self.kind: list = [None] * int(streams)   # None | "group" | engine of its lanes
        self.count = [0] * int(streams)
        self.draining: set = set()

    @property
    def idle(self) -> int:
        return sum(k is None for k in self.kind)

    def add(self, item) -> None:
        self.pending.append(item)

    def _idle_worker(self):
        return next((w for w, k in enumerate(self.kind) if k is None), None)

    def _lane_worker(self, engine: str):
        # an idle worker first (its own stream: runs beside the others), else
        # the worker with the fewest lanes of this engine
        best = self._idle_worker()
        if best is not None:
            return best
        for w, k in enumerate(self.kind):
            if (k == engine and w not in self.draining and self.count[w] < self.max_group
                    and (best is None or self.count[w] < self.count[best])):
                best = w
        return best

    def _drain_one(self, engine) -> None:
        """No worker can take an item of ``engine`` (None: a launch group):
        unless one already drains, stop feeding the busy worker with the
        fewest items of another kind, so it comes free."""
        if self.draining:
            return
        cands = [w for w, k in enumerate(self.kind) if k is not None and k != engine]
        if cands:
            self.draining.add(min(cands, key=lambda w: self.count[w]))

    def _reserve(self, nbytes: int) -> None:
        self.reserved += nbytes
        self.peak = max(self.peak, self.reserved)

    def take(self):
        blocked: set = set()   # kinds whose oldest waiting item has no worker
        for item in list(self.pending):
            if self.reserved + item[2] > self.budget:
                return None
            lane = item[6]
            kind = lane if lane is not None else ("group", item[3])
            if kind in blocked:
                continue
            if lane is not None:
                w = self._lane_worker(lane)
                if w is None:
                    blocked.add(kind)
                    self._drain_one(lane)
                    continue
                # by identity: the items hold requests whose dataclass
                # equality compares their NumPy arrays
                self.pending = deque(x for x in self.pending if x is not item)
                self._reserve(item[2])
                self.kind[w] = lane
                self.count[w] += 1
                return ("lane", w, [item])
            w = self._idle_worker()
            if w is None:
                blocked.add(kind)
                self._drain_one(None)
                continue
            group = [item]
            used = item[2]
            for other in self.pending:
                if len(group) >= self.max_group:
                    break
                if (other is not item and other[6] is None and other[3] == item[3]
                        and self.reserved + used + other[2] <= self.budget):
                    group.append(other)
                    used += other[2]
            taken = {id(x) for x in group}
            self.pending = deque(x for x in self.pending if id(x) not in taken)
            self._reserve(used)
            self.kind[w] = "group"
            self.count[w] = len(group)
            return ("group", w, group)
        return None

    def release(self, worker: int, items) -> None:
        """Items assigned to ``worker`` have finished: free their bytes."""
        self.reserved -= sum(item[2] for item in items)
        self.count[worker] -= len(items)
        if self.count[worker] <= 0:
            self.kind[worker] = None
            self.count[worker] = 0
            self.draining.discard(worker)


class ServingLoop:
    """Admit, batch and run migration requests on the current GPU.

    ``submit`` returns a Future of a ``ShotResult``.  ``streams`` workers run
    concurrently, each on its own stream; ``hbm_budget_gb`` caps the HBM
    reserved by admitted requests; ``max_group`` caps the shots one worker
    advances together.  With ``lanes`` (the default), acoustic and elastic
    zero_lag/normalized requests run as lanes (``ertm.lanes``): a worker
    advances all its lanes of one engine in shared launches whatever their
    nt, dt, grid and phase, and a new request joins a running worker's lanes
    within ``lane_ticks`` time steps.  Other requests (and all of them
    without ``lanes``) run in launch groups of requests with one launch key.
    """

    def __init__(self, hbm_budget_gb: float = 200.0, streams: int = 4, max_group: int = 64,
                 gpu_config: GpuConfig | None = None, lanes: bool = True,
                 lane_ticks: int = 256) -> None:
        if streams < 1 or max_group < 1 or hbm_budget_gb <= 0 or lane_ticks < 1:
            raise ValueError("streams, max_group, lane_ticks and hbm_budget_gb must be positive")
        self.budget = int(float(hbm_budget_gb) * _GIB)
        self.gpu_config = serving_gpu_config(gpu_config)
        self.lanes = bool(lanes)
        self.lane_ticks = int(lane_ticks)
        self._lock = threading.Lock()
        self._work = threading.Condition(self._lock)
        self._adm = Admission(self.budget, streams, max_group)
        self._inbox: list[list] = [[] for _ in range(int(streams))]
        self._open = True
        self._results: list[ShotResult] = []
        self._outstanding = 0
        self._stacks: dict[Any, dict[str, Any]] = {}
        self._launch_stats = [0, 0]   # lane steps, step-launch groups
        # host post-processing of finished lanes, beside the workers' steps
        self._post_pool = ThreadPoolExecutor(max_workers=4, thread_name_prefix="ertm-post")
        self._threads = [threading.Thread(target=self._worker, args=(k,), name=f"ertm-serve-{k}",
                                          daemon=True) for k in range(int(streams))]
        for t in self._threads:
            t.start()

    # ------------------------------------------------------------ public
    def open_stack(self, key: Any, count: int, *, normalize: bool = True) -> Future:
        """Stack the ``count`` requests submitted with ``stack=key`` (indices
        0 .. count-1).  The Future resolves to {mode: ImageResult}, each equal
        to ``stack_images`` of that mode's images in index order, or to a
        PipelineError naming the shots that failed."""
        if count < 1:
            raise ValueError("a stack needs at least one shot")
        fut: Future = Future()
        with self._lock:
            if key in self._stacks:
                raise ValueError(f"stack {key!r} is already open")
            self._stacks[key] = {"count": int(count), "normalize": bool(normalize), "accs": None,
                                 "seen": set(), "errors": [], "future": fut}
        return fut

    def submit(self, req: ShotRequest, arrival: float | None = None) -> Future:
        fut: Future = Future()
        arrival = time.perf_counter() if arrival is None else float(arrival)
        if req.stack is not None:
            with self._lock:
                st = self._stacks.get(req.stack)
                idx = int(req.stack_index)
                if st is None:
                    bad = "stack not open"
                elif not 0 <= idx < st["count"]:
                    bad = "index outside the stack"
                elif idx in st["seen"]:
                    bad = "index given twice"
                else:
                    bad = None
                    st["seen"].add(idx)
            if bad is not None:
                self._finish(fut, ShotResult(req.rid, None, f"rejected: stack {req.stack!r}: {bad}",
                                             arrival, finish=time.perf_counter()))
                return fut
        try:
            imager = self._imager(req)
            job = imager.prepare_batch([req.shot], req.model, modes=req.modes)
            nbytes = request_bytes(req, self.gpu_config)
        except EngineUnavailable:
            raise
        except (ERTMError, ValueError) as exc:
            self._finish(fut, ShotResult(req.rid, None, f"rejected: {exc}", arrival,
                                         finish=time.perf_counter()), req)
            return fut
        if nbytes > self.budget:
            self._finish(fut, ShotResult(req.rid, None,
                                         f"rejected: needs {nbytes / _GIB:.2f} GiB of HBM, "
                                         f"budget {self.budget / _GIB:.2f} GiB", arrival,
                                         finish=time.perf_counter(), hbm_bytes=nbytes), req)
            return fut
        key = _key(req, imager._batch_key(req.model, job.specs[0]))
        lane = req.solver.engine if self.lanes and lanes_supported(imager) else None
        with self._lock:
            if not self._open:
                raise RuntimeError("serving loop is closed")
            self._outstanding += 1
            self._adm.add((req, job, nbytes, key, arrival, fut, lane))
            self._admit()
        return fut

    def drain(self, timeout: float | None = None) -> None:
        """Wait until every submitted request has finished."""
        end = None if timeout is None else time.perf_counter() + timeout
        with self._lock:
            while self._outstanding > 0:
                left = None if end is None else end - time.perf_counter()
                if left is not None and left <= 0:
                    raise TimeoutError("serving loop did not drain in time")
                self._work.wait(left)

    def close(self) -> None:
        self.drain()
        with self._lock:
            self._open = False
            self._work.notify_all()
        for t in self._threads:
            t.join()
        self._post_pool.shutdown(wait=True)

    def __enter__(self) -> "ServingLoop":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def results(self) -> list[ShotResult]:
        with self._lock:
            return list(self._results)

    def stats(self) -> dict[str, Any]:
        res = self.results()
        done = [r for r in res if r.ok]
        out: dict[str, Any] = {"completed": len(done), "failed": len(res) - len(done),
                               "peak_reserved_gib": round(self._adm.peak / _GIB, 3)}
        if done:
            lat = np.array([r.latency for r in done])
            span = max(r.finish for r in done) - min(r.arrival for r in done)
            out.update({"shots_per_s": len(done) / span if span > 0 else math.inf,
                        "p50_s": float(np.percentile(lat, 50)),
                        "p99_s": float(np.percentile(lat, 99)),
                        "mean_group": float(np.mean([r.group_size for r in done]))})
        steps, groups = self._launch_stats
        if groups:
            # lanes advanced per step launch (over every lane tick so far)
            out["lanes_per_launch"] = steps / groups
        return out

    # ------------------------------------------------------------ internals
    def _imager(self, req: ShotRequest) -> RTMImager:
        # several worker threads drive the GPU at once: no captured time loops
        # (another thread's work can invalidate a capture in progress)
        gpu = dataclasses.replace(self.gpu_config, graph_steps=0)
        return RTMImager(req.imaging, req.solver, req.q, gpu)

    def _admit(self) -> None:
        """Hand every assignment that can start now to its worker (called
        with the lock held)."""
        while True:
            got = self._adm.take()
            if got is None:
                return
            kind, w, items = got
            self._inbox[w].append((kind, items))
            self._work.notify_all()

    def _finish(self, fut: Future, res: ShotResult, req: ShotRequest | None = None) -> None:
        with self._lock:
            self._results.append(res)
        if req is not None and req.stack is not None:
            self._fold(req, res)
        fut.set_result(res)

    def _fold(self, req: ShotRequest, res: ShotResult) -> None:
        """Fold a finished stacked request into its stack; resolve the
        stack's Future once all its shots are in."""
        with self._lock:
            st = self._stacks[req.stack]
            if res.ok:
                if st["accs"] is None:
                    st["accs"] = {m: StackAccumulator(st["count"], normalize=st["normalize"])
                                  for m in res.images}
                for m, acc in st["accs"].items():
                    acc.add(req.stack_index, res.images[m])
            else:
                st["errors"].append((req.stack_index, res.error))
            st["received"] = st.get("received", 0) + 1
            if st["received"] < st["count"]:
                return
            del self._stacks[req.stack]
        fut = st["future"]
        if st["errors"]:
            shots = ", ".join(f"{i}: {e}" for i, e in sorted(st["errors"], key=lambda x: x[0]))
            fut.set_exception(PipelineError(f"stack {req.stack!r}: failed shots {shots}"))
        else:
            fut.set_result({m: acc.result() for m, acc in st["accs"].items()})

    def _worker(self, w: int) -> None:
        import torch

        stream = torch.cuda.Stream()
        runner = LaneRunner(max_lanes=self._adm.max_group,
                            variant=self.gpu_config.kernel_variant)
        while True:
            with self._lock:
                while not self._inbox[w] and not runner.lanes and self._open:
                    self._work.wait()
                if not self._inbox[w] and not runner.lanes:
                    return
                inbox, self._inbox[w] = self._inbox[w], []
            with torch.cuda.stream(stream):
                for kind, items in inbox:
                    if kind == "group":
                        self._done(w, items, self._run(items))
                    else:
                        self._add_lane(w, runner, items[0])
                if runner.lanes:
                    self._run_lanes(w, runner)

    def _done(self, w: int, items, outcome) -> None:
        """Finish items assigned to worker ``w`` and admit what now fits."""
        self._release(w, items)
        self._complete(items, outcome)

    def _release(self, w: int, items) -> None:
        """The items' HBM is free: admit what now fits."""
        with self._lock:
            self._adm.release(w, items)
            self._admit()
            self._work.notify_all()

    def _complete(self, items, outcome) -> None:
        for item, res in zip(items, outcome):
            self._finish(item[5], res, item[0])
        with self._lock:
            self._outstanding -= len(items)
            self._work.notify_all()

    def _add_lane(self, w: int, runner: "LaneRunner", item) -> None:
        req, job = item[0], item[1]
        try:
            lane = Lane(self._imager(req), job, tag=item)
        except Exception:  # noqa: BLE001 - set-up failed: run it by itself
            self._done(w, [item], self._run([item]))
            return
        runner.add(lane)

    def _run_lanes(self, w: int, runner: "LaneRunner") -> None:
        """Advance the worker's lanes for up to ``lane_ticks`` steps; finished
        lanes return their results.  If the shared run fails, each lane's
        request is run again by itself, so a bad request fails only itself."""
        steps0, groups0 = runner.lane_steps, runner.step_groups
        try:
            finished = runner.run(max_ticks=self.lane_ticks, post=False)
        except Exception:  # noqa: BLE001 - isolate per request below
            items = [ln.tag for ln in runner.lanes]
            runner.close()
            for item in items:
                r = self._run([item])[0]
                r.rerun_alone = True
                self._done(w, [item], [r])
            return
        with self._lock:
            self._launch_stats[0] += runner.lane_steps - steps0
            self._launch_stats[1] += runner.step_groups - groups0
        for ln in finished:
            # the worker goes on with its other lanes; a post thread waits
            # for this one, frees its HBM and makes its images
            self._post_pool.submit(self._post_lane, w, ln)

    def _post_lane(self, w: int, ln: "Lane") -> None:
        item = ln.tag
        try:
            images = ln.post()
            bad = [m for m, r in images.items() if not np.isfinite(r.image).all()]
            err = "ImagingError: non-finite image" if bad else None
        except Exception as exc:  # noqa: BLE001 - fails this request only
            images, err = None, f"{type(exc).__name__}: {exc}"
            ln.release()
        released = time.perf_counter()
        self._release(w, [item])
        res = ShotResult(item[0].rid, None if err else images, err, item[4], ln.started,
                         time.perf_counter(), ln.mean_group, item[2], released=released)
        self._complete([item], [res])

    def _run(self, group) -> list[ShotResult]:
        start = time.perf_counter()
        try:
            out = self._migrate(group)
            finish = time.perf_counter()
            return [ShotResult(item[0].rid, images, None, item[4], start, finish, len(group),
                               item[2]) for item, images in zip(group, out)]
        except Exception as exc:  # noqa: BLE001 - isolate per request below
            if len(group) == 1:
                item = group[0]
                return [ShotResult(item[0].rid, None, f"{type(exc).__name__}: {exc}", item[4],
                                   start, time.perf_counter(), 1, item[2])]
            # a member broke the group: run each alone so only it fails
            res = []
            for item in group:
                r = self._run([item])[0]
                r.rerun_alone = True
                res.append(r)
            return res

    def _migrate(self, group) -> list[dict[str, ImageResult]]:
        req0, job0 = group[0][0], group[0][1]
        imager = self._imager(req0)
        jobs = [item[1] for item in group]
        merged = dataclasses.replace(job0, shots=[j.shots[0] for j in jobs],
                                     models=[j.models[0] for j in jobs],
                                     specs=[j.specs[0] for j in jobs],
                                     records=[j.records[0] for j in jobs],
                                     waves=[j.waves[0] for j in jobs])
        out = imager.run_batch(merged)
        for images in out:
            for r in images.values():
                if not np.isfinite(r.image).all():
                    raise ImagingError("non-finite image")
        return out
