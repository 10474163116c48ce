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
* **continuous batching**: when a request is admitted, waiting requests that
  can share its launches (same engine, configs, nt and checkpoint interval --
  see ``RTMImager.migrate_batch``) and that also fit join it, up to
  ``max_group`` shots advanced by one launch per time step;
* **concurrency**: several launch groups run at once, each driven by its own
  host thread on its own HIP stream; HBM freed by a finished group admits the
  next requests at once.  Wavefields, checkpoints and partial images stay on
  the device for a request's whole life;
* **isolation**: a request that fails validation fails alone; if a launch
  group fails while running, its members are re-run one by one, so a bad
  request fails only itself.  A shot's image does not depend on what else
  is running (each image equals the one ``migrate`` gives for that shot
  alone);
* **stacking as shots finish**: requests that name a stack (``stack``,
  ``stack_index``) are folded into that stack's float64 sum as they finish
  (``open_stack``), in stack order through a reorder buffer, so the stacked
  image equals ``stack_images`` over the shots in stack order
  (ertm/pipeline.py:133-161).

Results carry arrival, start and finish times; ``stats`` reports shots/s and
the p50/p99 latency.
"""

from __future__ import annotations

import dataclasses
import math
import threading
import time
from collections import deque
from concurrent.futures import Future
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import numerics as num
from ._lib import EngineUnavailable
from .config import GpuConfig, ImagingConfig, QConfig, SolverConfig
from .exceptions import ERTMError, ImagingError, PipelineError
from .imaging import RTMImager, checkpoint_interval
from .pipeline import StackAccumulator
from .propagation import PropagationSpec
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
    group_size: int = 0
    hbm_bytes: int = 0
    rerun_alone: bool = False   # its launch group failed; it was re-run by itself

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
    props = 2 * (2 * state + media)                               # forward + backward
    nt = int(spec.nt)
    interval = checkpoint_interval(nt, nf * nzp * nxp * 4, float(req.imaging.checkpoint_mb))
    arena = math.ceil(nt / interval) * state
    cells = nz * nx
    slots = min(interval, max(1, int(float(gpu.replay_budget_gb) * _GIB // (4 * cells))))
    if gpu.replay_slots > 0:
        slots = min(interval, int(gpu.replay_slots))
    sub = math.ceil(interval / slots) * state if slots < interval else 0
    nmodes = len(req.modes or (req.imaging.mode,))
    accum = (2 * nmodes + 1 + 3) * cells * 2                      # float64 -> 2 floats
    # receiver-field ring of the backward pass (elastic and acoustic engines)
    ring = int(gpu.image_ring) * nmodes * nz * ld if engine in ("elastic", "acoustic") else 0
    nrec = int(req.shot.receivers.nrec)
    ncomp = max(1, len(req.shot.components))
    record = nt * nrec * ncomp + 16 * nrec * ncomp
    floats = props + arena + slots * cells + sub + accum + ring + record
    return int(floats * 4 * 1.05) + (8 << 20)


def _key(req: ShotRequest, imager_key) -> tuple:
    def tup(c):
        return None if c is None else dataclasses.astuple(c)
    return (tup(req.solver), tup(req.imaging), tup(req.q),
            tuple(req.modes or (req.imaging.mode,)), imager_key)


class Admission:
    """First-come first-served admission by bytes, with same-key companions.

    Items are (request, job, bytes, key, arrival, future) tuples; only bytes
    and key are read here.  ``take`` returns the next launch group to start,
    or None when the oldest waiting item does not fit or no stream is idle.
    """

    def __init__(self, budget: int, streams: int, max_group: int) -> None:
        self.budget = int(budget)
        self.idle = int(streams)
        self.max_group = int(max_group)
        self.pending: deque = deque()
        self.reserved = 0
        self.peak = 0

    def add(self, item) -> None:
        self.pending.append(item)

    def take(self):
        if not self.pending or self.idle <= 0:
            return None
        head = self.pending[0]
        if self.reserved + head[2] > self.budget:
            return None
        group = [head]
        used = head[2]
        for item in list(self.pending)[1:]:
            if len(group) >= self.max_group:
                break
            if item[3] == head[3] and self.reserved + used + item[2] <= self.budget:
                group.append(item)
                used += item[2]
        taken = {id(item) for item in group}
        self.pending = deque(item for item in self.pending if id(item) not in taken)
        self.reserved += used
        self.peak = max(self.peak, self.reserved)
        self.idle -= 1
        return group

    def release(self, group) -> None:
        self.reserved -= sum(item[2] for item in group)
        self.idle += 1


class ServingLoop:
    """Admit, batch and run migration requests on the current GPU.

    ``submit`` returns a Future of a ``ShotResult``.  ``streams`` launch
    groups run concurrently; ``hbm_budget_gb`` caps the HBM reserved by
    admitted requests; ``max_group`` caps the shots per launch group.
    """

    def __init__(self, hbm_budget_gb: float = 200.0, streams: int = 4, max_group: int = 64,
                 gpu_config: GpuConfig | None = None) -> None:
        if streams < 1 or max_group < 1 or hbm_budget_gb <= 0:
            raise ValueError("streams, max_group and hbm_budget_gb must be positive")
        self.budget = int(float(hbm_budget_gb) * _GIB)
        self.gpu_config = gpu_config or GpuConfig()
        self._lock = threading.Lock()
        self._work = threading.Condition(self._lock)
        self._adm = Admission(self.budget, streams, max_group)
        self._groups: deque = deque()
        self._open = True
        self._results: list[ShotResult] = []
        self._outstanding = 0
        self._stacks: dict[Any, dict[str, Any]] = {}
        self._threads = [threading.Thread(target=self._worker, name=f"ertm-serve-{k}",
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
        with self._lock:
            if not self._open:
                raise RuntimeError("serving loop is closed")
            self._outstanding += 1
            self._adm.add((req, job, nbytes, key, arrival, fut))
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
        return out

    # ------------------------------------------------------------ internals
    def _imager(self, req: ShotRequest) -> RTMImager:
        # several worker threads drive the GPU at once: no captured time loops
        # (another thread's work can invalidate a capture in progress)
        gpu = dataclasses.replace(self.gpu_config, graph_steps=0)
        return RTMImager(req.imaging, req.solver, req.q, gpu)

    def _admit(self) -> None:
        """Start every group that can start now (called with the lock held)."""
        while True:
            group = self._adm.take()
            if group is None:
                return
            self._groups.append(group)
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

    def _worker(self) -> None:
        import torch

        stream = torch.cuda.Stream()
        while True:
            with self._lock:
                while not self._groups and self._open:
                    self._work.wait()
                if not self._groups:
                    return
                group = self._groups.popleft()
            with torch.cuda.stream(stream):
                outcome = self._run(group)
            with self._lock:
                self._adm.release(group)
            for item, res in zip(group, outcome):
                self._finish(item[5], res, item[0])
            with self._lock:
                self._outstanding -= len(group)
                self._admit()
                self._work.notify_all()

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
