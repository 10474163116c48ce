"""Lanes: shot migrations on independent schedules advanced in lockstep.

``RTMImager.run_batch`` advances shots that share nt and the checkpoint
interval, one ertm_run per pass.  A lane is one shot's whole migration
written as a list of passes (segments) of its two propagators: the
checkpointing forward pass, then per checkpoint block, last block first, the
replay that re-creates the source field into the replay slots and the
backward pass that injects the record and correlates (ertm/imaging.py:
113-162, :279-324; sub-checkpoints split a block whose replay slots exceed
the replay budget, exactly as ``RTMImager._run_shots`` does).

``LaneRunner`` keeps a set of lanes of one engine on the current stream and
runs them with ertm_run_lanes: every tick advances each lane by one step,
whatever its nt, dt, grid, checkpoint interval, block or pass, and the step
launches of a tick cover every lane in the same kind of pass.  Lanes join
between calls (continuous batching) and leave when their last pass ends.
A lane's image equals the one ``RTMImager.migrate`` gives for its shot
alone: each propagator runs the same kernels on the same inputs in the same
order; only which other shots share its launches changes.

Supported: acoustic and elastic engines, conditions ``zero_lag`` and
``normalized`` (the excitation condition and the viscoacoustic stabiliser
keep their own loops in ``RTMImager``).
"""

from __future__ import annotations

import ctypes as C
import math
import time
from typing import Any

import numpy as np

from . import _lib as L
from . import device as dev
from .exceptions import ImagingError
from .imaging import (CheckpointStore, PreparedBatch, RTMImager, checkpoint_interval,
                      replay_segments, slot_plan, sub_checkpoint_origin)
from .propagation import build_propagator
from .types import ImageResult

_GIB = 1024.0 ** 3
LANE_ENGINES = ("acoustic", "elastic")
SEG_DTYPE = np.dtype(L.ErtmSeg)   # the C layout of ErtmSeg


def lanes_supported(imager: RTMImager) -> bool:
    """Whether ``imager``'s migrations can run as lanes."""
    return (imager.solver_config.engine in LANE_ENGINES
            and str(imager.config.condition).lower() != "excitation")


def lane_slot_room(gpu, cells: int) -> int:
    """Replay slots one migration of ``cells`` cells may hold outside a
    batch: ``GpuConfig.replay_slots`` if set, else what the replay budget
    holds (the serving loop admits requests by these bytes)."""
    if gpu.replay_slots > 0:
        return max(1, int(gpu.replay_slots))
    return max(1, int(float(gpu.replay_budget_gb) * _GIB // (4 * max(cells, 1))))


def lane_replay_slots(imager: RTMImager, interval: int, cells: int, wave=None,
                      nt: int = 0) -> tuple[int, int | None]:
    """(replay slots, kept tail start or None) of one lane (``slot_plan``
    for ``lane_slot_room``).  They only set which steps are replayed and how
    a block is split into replay segments; the image does not depend on
    them."""
    gpu = imager.gpu_config
    room = lane_slot_room(gpu, cells)
    if wave is None:
        return max(1, min(interval, room)), None
    return slot_plan([wave], nt, interval, room, gpu, cells)


class Lane:
    """One shot's migration as propagators, buffers and a pass program.

    Built from an imager and a one-shot ``PreparedBatch`` (records already
    in HBM); ``tag`` is returned with the result."""

    def __init__(self, imager: RTMImager, job: PreparedBatch, tag: Any = None) -> None:
        import torch

        if len(job.shots) != 1:
            raise ImagingError("a lane migrates one shot")
        if not lanes_supported(imager):
            raise ImagingError("lanes run acoustic and elastic zero_lag/normalized migrations")
        self.imager, self.job, self.tag = imager, job, tag
        self.engine = imager.solver_config.engine
        cfg = imager.config
        shot, model, spec = job.shots[0], job.models[0], job.specs[0]
        record = job.records[0]
        self.condition = str(cfg.condition).lower()
        with dev.trusted_models([model]):
            fwd = build_propagator(model, spec, engine=self.engine,
                                   q_options=imager.q_options(compensation=False))
            fwd.set_source(shot.source)
            fwd.set_receivers(shot.receivers)
            back = build_propagator(model, spec, engine=self.engine,
                                    q_options=imager.q_options(compensation=True))
            back.set_source(shot.source)
            back.set_receivers(shot.receivers)
        if fwd._run_kwargs() or back._run_kwargs():
            raise ImagingError("lanes do not run the viscoacoustic stabiliser")
        names = list(record)
        tab = dev.RecvTables()
        tab.build_injection(back._core, shot.receivers.x, shot.receivers.z,
                            [imager._routing(nm) for nm in names], back._k_relaxed_host())
        tab.set_records([record[nm] for nm in names])
        self.fwd, self.back, self.tab = fwd, back, tab
        wave = np.ascontiguousarray(job.waves[0], dtype=np.float32)
        nt = int(len(wave))
        self.nt = nt
        # a kept tail's last forward step (nt) takes the constant last value
        self.wave = np.ascontiguousarray(np.append(wave, wave[-1:]))
        # the interval from the fields-only state size before the first step
        # (ertm/imaging.py:121-124), as record_source_wavefield takes it
        self.per_state = max(int(fwd.state_bytes()), 1)
        interval = checkpoint_interval(nt, self.per_state, float(cfg.checkpoint_mb))
        starts = list(range(0, nt, interval))
        fc, bc = fwd._core, back._core
        self.stride = int(fc.state_floats)
        nz, nx = model.shape
        self.slots, kept = lane_replay_slots(imager, interval, nz * nx, wave, nt)
        self.t_keep = nt if kept is None else kept
        self.arena = torch.empty((len(starts), self.stride), dtype=torch.float32, device="cuda")
        self.store = CheckpointStore(states=[(s, k) for k, s in enumerate(starts)],
                                     interval=interval, bytes_per_state=self.per_state,
                                     arena=self.arena, state_floats=self.stride,
                                     forward_steps=starts[-1] + 1 if kept is None else nt + 1)
        st = int(np.floor((float(shot.source.delay) + float(cfg.start_periods)
                           / max(float(shot.source.f0), 1e-6)) / spec.dt))
        self.start_step = int(np.clip(st, 0, nt))
        rec_illum = L.ILLUM_REC if cfg.receiver_normalization else 0
        illum = rec_illum | (L.ILLUM_SRC if self.condition == "normalized" and cfg.illumination
                             else 0)
        self.rbuf = torch.empty((self.slots, nz, nx), dtype=torch.float32, device="cuda")
        ring = int(imager.gpu_config.image_ring)
        self.ib = dev.ImageBuffers(nz, nx, cfg.separation, job.swave, job.rwaves, self.start_step,
                                   illum, rbuf=self.rbuf, nslots=self.slots, ring=ring,
                                   ring_ld=bc.ld)
        self.fimg = dev.ImageBuffers(nz, nx, cfg.separation, job.swave, (), 0, 0, rbuf=self.rbuf,
                                     nslots=self.slots)
        self.subs = None
        if self.slots < interval:
            nsub = int(math.ceil(interval / self.slots))
            self.subs = torch.empty((nsub, self.stride), dtype=torch.float32, device="cuda")
        self.program = self._program()
        self.segs = self.segments()
        self.replayed = sum(e - a for ph, a, e, *_ in self.program if ph == 1)
        self.back_steps = sum(e - a for ph, a, e, *_ in self.program if ph == 2)
        self.lane_steps = 0      # ticks this lane ran
        self.shared_steps = 0    # sum over those ticks of the lanes advanced with it
        self.started = None
        self.finished = None
        self.result = None
        self._event = None

    def _program(self) -> list[tuple]:
        """(phase, t_begin, t_end, op, img_t0, ck_interval, ck_origin, save,
        restore) per pass, in order; save/restore are (tensor, slot) or None.
        The same passes ``RTMImager._run_shots`` makes for this shot."""
        st, nt, slots, tk = self.store, self.nt, self.slots, self.t_keep
        if tk < nt:
            # checkpoints up to the state after step tk, then the kept tail:
            # forward steps tk + 1 .. nt extract what replay steps tk .. nt - 1
            # would, and the backward pass over it reads them
            prog = [(0, 0, tk + 1, -1, 0, st.interval, 0, (self.arena, 0), None),
                    (0, tk + 1, nt + 1, L.IMG_EXTRACT, tk + 1, 0, 0, None, None),
                    (2, tk, nt, L.IMG_CORRELATE, tk, 0, 0, None, None)]
        else:
            prog = [(0, 0, st.forward_steps, -1, 0, st.interval, 0, (self.arena, 0), None)]
        for blk in range(len(st.states) - 1, -1, -1):
            s0, s1 = st.bounds(blk, nt)
            s1 = min(s1, tk)
            if s1 <= s0:
                continue
            slot = st.states[blk][1]
            segs = replay_segments(s0, s1, slots)
            if len(segs) > 1:
                # sub-checkpoints: the state just before w[a] of each later segment
                prog.append((1, s0, segs[-1][0], -1, 0, slots, sub_checkpoint_origin(segs),
                             (self.subs, 0), (self.arena, slot)))
            for k in range(len(segs) - 1, -1, -1):
                a, e = segs[k]
                src = (self.arena, slot) if k == 0 else (self.subs, k - 1)
                prog.append((1, a, e, L.IMG_EXTRACT, a, 0, 0, None, src))
                prog.append((2, a, e, L.IMG_CORRELATE, a, 0, 0, None, None))
        return prog

    def segments(self) -> np.ndarray:
        """The program as ErtmSeg records (a NumPy structured array of the
        C layout) with ``prop`` 0 for the forward propagator and 1 for the
        backward one; the runner offsets them to the batch's indices."""
        prog = self.program
        segs = np.zeros(len(prog), dtype=SEG_DTYPE)
        segs["phase"] = [g[0] for g in prog]
        segs["prop"] = (segs["phase"] == 2).astype(np.int32)
        segs["t_begin"] = [g[1] for g in prog]
        segs["t_end"] = [g[2] for g in prog]
        segs["op_after"] = [g[3] for g in prog]
        segs["img_t0"] = [g[4] for g in prog]
        segs["ck_interval"] = [g[5] for g in prog]
        segs["ck_origin"] = [g[6] for g in prog]
        row = 4 * self.stride   # bytes per saved state
        segs["save_arena"] = [0 if g[7] is None else L.ptr(g[7][0]) + row * g[7][1] for g in prog]
        segs["save_stride"] = [0 if g[7] is None else self.stride for g in prog]
        segs["restore"] = [0 if g[8] is None else L.ptr(g[8][0]) + row * g[8][1] for g in prog]
        segs["wavelet"] = np.where(segs["phase"] != 2, self.wave.ctypes.data, 0)
        return segs

    @property
    def done(self) -> bool:
        last = self.segs[-1]
        return bool(last["done"] >= last["t_end"] - last["t_begin"])

    @property
    def cell_updates(self) -> float:
        nz, nx = self.job.models[0].shape
        return float((self.store.forward_steps + self.replayed + self.back_steps) * nz * nx)

    @property
    def mean_group(self) -> float:
        """Lanes advanced per tick, averaged over this lane's ticks."""
        return self.shared_steps / self.lane_steps if self.lane_steps else 1.0

    def collect(self, done_event=None) -> None:
        """Record that the lane's passes are all enqueued on the current
        stream (``done_event`` recorded after them, or a new one): the
        metadata is taken now; ``post`` waits for the event, copies the
        accumulators to the host, releases the device buffers and makes the
        images.  With device post-processing the images are made here, on
        the current stream."""
        import torch

        im, job = self.imager, self.job
        for p, steps in ((self.fwd, self.store.forward_steps), (self.back, self.back_steps)):
            p.pml.touched = True
            p.step_index += steps
        info = dict(store=self.store, start_step=self.start_step, prop=self.fwd, back=self.back,
                    replayed_steps=self.replayed, nt=self.nt, cell_updates=self.cell_updates)
        self._metas = [[im._meta(job.specs[0], self.engine, job.shots[0], info, self.condition, m)
                        for m in job.modes]]
        if im.gpu_config.device_postprocess:
            ib = self.ib
            self.result = self._post([("device", list(ib.num), ib.ill, list(ib.rill))])
            torch.cuda.current_stream().synchronize()
            self.release()
            self._event = None
            return
        if done_event is None:
            done_event = torch.cuda.Event()
            done_event.record()
        self._event = done_event

    def post(self) -> dict[str, ImageResult]:
        """The normalised, post-processed images {mode: ImageResult} after
        ``collect``; any thread.  The device buffers are released here."""
        if self.result is not None:
            return self.result
        import torch

        im, ib = self.imager, self.ib
        self._event.synchronize()
        use_ill = self.condition == "normalized" and bool(im.config.illumination)
        use_rill = bool(im.config.receiver_normalization)
        # the copies run on this thread's side stream: the lane's work is done
        raw = ("fetch", list(ib.num), ib.ill if use_ill else None,
               list(ib.rill) if use_rill else None, torch.cuda.default_stream())
        raws = im._fetch_host([raw])
        self.release()
        self._event = None
        self.result = self._post(raws)
        return self.result

    def _post(self, raws) -> dict[str, ImageResult]:
        results: list = [None]
        self.imager._post_group(self.job, [0], raws, self._metas, self.condition, results)
        return results[0]

    def finish(self) -> dict[str, ImageResult]:
        """``collect`` then ``post``."""
        self.collect()
        return self.post()

    def release(self) -> None:
        """Drop the lane's device buffers (its image has been taken)."""
        self.arena = self.rbuf = self.subs = self.ib = self.fimg = self.tab = None
        self.fwd = self.back = None
        self.store.arena = None


class LaneRunner:
    """Lanes of one engine advanced together on the current stream.

    ``add`` lanes at any time; ``run`` advances them until one finishes (or
    ``max_ticks``) and returns the finished ones with their images."""

    def __init__(self, max_lanes: int = 64, variant: tuple[int, int] | None = None) -> None:
        if max_lanes < 1:
            raise ValueError("max_lanes must be positive")
        self.max_lanes = int(max_lanes)
        self.variant = variant
        self.lanes: list[Lane] = []
        self._batch = None
        self._curs: list[int] = []
        self._next_id = 0
        self.ticks = 0
        self.lane_steps = 0
        self.step_groups = 0

    @property
    def engine(self) -> str | None:
        return self.lanes[0].engine if self.lanes else None

    def __len__(self) -> int:
        return len(self.lanes)

    def can_take(self, engine: str) -> bool:
        return len(self.lanes) < self.max_lanes and self.engine in (None, engine)

    def add(self, lane: Lane) -> None:
        if not self.can_take(lane.engine):
            raise ImagingError("lane runner is full or runs another engine")
        lane.lane_id = self._next_id
        self._next_id += 1
        self.lanes.append(lane)
        self._close_batch()

    def _close_batch(self) -> None:
        """Close the current batch (its device tables return to the library's
        pool once the launches already queued have run: no wait here)."""
        if self._batch is not None:
            for k, c in enumerate(self._batch.cores):
                c.cur = self._curs[k]
            self._batch.close()
            self._batch = None

    def _open_batch(self):
        cores, recvs, imgs = [], [], []
        for ln in self.lanes:
            cores += [ln.fwd._core, ln.back._core]
            recvs += [dev.RecvTables(), ln.tab]
            imgs += [ln.fimg, ln.ib]
        variant = self.variant
        if variant is None:
            from . import tuning

            c0 = cores[0]
            # a fused shape: ertm_run_lanes sizes each launch's chunks itself
            if tuning.fused_supported(c0.engine, c0.order, c0.nmech):
                variant = (2, 48)
        self._batch = dev.Batch(cores, recvs=recvs, imgs=imgs, variant=variant)
        self._curs = [int(c.cur) for c in cores]
        # every lane's program in one array: lane k's propagators are 2k, 2k+1
        counts = np.array([len(ln.segs) for ln in self.lanes])
        self._arr = np.concatenate([ln.segs for ln in self.lanes])
        k = np.repeat(np.arange(len(self.lanes), dtype=np.int32), counts)
        self._arr["prop"] += 2 * k
        self._arr["lane"] = np.repeat(np.array([ln.lane_id for ln in self.lanes], np.int32), counts)
        self._bounds = np.concatenate([[0], np.cumsum(counts)])

    def run(self, max_ticks: int = 0, post: bool = True) -> list[Lane]:
        """Advance every lane until one finishes (or ``max_ticks`` ticks);
        returns the finished lanes (removed from the runner), images in
        ``lane.result``.  With ``post`` False the call does not wait for
        them: ``lane.post()`` (any thread) waits, takes the images and
        releases the lane's device memory, while the other lanes go on."""
        import torch

        if not self.lanes:
            return []
        if self._batch is None:
            self._open_batch()
        now = time.perf_counter()
        for ln in self.lanes:
            if ln.started is None:
                ln.started = now
        nlanes = len(self.lanes)
        segs = self._arr.ctypes.data_as(C.POINTER(L.ErtmSeg))
        ticks, steps, groups = self._batch.run_lanes(segs, len(self._arr), self._curs,
                                                     max_ticks=max_ticks, stop_at_lane_end=True)
        self.ticks += ticks
        self.lane_steps += steps
        self.step_groups += groups
        done = []
        for k, ln in enumerate(self.lanes):
            ln.segs["done"] = self._arr["done"][self._bounds[k]:self._bounds[k + 1]]
            # the call ends at the tick a lane finishes: every lane ran every tick
            ln.lane_steps += ticks
            ln.shared_steps += ticks * nlanes
            if ln.done:
                done.append(ln)
        if not done:
            return []
        ev = torch.cuda.Event()
        ev.record()
        self._close_batch()
        self.lanes = [ln for ln in self.lanes if ln not in done]
        now = time.perf_counter()
        for ln in done:
            ln.finished = now
            ln.collect(ev)
            if post:
                ln.post()
        return done

    def run_all(self) -> list[Lane]:
        """Run until every lane has finished; lanes in finishing order."""
        out = []
        while self.lanes:
            out += self.run()
        return out

    def close(self) -> None:
        self._close_batch()
        self.lanes = []


def migrate_lanes(items: list[tuple[RTMImager, PreparedBatch]], *, max_lanes: int = 64
                  ) -> list[dict[str, ImageResult]]:
    """Migrate one-shot prepared jobs (each with its own imager: nt, dt,
    grid and configs may all differ) as lanes; one runner per engine, run
    one after the other.  Returns {mode: ImageResult} per item, in order."""
    out: list = [None] * len(items)
    by_engine: dict[str, list[int]] = {}
    for i, (im, _job) in enumerate(items):
        by_engine.setdefault(im.solver_config.engine, []).append(i)
    for idx in by_engine.values():
        for c0 in range(0, len(idx), max_lanes):
            runner = LaneRunner(max_lanes=max_lanes)
            for i in idx[c0:c0 + max_lanes]:
                im, job = items[i]
                runner.add(Lane(im, job, tag=i))
            for ln in runner.run_all():
                out[ln.tag] = ln.result
            runner.close()
    return out
