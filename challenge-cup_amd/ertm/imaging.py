"""Reverse-time migration of one shot on the GPU (ertm/imaging.py).

``RTMImager.migrate`` keeps the reference's algorithm and quirks
(ertm/imaging.py:205-349):

1. forward propagation of the source wavelet with checkpoints every
   ``interval`` steps, the interval coming from the reference's budget rule
   on the fields-only state size (:113-137);
2. blocks in reverse: each block is replayed from its checkpoint, with the
   reference's replay convention (the state saved after step s is stepped
   with w[s] again, :140-162), writing the source-side image field of every
   step into HBM replay slots;
3. the backward propagator, fed the time-reversed record at the receivers,
   steps through the block in reverse while the imaging condition
   (zero_lag / normalized / excitation) is accumulated in float64 on the GPU;
4. normalisation, Laplacian, wavenumber low-pass and the PML mask (once per
   shot, float64 on the host, as the reference does).

Everything between the upload of the record and the download of the
accumulators stays in HBM: checkpoints are device-to-device copies into an
arena, replayed image fields live in replay slots, accumulators never leave
the device until the end.  When one block's replay slots would not fit the
replay budget, the block is replayed in segments from sub-checkpoints
(nested replay); the segments reproduce the same fields bit for bit.
"""

from __future__ import annotations

import math
import os
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Any, Callable

import numpy as np

from . import _lib as L
from . import device as dev
from . import numerics as num
from .config import GpuConfig, ImagingConfig, QConfig, SolverConfig
from .exceptions import ImagingError
from .propagation import (
    PropagationSpec,
    Propagator2D,
    QOptions,
    batch_after_run,
    batch_run_kwargs,
    build_propagator,
    check_hard_stability,
    check_stability,
    source_wavelet,
)
from .types import ElasticModel, ImageResult, ShotRecord

#: imaging mode -> (source wave, receiver wave), ertm/imaging.py:47-52
WAVE_MODES: dict[str, tuple[str, str]] = {
    "pp": ("P", "P"),
    "ps": ("P", "S"),
    "sp": ("S", "P"),
    "ss": ("S", "S"),
}
_ACOUSTIC_MODES = {"pp"}


def zero_lag_condition(source_wavefield: np.ndarray, receiver_wavefield: np.ndarray) -> np.ndarray:
    """Integrand W_s * W_r of the zero-lag cross-correlation (ertm/imaging.py:61-66)."""
    return np.asarray(source_wavefield) * np.asarray(receiver_wavefield)


def normalized_terms(source_wavefield, receiver_wavefield):
    """(W_s W_r, W_s^2, W_r^2) (ertm/imaging.py:69-79)."""
    s = np.asarray(source_wavefield)
    r = np.asarray(receiver_wavefield)
    return s * r, s * s, r * r


def excitation_terms(source_amplitude, receiver_wavefield) -> np.ndarray:
    """A_s(x) W_r(x, t_exc) (ertm/imaging.py:82-87)."""
    return np.asarray(source_amplitude) * np.asarray(receiver_wavefield)


# ---------------------------------------------------------------------------
# checkpoints
# ---------------------------------------------------------------------------
@dataclass
class CheckpointStore:
    """Source-wavefield checkpoints held in an HBM arena (ertm/imaging.py:93-110).

    ``states`` lists (step, arena slot); ``bytes_per_state`` is the reference's
    budgeting size (fields only, see record_source_wavefield); ``arena_bytes``
    is what the arena really holds (fields + CPML strips per state);
    ``state_floats`` is the arena's slot stride (the largest state of the
    shots recorded together).
    """

    states: list[tuple[int, int]] = field(default_factory=list)
    interval: int = 1
    bytes_per_state: int = 0
    arena: Any = None
    state_floats: int = 0
    forward_steps: int = 0  # steps the checkpointing pass ran

    @property
    def memory_mb(self) -> float:
        return len(self.states) * self.bytes_per_state / (1024.0 ** 2)

    @property
    def arena_bytes(self) -> int:
        return 0 if self.arena is None else int(self.arena.numel() * 4)

    def summary(self) -> dict[str, Any]:
        return {"interval": int(self.interval), "n_states": int(len(self.states)),
                "memory_mb": round(self.memory_mb, 3),
                "device_arena_mb": round(self.arena_bytes / 1024.0 ** 2, 3)}

    def bounds(self, b: int, nt: int) -> tuple[int, int]:
        start = self.states[b][0]
        end = self.states[b + 1][0] if b + 1 < len(self.states) else nt
        return start, end


@dataclass
class PreparedBatch:
    """Shots ready to migrate: validated, specs resolved, records aligned to
    (nt, dt) and resident in HBM as float32 (nt, nrec) tensors."""

    shots: list
    models: list
    specs: list
    records: list          # per shot: {record name: device tensor}
    waves: list            # per shot: float32 (nt,) source wavelet
    modes: tuple
    swave: str
    rwaves: tuple


def checkpoint_interval(nt: int, per_state: int, budget_mb: float) -> int:
    """ertm/imaging.py:121-124."""
    budget = max(float(budget_mb), 1.0) * 1024.0 ** 2
    max_states = max(1, int(budget // max(int(per_state), 1)))
    return max(1, int(np.ceil(nt / max_states)))


def replay_segments(s0: int, s1: int, slots: int) -> list[tuple[int, int]]:
    """Block [s0, s1) split into replay segments of at most ``slots`` steps,
    the short one (if any) first.

    A block longer than its replay slots is replayed segment by segment,
    last segment first, each from a sub-checkpoint that one pass from s0 to
    the start of the last segment saves.  That pass is the overhead: with
    the short segment first it runs s1 - s0 - slots steps, with the short
    segment last up to s1 - s0 - 1.  Later segments all hold ``slots``
    steps, so their sub-checkpoints are ``slots`` steps apart, the first
    just before the second segment (``sub_checkpoint_origin``)."""
    slots = max(1, int(slots))
    first = (s1 - s0) % slots or slots
    bounds = [s0] + list(range(s0 + first, s1 + 1, slots))
    return list(zip(bounds[:-1], bounds[1:]))


def sub_checkpoint_origin(segs: list[tuple[int, int]]) -> int:
    """The step whose state the first sub-checkpoint holds (the state just
    before the second segment's first step)."""
    return segs[1][0] - 1


def constant_block_start(wavelets, nt: int, interval: int) -> int | None:
    """The first checkpoint block start s0 from which every wavelet is
    bitwise constant up to step nt - 1, or None when there is none before
    nt.  From s0 on, replaying a block restarted from the state after step
    s0 steps it with wavelet[s0], wavelet[s0 + 1], ... (ertm/imaging.py:
    150-160) exactly as the forward pass stepped it with wavelet[s0 + 1],
    ...: replay step t repeats forward step t + 1, and a replay run on past
    a block's end produces the next block's replayed fields."""
    nt, interval = int(nt), max(1, int(interval))
    z = 0
    for w in wavelets:
        u = np.ascontiguousarray(np.asarray(w, dtype=np.float32)[:nt]).view(np.uint32)
        if u.size < nt:
            return None
        diff = np.nonzero(u != u[nt - 1])[0]
        z = max(z, int(diff[-1]) + 1 if diff.size else 0)
    first = -(-z // interval) * interval
    return first if first < nt else None


def keep_tail_start(wavelets, nt: int, interval: int, room: int) -> int | None:
    """First step T of the tail whose source fields the checkpointing pass
    keeps (``record_source_wavefields``' ``tail``), or None.

    Where every wavelet is bitwise constant from a block start on
    (``constant_block_start``), replay step t extracts what forward step
    t + 1 does.  T is the latest of that block start and nt - ``room`` (the
    replay slots there are); None when the tail would hold no more than the
    last block, whose forward steps the pass otherwise skips."""
    nt, interval = int(nt), max(1, int(interval))
    last = (nt - 1) // interval * interval
    first = constant_block_start(wavelets, nt, interval)
    if first is None:
        return None
    t = max(first, nt - max(int(room), 0))
    return t if t < last else None


def slot_plan(wavelets, nt: int, interval: int, room: int, gpu: GpuConfig,
              cells: int) -> tuple[int, int | None]:
    """(replay slots, kept tail start or None) of a migration of ``cells``
    cells (summed over the shots advanced together) whose slots may hold
    ``room`` steps: a whole block (fewer when ``room`` is smaller), or the
    kept tail when it is longer; the tail fills at most
    ``gpu.tail_budget_gb`` (None: all of ``room``; 0: the whole-block
    slots)."""
    slots = max(1, min(int(interval), int(room)))
    t_keep = None
    if gpu.keep_tail:
        t_room = int(room)
        if gpu.tail_budget_gb is not None:
            extra = int(float(gpu.tail_budget_gb) * 1024 ** 3 // (4 * max(int(cells), 1)))
            t_room = min(t_room, max(slots, extra))
        t_keep = keep_tail_start(wavelets, nt, interval, t_room)
    if t_keep is not None:
        slots = max(slots, int(nt) - t_keep)
    return slots, t_keep


def record_source_wavefield(prop: Propagator2D, wavelet: np.ndarray, *,
                            budget_mb: float) -> CheckpointStore:
    """Forward-propagate and checkpoint into an HBM arena (ertm/imaging.py:113-137).

    The interval is computed from ``state_bytes()`` taken before the first
    step (fields only: the CPML memories do not exist yet), exactly as the
    reference does; the arena itself stores fields and CPML strips.
    """
    return record_source_wavefields([prop], [wavelet], budget_mb=budget_mb)[0]


def record_source_wavefields(props: list[Propagator2D], wavelets: list[np.ndarray], *,
                             budget_mb: float, variant=None, graph_steps: int = 0,
                             during=None, to_last_checkpoint: bool = False,
                             two_step=(0, 0), tail=None,
                             end_after: int | None = None) -> list[CheckpointStore]:
    """``record_source_wavefield`` for several shots advanced in one launch per step.

    All shots must share nt and the checkpoint interval (the interval follows
    from nt and the fields-only state size, so shots on grids of the same
    size always do).  ``during``, if given, is called once the pass has been
    launched and before waiting for it (host set-up to overlap with it).
    ``to_last_checkpoint`` ends the pass at the step that saves the last
    checkpoint instead of at nt: a migration restarts every block from its
    checkpoint, so the steps after the last one produce nothing it reads
    (the checkpoints are the same either way).  ``end_after`` (a block
    start) ends it after the step that saves that block's checkpoint.

    ``tail = (T, imgs)`` (``keep_tail_start``): the pass saves the
    checkpoints up to the state after step T, then runs steps T + 1 .. nt
    with the source-field extraction into the replay slots of ``imgs``
    (their t0 = T + 1), so slot t - T holds what replaying step t would
    extract; step nt takes the wavelet's constant last value.
    """
    import torch

    nts = {int(len(w)) for w in wavelets}
    if len(nts) != 1:
        raise ImagingError("batched shots must share nt")
    nt = nts.pop()
    intervals = {checkpoint_interval(nt, max(int(p.state_bytes()), 1), budget_mb) for p in props}
    if len(intervals) != 1:
        raise ImagingError("batched shots must share the checkpoint interval")
    interval = intervals.pop()
    starts = list(range(0, nt, interval))
    cores = [p._core for p in props]
    # one stride for every arena (grids may differ): the largest state
    stride = max(c.state_floats for c in cores)
    arenas = [torch.empty((len(starts), stride), dtype=torch.float32, device="cuda")
              for _c in cores]
    per_state = [max(int(p.state_bytes()), 1) for p in props]
    step0 = {p.step_index for p in props}
    curs = {c.cur for c in cores}
    if len(step0) != 1 or len(curs) != 1:
        raise ImagingError("batched propagators must be at the same step")
    if tail is not None and end_after is not None:
        raise ImagingError("a kept tail runs the pass to nt: no end_after with it")
    if tail is not None and not 0 <= int(tail[0]) < nt - 1:
        raise ImagingError(f"kept tail start {tail[0]} outside [0, {nt - 1})")
    if end_after is not None and int(end_after) not in starts:
        raise ImagingError(f"end_after {end_after} is not a checkpoint block start")
    b = dev.Batch(cores, imgs=tail[1] if tail is not None else None, variant=variant,
                  graph_steps=graph_steps, two_step=two_step)
    if variant is None:
        from . import tuning

        # one tuning pass per shape and process; later batches look it up
        if tuning.maybe_autotune(b, 3.0 * nt * sum(c.nz * c.nx for c in cores)) is not None:
            b.set_variant(*tuning.lookup(cores[0].engine, cores[0].order,
                                         max(c.nzp for c in cores), max(c.nxp for c in cores),
                                         len(cores), cores[0].nmech))
    ws = [np.asarray(w, dtype=np.float32) for w in wavelets]
    kw = batch_run_kwargs(props)
    # the checkpoint of start s holds the state after step s
    si0 = step0.pop()
    if tail is None:
        t_end = starts[-1] + 1 if to_last_checkpoint else nt
        if end_after is not None:
            t_end = int(end_after) + 1
        cur = b.run(0, 0, t_end, cores[0].cur, wavelets=ws, ck_interval=interval, ck_origin=0,
                    save=arenas, save_stride=stride, step_index0=si0, **kw)
    else:
        t_keep = int(tail[0])
        cur = b.run(0, 0, t_keep + 1, cores[0].cur, wavelets=ws, ck_interval=interval,
                    ck_origin=0, save=arenas, save_stride=stride, step_index0=si0, **kw)
        # no checkpoint is read from here on: every later block lies in the tail
        wz = [np.append(w, w[-1:]) for w in ws]
        t_end = nt + 1
        cur = b.run(0, t_keep + 1, t_end, cur, wavelets=wz, op_after=L.IMG_EXTRACT,
                    step_index0=si0 + t_keep + 1, **kw)
    try:
        if during is not None:
            during()
    finally:
        b.close()
    batch_after_run(props, kw)
    stores = []
    for p, c, arena, ps in zip(props, cores, arenas, per_state):
        c.cur = cur
        p.pml.touched = True
        p.step_index += t_end
        stores.append(CheckpointStore(states=[(s, k) for k, s in enumerate(starts)],
                                      interval=interval, bytes_per_state=ps, arena=arena,
                                      state_floats=stride, forward_steps=t_end))
    return stores


_SIDE_STREAMS: dict = {}
_SIDE_LOCK = threading.Lock()
# Share of the free HBM the replay slots of one batch may take once the
# checkpoints, rings, backward states and accumulators are set aside (the
# rest covers sub-checkpoints, the stabiliser's FFT buffers, receiver tables,
# post-processing and the caching allocator's rounding).
_REPLAY_FREE_SHARE = 0.9


def side_stream(role: str, launch=None):
    """The stream for set-up or copy work beside passes launched on ``launch``
    (default: the current stream).

    For the device's default stream: a side stream, one per role and device,
    kept for the life of the process (default priority), so that the work
    overlaps passes queued on the default stream and device blocks the
    caching allocator keeps for it are reused.  For any other stream:
    ``launch`` itself.  Callers that launch on streams of their own (the
    serving workers) already run several streams, and with a few hardware
    queues per process more streams would share queues with other workers'
    kernels and wait behind them."""
    import torch

    cur = torch.cuda.current_stream() if launch is None else launch
    if cur.cuda_stream != torch.cuda.default_stream(cur.device).cuda_stream:
        return cur
    key = (role, cur.device_index)
    with _SIDE_LOCK:
        s = _SIDE_STREAMS.get(key)
        if s is None:
            s = torch.cuda.Stream(device=cur.device)
            _SIDE_STREAMS[key] = s
    return s


class _HostTask:
    """``fn()`` on a host thread, on the caller's device; ``join()`` waits
    and re-raises what it raised.

    For set-up that overlaps a pass: a pass is one library call, which
    releases the GIL while it queues its launches, and that call returns
    only once the GPU's queue has drained to a few hundred launches, so
    work started after it overlaps only that tail."""

    def __init__(self, fn: Callable[[], None]):
        import torch

        device = torch.cuda.current_device()
        self._exc: BaseException | None = None

        def run():
            try:
                torch.cuda.set_device(device)
                fn()
            except BaseException as e:  # re-raised by join()
                self._exc = e

        self._thread = threading.Thread(target=run, name="ertm-setup", daemon=True)
        self._thread.start()

    def wait(self) -> None:
        self._thread.join()

    def join(self) -> None:
        self._thread.join()
        if self._exc is not None:
            raise self._exc


def _host_workers(n_items: int) -> int:
    return max(1, min(n_items, 16, os.cpu_count() or 1))


def _common(values: set, what: str):
    if len(values) != 1:
        raise ImagingError(f"batched shots must share the {what}")
    return next(iter(values))


def replay_block(prop: Propagator2D, wavelet: np.ndarray, store: CheckpointStore,
                 block_index: int, separation: str) -> list[dict[str, np.ndarray]]:
    """Host copies of the image fields of one replayed block (ertm/imaging.py:140-162).

    Drop-in API for callers that want the fields on the host; migrate keeps
    them in HBM replay slots instead.
    """
    import torch

    core = prop._core
    nt = int(len(wavelet))
    start, end = store.bounds(block_index, nt)
    k = end - start
    waves = ("P", "S") if prop._engine == "elastic" else ("P",)
    bufs = {wv: torch.zeros((max(k, 1), core.nz, core.nx), dtype=torch.float32, device="cuda")
            for wv in waves}
    w = np.asarray(wavelet, dtype=np.float32)
    # first pass: the P (or only) field; second pass (elastic): S
    for n, wv in enumerate(waves):
        img = dev.ImageBuffers(core.nz, core.nx, separation, wv, (), 0, False, rbuf=bufs[wv],
                               nslots=k)
        img.s.t0 = start
        b = dev.Batch([core], imgs=[img])
        core.cur = b.run(1, start, end, core.cur, wavelets=[w], op_after=L.IMG_EXTRACT,
                         restore=[store.arena], restore_stride=store.state_floats,
                         restore_slot=store.states[block_index][1])
        b.close()
    prop.step_index = store.states[block_index][0] + 1 + k
    host = {wv: bufs[wv].cpu().numpy() for wv in waves}
    return [{wv: host[wv][i] for wv in waves} for i in range(k)]


# ---------------------------------------------------------------------------
# imager
# ---------------------------------------------------------------------------
class RTMImager:
    """Single-shot two-way-wave RTM on the GPU (ertm/imaging.py:168-563)."""

    name = "rtm-imager"

    def __init__(self, imaging_config: ImagingConfig, solver_config: SolverConfig | None = None,
                 q_config: QConfig | None = None, gpu_config: GpuConfig | None = None) -> None:
        self.config = imaging_config
        self.config.validate()
        self.solver_config = solver_config or SolverConfig()
        self.solver_config.validate()
        self.q_config = q_config or QConfig()
        self.q_config.validate()
        self.gpu_config = gpu_config or GpuConfig()
        self.gpu_config.validate()
        self._validate_engine_mode()
        self.last_timing: dict[str, float] = {}
        # replay slots per batch shape: set by its first batch, never grown
        self._rooms: dict[tuple, int] = {}

    def q_options(self, *, compensation: bool = False) -> QOptions | None:
        """GSLS options for the viscoacoustic engine (ertm/imaging.py:187-202)."""
        if self.solver_config.engine != "viscoacoustic":
            return None
        q = self.q_config
        return QOptions(enabled=True, f_ref=float(q.f_ref), n_mechanisms=int(q.n_mechanisms),
                        band_ratio=float(q.band_ratio),
                        compensation=bool(compensation and q.enabled), alpha_c=float(q.alpha_c),
                        amplitude_cap=float(q.amplitude_cap),
                        lowpass_interval=int(q.lowpass_interval),
                        f_max=None if q.f_max is None else float(q.f_max),
                        gain_budget_db=float(q.gain_budget_db))

    # ------------------------------------------------------------ public API
    def migrate(self, shot: ShotRecord, model: ElasticModel, *, wavelet: np.ndarray | None = None,
                preprocess: Callable[[np.ndarray, float], np.ndarray] | None = None) -> ImageResult:
        """Migrate one shot record into a depth image (ertm/imaging.py:205-349)."""
        return self.migrate_modes(shot, model, (self.config.mode,), wavelet=wavelet,
                                  preprocess=preprocess)[self.config.mode]

    def migrate_modes(self, shot: ShotRecord, model: ElasticModel, modes: tuple[str, ...], *,
                      wavelet: np.ndarray | None = None,
                      preprocess: Callable[[np.ndarray, float], np.ndarray] | None = None
                      ) -> dict[str, ImageResult]:
        """Several wave modes sharing the source wave (e.g. pp + ps) in one pass.

        Each image equals what ``migrate`` returns for that mode alone; the
        forward, replay and backward propagations are shared.
        """
        return self.migrate_batch([shot], [model], modes=modes,
                                  wavelets=None if wavelet is None else [wavelet],
                                  preprocess=preprocess)[0]

    def migrate_batch(self, shots: list[ShotRecord], models: list[ElasticModel] | ElasticModel,
                      *, modes: tuple[str, ...] | None = None,
                      wavelets: list[np.ndarray] | None = None,
                      preprocess: Callable[[np.ndarray, float], np.ndarray] | None = None
                      ) -> list[dict[str, ImageResult]]:
        """Migrate many shots; shots that share nt and the checkpoint interval
        are advanced together, up to ``gpu.max_batch`` per launch, on grids of
        any (also different) sizes.  Every result equals what
        ``migrate_modes`` returns for that shot alone."""
        t_start = time.perf_counter()
        job = self.prepare_batch(shots, models, modes=modes, wavelets=wavelets,
                                 preprocess=preprocess)
        t_ready = time.perf_counter()
        out = self.run_batch(job)
        self.last_timing["prepare_s"] = t_ready - t_start
        self.last_timing["total_s"] = time.perf_counter() - t_start
        for res in out:
            for r in res.values():
                r.meta["perf"] = dict(self.last_timing)
        return out

    def prepare_batch(self, shots: list[ShotRecord], models: list[ElasticModel] | ElasticModel,
                      *, modes: tuple[str, ...] | None = None,
                      wavelets: list[np.ndarray] | None = None,
                      preprocess: Callable[[np.ndarray, float], np.ndarray] | None = None
                      ) -> "PreparedBatch":
        """Validate, align the records to (nt, dt) and upload them to HBM
        (ertm/imaging.py:205-245, :386-423); ``run_batch`` does the rest."""
        if isinstance(models, ElasticModel):
            models = [models] * len(shots)
        if len(models) != len(shots):
            raise ImagingError("one model per shot expected")
        if wavelets is not None and len(wavelets) != len(shots):
            raise ImagingError("one wavelet per shot expected")
        modes = tuple(str(m).lower() for m in (modes or (self.config.mode,)))
        engine = self.solver_config.engine
        for m in modes:
            if m not in WAVE_MODES:
                raise ImagingError(f"unsupported imaging mode {m!r}")
            if engine in ("acoustic", "viscoacoustic") and m not in _ACOUSTIC_MODES:
                raise ImagingError(f"acoustic/viscoacoustic engines image PP only, got {m!r}")
        swaves = {WAVE_MODES[m][0] for m in modes}
        if len(swaves) != 1 or len(modes) > 2:
            raise ImagingError("fused modes must share the source wave (at most two modes)")
        specs, records, waves = [], [], []
        for i, (shot, model) in enumerate(zip(shots, models)):
            try:
                shot.validate()
                model.validate()
            except ValueError as exc:
                raise ImagingError(f"invalid imaging input: {exc}") from exc
            spec = PropagationSpec.from_config(self.solver_config, model)
            try:
                if self.gpu_config.strict_cfl:
                    check_stability(model, spec)
                else:
                    check_hard_stability(model, spec)
            except Exception as exc:  # SolverError -> ImagingError
                raise ImagingError(str(exc)) from exc
            record = self._prepare_record(shot, spec, preprocess)
            if not record:
                raise ImagingError("empty shot record: nothing to back-propagate")
            specs.append(spec)
            records.append({nm: v if not isinstance(v, np.ndarray)
                            else dev.to_device(np.asarray(v, dtype=np.float32))
                            for nm, v in record.items()})
            waves.append(self._prepare_wavelet(shot, spec,
                                               None if wavelets is None else wavelets[i]))
        return PreparedBatch(shots=list(shots), models=list(models), specs=specs,
                             records=records, waves=waves, modes=modes, swave=swaves.pop(),
                             rwaves=tuple(WAVE_MODES[m][1] for m in modes))

    def run_batch(self, job: "PreparedBatch") -> list[dict[str, ImageResult]]:
        """Forward, replay, backward and imaging of prepared shots; records are
        already in HBM.  Returns one {mode: ImageResult} per shot."""
        return self.run_batches([job])[0]

    def run_batches(self, jobs: list["PreparedBatch"]) -> list[list[dict[str, ImageResult]]]:
        """``run_batch`` over several prepared batches, back to back: the host
        post-processing of one launch group runs in a background thread while
        the GPU advances the next, so a survey split into batches keeps the
        GPU busy.  Each result equals what ``run_batch`` returns for its job."""
        t_start = time.perf_counter()
        engine = self.solver_config.engine
        condition = str(self.config.condition).lower()
        timing = {"forward_s": 0.0, "replay_backward_s": 0.0, "postprocess_s": 0.0,
                  "cell_updates": 0.0, "shots": 0.0, "launch_groups": 0.0,
                  "kept_tail_steps": 0.0}
        cap = int(self.gpu_config.max_batch)
        outs: list[list[dict[str, ImageResult] | None]] = [[None] * len(j.shots) for j in jobs]
        units = []
        for jn, job in enumerate(jobs):
            groups: dict[tuple[int, int], list[int]] = {}
            for i, (model, spec) in enumerate(zip(job.models, job.specs)):
                groups.setdefault(self._batch_key(model, spec), []).append(i)
            for members in groups.values():
                units.extend((jn, members[c0:c0 + cap]) for c0 in range(0, len(members), cap))
        # device post-processing launches on this thread's stream: no overlap
        overlap = not self.gpu_config.device_postprocess
        pending = []
        with ThreadPoolExecutor(max_workers=1) as poster:
            for jn, idx in units:
                job = jobs[jn]
                # prepare_batch validated the models: the propagators built on
                # them skip re-validation and share each array's digest
                with dev.trusted_models([job.models[i] for i in idx]):
                    imgs, infos = self._run_shots(
                        [job.models[i] for i in idx], [job.specs[i] for i in idx], engine,
                        [job.shots[i] for i in idx], [job.records[i] for i in idx],
                        [job.waves[i] for i in idx], job.swave, job.rwaves, condition,
                        self.config.separation)
                for key in ("forward_s", "replay_backward_s", "cell_updates", "shots",
                            "kept_tail_steps"):
                    timing[key] += self.last_timing.get(key, 0.0)
                timing["launch_groups"] += 1.0
                # metadata now, so that no device buffer of this group outlives it
                metas = [[self._meta(job.specs[i], engine, job.shots[i], infos[j], condition, m)
                          for m in job.modes] for j, i in enumerate(idx)]
                del infos
                args = (job, idx, imgs, metas, condition, outs[jn])
                if overlap:
                    pending.append(poster.submit(self._post_group, *args))
                else:
                    timing["postprocess_s"] += self._post_group(*args)
            for fut in pending:
                timing["postprocess_s"] += fut.result()
        timing["run_s"] = time.perf_counter() - t_start
        self.last_timing = timing
        return outs

    def _post_group(self, job: "PreparedBatch", idx: list[int], imgs, metas, condition: str,
                    results: list) -> float:
        """Normalisation and post-processing of one launch group's images, in
        host threads (NumPy releases the GIL in these array operations; every
        image is computed exactly as in a serial loop).  Returns the time."""
        t_post = time.perf_counter()
        imgs = self._fetch_host(imgs)
        work = [(j, i, k) for j, i in enumerate(idx) for k in range(len(job.modes))]
        finish = lambda w: self._finish_image(  # noqa: E731
            imgs[w[0]], w[2], condition, job.models[w[1]], job.specs[w[1]])
        if self.gpu_config.device_postprocess:
            images = [finish(w) for w in work]  # kernels on this thread's stream
        else:
            with ThreadPoolExecutor(max_workers=_host_workers(len(work))) as pool:
                images = list(pool.map(finish, work))
        for (j, i, k), image in zip(work, images):
            m = job.modes[k]
            res = ImageResult(image=image, grid=job.models[i].grid, mode=m, meta=metas[j][k])
            res.validate()
            if results[i] is None:
                results[i] = {}
            results[i][m] = res
        return time.perf_counter() - t_post

    def _batch_key(self, model: ElasticModel, spec: PropagationSpec) -> tuple[int, int]:
        """(nt, checkpoint interval): shots with equal keys run in one launch.
        The interval follows the reference's rule on the fields-only state
        size (ertm/imaging.py:121-124)."""
        engine = self.solver_config.engine
        nf = {"acoustic": 3, "elastic": 5}.get(engine, 4 + int(self.q_config.n_mechanisms))
        pad = num.ghost_width(int(spec.order))
        nz, nx = model.shape
        per_state = nf * (nz + 2 * pad) * (nx + 2 * pad) * 4
        return int(spec.nt), checkpoint_interval(int(spec.nt), per_state,
                                                 float(self.config.checkpoint_mb))

    def image_condition(self, source_wavefield: np.ndarray, receiver_wavefield: np.ndarray):
        """Per-time-step integrand (ertm/imaging.py:351-368)."""
        s = np.asarray(source_wavefield)
        r = np.asarray(receiver_wavefield)
        if s.shape != r.shape:
            raise ImagingError(f"imaging-condition inputs differ in shape: {s.shape} vs {r.shape}")
        return zero_lag_condition(s, r).astype(np.float32)

    # ------------------------------------------------------------ engine
    def _routing(self, name: str) -> str:
        """Record name -> injected component (ertm/imaging.py:440-445)."""
        if name in ("primary", "p") and self.solver_config.engine == "acoustic":
            return "p"
        if name == "primary":
            return "vz"
        return name

    def _run_shots(self, models, specs, engine, shots, records, waves, swave, rwaves, condition,
                   separation):
        """Forward, replay and backward passes of several shots, each pass
        advancing every shot with one launch per time step.  The shots share
        nt and the checkpoint interval; grids may differ in size.  Each
        shot's result is the one it gets alone."""
        import torch

        timing = {}
        t0 = time.perf_counter()
        n = len(shots)
        nt = _common({int(sp.nt) for sp in specs}, "number of time steps")
        fwds, backs, tabs = [], [], []
        for model, spec, shot in zip(models, specs, shots):
            fwd = build_propagator(model, spec, engine=engine,
                                   q_options=self.q_options(compensation=False))
            fwd.set_source(shot.source)
            fwd.set_receivers(shot.receivers)
            fwds.append(fwd)
        kv = self.gpu_config.kernel_variant
        gsteps = int(self.gpu_config.graph_steps)
        ts = self.gpu_config.two_step
        side = side_stream("setup")

        def build_backs():
            # while the forward pass runs: the uploads of this set-up go
            # through a stream of their own instead of waiting behind it
            with torch.cuda.stream(side):
                for model, spec, shot, record in zip(models, specs, shots, records):
                    back = build_propagator(model, spec, engine=engine,
                                            q_options=self.q_options(compensation=True))
                    back.set_source(shot.source)
                    back.set_receivers(shot.receivers)
                    names = list(record)
                    recs = [record[nm] for nm in names]
                    tab = dev.RecvTables()
                    tab.build_injection(back._core, shot.receivers.x, shot.receivers.z,
                                        [self._routing(nm) for nm in names],
                                        back._k_relaxed_host())
                    tab.set_records(recs)
                    backs.append(back)
                    tabs.append(tab)

        # Replay slots, allocated before the forward pass: it fills the
        # first ones when it keeps the source fields of the last steps.
        tail = None
        t_keep = nt
        # excitation: from the first block start at which every wavelet is
        # constant, one replay run carried on to nt replays all later blocks
        # (and their checkpoints are not needed)
        t_chain = nt
        interval = _common({checkpoint_interval(nt, max(int(f.state_bytes()), 1),
                                                float(self.config.checkpoint_mb))
                            for f in fwds}, "checkpoint interval")
        if condition == "excitation":
            first = (constant_block_start(waves, nt, interval)
                     if self.gpu_config.keep_tail else None)
            if first is not None:
                t_chain = first
        else:
            cells = sum(m.shape[0] * m.shape[1] for m in models)
            arena_bytes = (4 * len(range(0, nt, interval))
                           * max(f._core.state_floats for f in fwds) * n)
            ring = int(self.gpu_config.image_ring)
            # folds beside the backward steps: a ring of two fold periods
            compensated = engine == "viscoacoustic" and bool(self.q_config.enabled)
            fold_beside = (self.gpu_config.overlap_flush and ring > 0 and bool(rwaves)
                           and gsteps == 0 and not compensated)
            ring_len = 2 * ring if fold_beside else ring
            # what the migration allocates beside the slots: its checkpoints,
            # the receiver rings, the backward states, and the float64
            # accumulators twice (the previous batch's are copied to the
            # host while this one runs)
            beside = arena_bytes + sum(
                4 * ring_len * len(rwaves) * m.shape[0] * f._core.ld
                + 8 * f._core.state_floats
                + 2 * 8 * (2 * len(rwaves) + 1) * m.shape[0] * m.shape[1]
                for m, f in zip(models, fwds))
            # Later batches of a shape take no more slots than the first: a
            # larger request than the block the caching allocator holds from
            # the previous batch makes it free its cache (a device
            # synchronisation) and map the slots anew.
            key = (nt, interval, tuple((m.shape[0], m.shape[1]) for m in models))
            room = min(self._slot_room(cells, reserve=beside),
                       self._rooms.get(key, 1 << 62))
            self._rooms[key] = room
            slots, kept = slot_plan(waves, nt, interval, room, self.gpu_config, cells)
            if kept is not None:
                t_keep = kept
            # the receiver-field rings before the replay slots: the step
            # writes a new ring slot every backward step
            rrings = [torch.empty((ring_len, len(rwaves), m.shape[0], f._core.ld),
                                  dtype=torch.float32, device="cuda")
                      if ring > 0 and rwaves else None for m, f in zip(models, fwds)]
            rbufs = [torch.empty((slots, m.shape[0], m.shape[1]), dtype=torch.float32,
                                 device="cuda") for m in models]
            fimgs = [dev.ImageBuffers(m.shape[0], m.shape[1], separation, swave, (), 0, 0,
                                      rbuf=rb, nslots=slots) for m, rb in zip(models, rbufs)]
            if kept is not None:
                for fi in fimgs:
                    fi.s.t0 = t_keep + 1
                tail = (t_keep, fimgs)

        setup = None
        if side.cuda_stream != torch.cuda.current_stream().cuda_stream:
            # the set-up has a stream of its own: run it beside the whole
            # forward pass, from before its launch
            setup = _HostTask(build_backs)
        try:
            stores = record_source_wavefields(
                fwds, waves, budget_mb=float(self.config.checkpoint_mb), variant=kv,
                graph_steps=gsteps, during=setup.join if setup else build_backs,
                to_last_checkpoint=True, two_step=ts, tail=tail,
                end_after=t_chain if t_chain < nt else None)
        except BaseException:
            if setup is not None:
                setup.wait()
            raise
        torch.cuda.current_stream().wait_stream(side)
        starts = []
        for shot, spec in zip(shots, specs):
            st = int(np.floor((float(shot.source.delay) + float(self.config.start_periods)
                               / max(float(shot.source.f0), 1e-6)) / spec.dt))
            starts.append(int(np.clip(st, 0, nt)))
        # accumulate only the illuminations the normalisation reads
        rec_illum = L.ILLUM_REC if self.config.receiver_normalization else 0
        illum = rec_illum | (L.ILLUM_SRC if condition == "normalized" and self.config.illumination
                             else 0)
        fcores = [f._core for f in fwds]
        bcores = [bk._core for bk in backs]
        bkw = batch_run_kwargs(backs)
        back_steps = 0
        store0 = stores[0]
        nblocks = len(store0.states)
        ws = [np.asarray(w, dtype=np.float32) for w in waves]
        arenas = [st.arena for st in stores]
        astride = _common({st.state_floats for st in stores}, "state size")
        fcur = _common({c.cur for c in fcores}, "state parity")
        bcur = _common({c.cur for c in bcores}, "state parity")
        replayed = 0
        torch.cuda.current_stream().synchronize()
        timing["forward_s"] = time.perf_counter() - t0
        timing["kept_tail_steps"] = float(nt - t_keep)
        t1 = time.perf_counter()
        if condition == "excitation":
            ibs = [dev.ImageBuffers(m.shape[0], m.shape[1], separation, "P", rwaves[:1], st,
                                    rec_illum, excitation=True) for m, st in zip(models, starts)]
            fb = dev.Batch(fcores, imgs=ibs, variant=kv, graph_steps=gsteps)
            for blk in range(nblocks):
                s0, s1 = store0.bounds(blk, nt)
                if s0 > t_chain:
                    break
                if s0 == t_chain:
                    s1 = nt  # the chained replay of every later block
                fcur = fb.run(1, s0, s1, fcur, wavelets=ws, op_after=L.IMG_EXC_PEAK,
                              restore=arenas, restore_stride=astride,
                              restore_slot=store0.states[blk][1])
                replayed += s1 - s0
            fb.close()
            bb = dev.Batch(bcores, recvs=tabs, imgs=ibs, variant=kv, graph_steps=gsteps)
            for blk in range(nblocks - 1, -1, -1):
                s0, s1 = store0.bounds(blk, nt)
                bcur = bb.run(2, s0, s1, bcur, op_after=L.IMG_EXC_APPLY,
                              step_index0=back_steps, **bkw)
                back_steps += s1 - s0
            bb.close()
            accs = [(ib.num, ib.peak_amp ** 2, ib.rill) for ib in ibs]
        else:
            ibs = [dev.ImageBuffers(m.shape[0], m.shape[1], separation, swave, rwaves, st, illum,
                                    rbuf=rb, nslots=slots, ring=ring_len, ring_ld=bc.ld, rring=rr)
                   for m, st, rb, bc, rr in zip(models, starts, rbufs, bcores, rrings)]
            fb = dev.Batch(fcores, imgs=fimgs, variant=kv, graph_steps=gsteps, two_step=ts)
            bb = dev.Batch(bcores, recvs=tabs, imgs=ibs, variant=kv, graph_steps=gsteps)
            if fold_beside:
                bb.set_ring_flush(ring)
            subs = None
            if slots < store0.interval:
                nsub = int(math.ceil(store0.interval / slots))
                subs = [torch.empty((nsub, astride), dtype=torch.float32, device="cuda")
                        for _c in fcores]
            if t_keep < nt:
                # the kept tail: its source fields are in the slots already
                for ib in ibs:
                    ib.s.t0 = t_keep
                bb.update_images(ibs)
                bcur = bb.run(2, t_keep, nt, bcur, op_after=L.IMG_CORRELATE,
                              step_index0=back_steps, **bkw)
                back_steps += nt - t_keep
            for blk in range(nblocks - 1, -1, -1):
                s0, s1 = store0.bounds(blk, nt)
                s1 = min(s1, t_keep)
                if s1 <= s0:
                    continue
                slot = store0.states[blk][1]
                segs = replay_segments(s0, s1, slots)
                if len(segs) > 1:
                    # sub-checkpoints: replay state just before w[a] of each later segment
                    fcur = fb.run(1, s0, segs[-1][0], fcur, wavelets=ws, ck_interval=slots,
                                  ck_origin=sub_checkpoint_origin(segs), save=subs,
                                  save_stride=astride,
                                  restore=arenas, restore_stride=astride, restore_slot=slot)
                    replayed += segs[-1][0] - s0
                for k in range(len(segs) - 1, -1, -1):
                    a, e = segs[k]
                    for fi, ib in zip(fimgs, ibs):
                        fi.s.t0 = a
                        ib.s.t0 = a
                    fb.update_images(fimgs)
                    bb.update_images(ibs)
                    if k == 0:
                        src = dict(restore=arenas, restore_stride=astride, restore_slot=slot)
                    else:
                        src = dict(restore=subs, restore_stride=astride, restore_slot=k - 1)
                    fcur = fb.run(1, a, e, fcur, wavelets=ws, op_after=L.IMG_EXTRACT, **src)
                    replayed += e - a
                    bcur = bb.run(2, a, e, bcur, op_after=L.IMG_CORRELATE,
                                  step_index0=back_steps, **bkw)
                    back_steps += e - a
            fb.close()
            bb.close()
            accs = [(ib.num, ib.ill, ib.rill) for ib in ibs]
        batch_after_run(backs, bkw)
        torch.cuda.current_stream().synchronize()
        timing["replay_backward_s"] = time.perf_counter() - t1
        for c in fcores:
            c.cur = fcur
        for c in bcores:
            c.cur = bcur
        use_ill = condition == "normalized" and bool(self.config.illumination)
        use_rill = bool(self.config.receiver_normalization)
        all_outs, infos = [], []
        for j in range(n):
            num_t, ill, rill = accs[j]
            backs[j].step_index = back_steps
            backs[j].pml.touched = True
            if self.gpu_config.device_postprocess:
                # accumulators stay in HBM; _finish_image runs ertm_postprocess
                all_outs.append(("device", list(num_t), ill, list(rill)))
            else:
                # raw float64 accumulators, copied to the host by _fetch_host
                # and normalised by _finish_image
                all_outs.append(("fetch", list(num_t), ill if use_ill else None,
                                 list(rill) if use_rill else None,
                                 torch.cuda.current_stream()))
            nz, nx = models[j].shape
            infos.append(dict(store=stores[j], start_step=starts[j], prop=fwds[j], back=backs[j],
                              replayed_steps=replayed, nt=nt,
                              cell_updates=(stores[j].forward_steps + replayed + back_steps)
                              * nz * nx))
        timing["cell_updates"] = float(sum(i["cell_updates"] for i in infos))
        timing["shots"] = float(n)
        self.last_timing = timing
        return all_outs, infos

    def _slot_room(self, cells: int, reserve: int = 0) -> int:
        """Replay slots (steps of ``cells`` float32 fields, summed over the
        shots advanced together) that fit the replay budget and
        ``_REPLAY_FREE_SHARE`` of the free HBM left after ``reserve`` bytes
        still to be allocated.  Free HBM counts the blocks torch's caching
        allocator holds unused (an earlier batch's slots), which these
        tensors reuse.  Over it a block is replayed in segments."""
        import torch

        if self.gpu_config.replay_slots > 0:
            return max(1, int(self.gpu_config.replay_slots))
        free, _total = torch.cuda.mem_get_info()
        free += max(0, torch.cuda.memory_reserved() - torch.cuda.memory_allocated())
        budget = min(float(self.gpu_config.replay_budget_gb) * 1024 ** 3,
                     _REPLAY_FREE_SHARE * max(0, free - int(reserve)))
        return max(1, int(budget // (4 * cells)))

    # ------------------------------------------------------------ helpers
    def _validate_engine_mode(self) -> None:
        if (self.solver_config.engine in ("acoustic", "viscoacoustic")
                and self.config.mode not in _ACOUSTIC_MODES):
            raise ImagingError(
                f"acoustic/viscoacoustic engines image PP only, got imaging.mode="
                f"{self.config.mode!r}; use solver.engine='elastic' for converted waves")

    def _resolve_mode(self) -> tuple[str, str]:
        mode = str(self.config.mode).lower()
        if mode not in WAVE_MODES:
            raise ImagingError(f"unsupported imaging mode {mode!r}")
        return WAVE_MODES[mode]

    def _prepare_wavelet(self, shot: ShotRecord, spec: PropagationSpec,
                         wavelet: np.ndarray | None) -> np.ndarray:
        if wavelet is not None:
            return np.asarray(wavelet, dtype=np.float32)
        return source_wavelet(shot.source, spec.dt, spec.nt)

    def _prepare_record(self, shot: ShotRecord, spec: PropagationSpec,
                        preprocess: Callable[[np.ndarray, float], np.ndarray] | None
                        ) -> dict[str, np.ndarray]:
        """Align the record to (nt, dt): sinc resampling, optional preprocess,
        truncate / zero-pad (ertm/imaging.py:396-423).  Host float64 arrays,
        or, for float32 records of two or more traces resampled without a
        preprocess callable (``GpuConfig.device_resample``), (nt, nrec)
        float32 device tensors resampled and fitted on the GPU."""
        comps = dict(shot.components) or {"primary": shot.data}
        out: dict[str, np.ndarray] = {}
        for name, arr in comps.items():
            resample = abs(float(shot.dt) - spec.dt) > 1e-12 and np.shape(arr)[0] > 1
            if (resample and preprocess is None and self.gpu_config.device_resample
                    and getattr(arr, "dtype", None) == np.float32 and np.ndim(arr) == 2
                    and np.shape(arr)[1] >= 2 and np.shape(arr)[0] >= 3):
                # on the device: a (nt, nrec) float32 tensor, already fitted
                from .preprocess import resample_record_on_device

                out[name] = resample_record_on_device(arr, float(shot.dt), spec.dt,
                                                      int(spec.nt))
                continue
            data = np.asarray(arr, dtype=np.float64)
            if resample:
                data = num.resample_sinc(data, float(shot.dt), spec.dt)
            if preprocess is not None:
                data = np.asarray(preprocess(data, spec.dt), dtype=np.float64)
            out[name] = self._fit_length(data, int(spec.nt))
        return out

    @staticmethod
    def _fit_length(data: np.ndarray, nt: int) -> np.ndarray:
        if data.shape[0] >= nt:
            return data[:nt]
        padded = np.zeros((nt, data.shape[1]), dtype=data.dtype)
        padded[: data.shape[0]] = data
        return padded

    @staticmethod
    def _fetch_host(imgs: list) -> list:
        """Host copies of a launch group's float64 accumulators.

        For groups launched on the default stream the copies run on a side
        stream: in ``run_batches`` the next group's launches are already
        queued on the launching stream, and the copies must not wait behind
        them (``side_stream`` says when the launching stream is used
        instead).  ``_run_shots`` synchronised its stream before returning,
        so the accumulators are complete; the copy stream is synchronised
        before the device tensors are released."""
        if not any(raw[0] == "fetch" for raw in imgs):
            return imgs
        import torch

        launch = next(raw[4] for raw in imgs if raw[0] == "fetch")
        side = side_stream("fetch", launch)
        out = []
        with torch.cuda.stream(side):
            for raw in imgs:
                if raw[0] != "fetch":
                    out.append(raw)
                    continue
                _tag, nums, ill, rill, _launch = raw
                out.append(([x.cpu().numpy() for x in nums],
                            None if ill is None else ill.cpu().numpy(),
                            None if rill is None else [r.cpu().numpy() for r in rill]))
        side.synchronize()
        return out

    def _finish_image(self, raw, k: int, condition: str, model: ElasticModel,
                      spec: PropagationSpec) -> np.ndarray:
        """Normalise receiver wave k's accumulator and post-process it
        (ertm/imaging.py:326-349)."""
        if raw[0] == "device":
            _tag, nums, ill, rill = raw
            out = dev.postprocess(
                nums[k], ill, rill[k], dx=model.grid.dx, dz=model.grid.dz,
                eps=float(self.config.eps),
                div_ill=condition == "normalized" and bool(self.config.illumination),
                div_rill=bool(self.config.receiver_normalization),
                laplacian=bool(self.config.laplacian),
                lowpass=self.config.lowpass_fraction < 1.0,
                fraction=float(self.config.lowpass_fraction), npml=int(spec.npml),
                mask_top=spec.boundary == "cpml")
            return out.cpu().numpy()
        nums, ill, rill = raw
        eps = float(self.config.eps)
        image = nums[k]
        if condition == "normalized" and self.config.illumination:
            image = image / (ill + eps)
        if self.config.receiver_normalization:
            image = image / (rill[k] + eps)
        return np.asarray(self._postprocess(image, model, spec), dtype=np.float32)

    def _postprocess(self, image: np.ndarray, model: ElasticModel, spec: PropagationSpec):
        """Laplacian, wavenumber low-pass, PML mask (ertm/imaging.py:476-490)."""
        data = np.asarray(image, dtype=np.float64)
        if self.config.laplacian:
            data = num.laplacian(data, model.grid.dx, model.grid.dz)
        if self.config.lowpass_fraction < 1.0:
            data = num.lowpass_wavenumber(data, model.grid.dx, model.grid.dz,
                                          self.config.lowpass_fraction)
        return data * self._pml_mask(model, spec)

    @staticmethod
    def _pml_mask(model: ElasticModel, spec: PropagationSpec) -> np.ndarray:
        """Zero npml physical cells at the absorbing sides (ertm/imaging.py:492-502)."""
        nz, nx = model.shape
        n = int(spec.npml)
        mask = np.ones((nz, nx), dtype=np.float64)
        mask[:, :n] = 0.0
        mask[:, -n:] = 0.0
        mask[-n:, :] = 0.0
        if spec.boundary == "cpml":
            mask[:n, :] = 0.0
        return mask

    def _meta(self, spec, engine, shot, info, condition, mode) -> dict[str, Any]:
        q = self.q_config
        q_meta: dict[str, Any] = {"model": str(q.model), "enabled": bool(q.enabled),
                                  "alpha_c": float(q.alpha_c), "f_ref": float(q.f_ref)}
        if engine == "viscoacoustic":
            fs = getattr(info["prop"], "quality_summary", None)
            if isinstance(fs, dict):
                q_meta.update(fs)
            bs = getattr(info["back"], "quality_summary", None)
            if isinstance(bs, dict):
                for key in ("compensation", "alpha_c", "alpha_mechanisms", "f_peak_mechanisms_hz",
                            "f_max", "gain_budget_db"):
                    if key in bs:
                        q_meta[key] = bs[key]
                q_meta["clipped_steps"] = int(bs.get("clipped", 0))
                q_meta["lowpass_steps"] = int(bs.get("lowpass_steps", 0))
        store: CheckpointStore = info["store"]
        return {
            "placeholder": False,
            "method": "reverse-time migration (two-way wave equation)",
            "imager": self.name,
            "engine": engine,
            "device": "hip-gfx950",
            "condition": condition,
            "separation": self.config.separation,
            "modes": [mode],
            "illumination": bool(self.config.illumination),
            "receiver_normalization": bool(self.config.receiver_normalization),
            "laplacian": bool(self.config.laplacian),
            "lowpass_fraction": float(self.config.lowpass_fraction),
            "checkpoints": store.summary(),
            "start_step": int(info["start_step"]),
            "start_periods": float(self.config.start_periods),
            "q": q_meta,
            "nt": int(spec.nt),
            "dt": float(spec.dt),
            "order": int(spec.order),
            "npml": int(spec.npml),
            "source_x": float(shot.source.x),
            "source_z": float(shot.source.z),
            "n_receivers": int(shot.receivers.nrec),
            "replayed_steps": int(info["replayed_steps"]),
            "replay_overhead": float(info["replayed_steps"]) / max(int(spec.nt), 1),
        }


def build_imager(config: ImagingConfig, solver_config: SolverConfig | None = None,
                 q_config: QConfig | None = None, gpu_config: GpuConfig | None = None) -> RTMImager:
    """ertm/imaging.py:566-572."""
    return RTMImager(config, solver_config, q_config, gpu_config)
