"""Stencil-kernel variant selection and autotuning.

Variants (see ertm_batch_set_variant): 0 = two launches per step (any
engine/order); 1/2/3/4 = fused single launch with 256/128/64/192 threads per
block, each block streaming a chunk of ``zc`` rows.  Every variant gives bit-identical
results, so the choice is purely a performance decision: the heuristic picks
the widest block that still yields enough blocks to keep all 256 CUs busy,
and ``autotune`` times the candidates on the real shapes.
"""

from __future__ import annotations

import json
import math
import os
import threading

FUSED_THREADS = {1: 256, 2: 128, 3: 64, 4: 192}


def threads(variant: int) -> int:
    return FUSED_THREADS[variant]
_cache: dict = {}
_lock = threading.Lock()
CACHE_FILE = os.environ.get("ERTM_TUNE_CACHE", "")


def fused_supported(engine: str, order: int, nmech: int = 0) -> bool:
    """Single-launch kernels: order 8, elastic, acoustic, or viscoacoustic
    with 1..4 relaxation mechanisms."""
    if int(order) != 8:
        return False
    return engine in ("elastic", "acoustic") or (engine == "viscoacoustic" and 1 <= nmech <= 4)


def blocks(variant: int, zc: int, nzp: int, nxp: int, nbatch: int) -> int:
    tx = threads(variant) - 16
    return math.ceil(nxp / tx) * math.ceil(nzp / zc) * nbatch


def heuristic(engine: str, order: int, nzp: int, nxp: int, nbatch: int = 1,
              n_cu: int = 256, nmech: int = 0) -> tuple[int, int]:
    """192-lane fused blocks with 48-row chunks when they give 16 blocks per
    CU, else 128-lane blocks with the longest chunk (up to 48 rows) that
    does: enough row streams keep loads in flight, and alternating chunk
    directions let L2 absorb part of the halo re-reads (measured on 64 shots
    of 610 x 1210 sharing one medium, 128 lanes: 48-row chunks 646 us per
    step, 24-row 731 us, 96-row 734 us); batches too small to fill the chip
    use the two-launch kernels, whose 2-D tiles need no halo recompute."""
    if not fused_supported(engine, order, nmech):
        return (0, 64)
    # 192-lane blocks (176 owned columns) re-read fewer halo columns than
    # 128-lane ones: 670 us against 714 us per step on 64 shots of 610 x 1210
    if blocks(4, 48, nzp, nxp, nbatch) >= 16 * n_cu:
        return (4, 48)
    for zc in (48, 40, 32, 24, 19, 16, 10):
        if blocks(2, zc, nzp, nxp, nbatch) >= 16 * n_cu:
            return (2, zc)
    if blocks(2, 10, nzp, nxp, nbatch) >= 4 * n_cu:
        return (2, 10)
    return (0, 64)


def two_step_lookup(engine: str, order: int, nzp: int, nxp: int, nbatch: int = 1
                    ) -> tuple[int, int]:
    """Launch shape of the two-steps-per-launch elastic kernel for a batch
    (forward and replay passes), (0, 0) = one step per launch."""
    if engine != "elastic" or int(order) != 8:
        return (0, 0)
    return TWO_STEP_DEFAULT


# (variant, zc) of the two-step kernel when the batch allows it
TWO_STEP_DEFAULT = (0, 0)


def key(engine, order, nzp, nxp, nbatch) -> str:
    return f"{engine}:{order}:{nzp}x{nxp}:b{nbatch}"


def lookup(engine: str, order: int, nzp: int, nxp: int, nbatch: int = 1,
           nmech: int = 0) -> tuple[int, int]:
    k = key(engine, order, nzp, nxp, nbatch)
    with _lock:
        if not _cache and CACHE_FILE and os.path.exists(CACHE_FILE):
            try:
                _cache.update({kk: tuple(v) for kk, v in json.load(open(CACHE_FILE)).items()})
            except (OSError, ValueError):
                pass
        if k in _cache:
            return _cache[k]
    return heuristic(engine, order, nzp, nxp, nbatch, nmech=nmech)


def record(engine, order, nzp, nxp, nbatch, choice) -> None:
    with _lock:
        _cache[key(engine, order, nzp, nxp, nbatch)] = tuple(choice)
        if CACHE_FILE:
            try:
                json.dump({k: list(v) for k, v in _cache.items()}, open(CACHE_FILE, "w"))
            except OSError:
                pass


def candidates(engine: str, order: int, nmech: int = 0) -> list[tuple[int, int]]:
    if not fused_supported(engine, order, nmech):
        return [(0, 64)]
    return [(0, 64)] + [(v, zc) for v in (1, 2, 3, 4)
                        for zc in (16, 24, 32, 48, 64, 96, 128, 160, 208, 320)]


def autotune(batch, steps: int = 10, cands=None) -> tuple[tuple[int, int], dict]:
    """Time ``steps`` steps of each candidate on ``batch`` (its state is
    advanced; results do not depend on the variant).  Returns the fastest
    (variant, zc) and all timings in ms/step; records the choice."""
    import torch

    c0 = batch.cores[0]
    cands = cands or candidates(c0.engine, c0.order, c0.nmech)
    times = {}
    amps = [0.0] * batch.n
    for cand in cands:
        batch.set_variant(*cand)
        for _ in range(2):
            batch.step(c0.cur, amps)
            for c in batch.cores:
                c.cur ^= 1
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            batch.step(c0.cur, amps)
            for c in batch.cores:
                c.cur ^= 1
        e1.record()
        torch.cuda.synchronize()
        times[cand] = e0.elapsed_time(e1) / steps
    best = min(times, key=times.get)
    batch.set_variant(*best)
    record(c0.engine, c0.order, c0.nzp, c0.nxp, batch.n, best)
    return best, times


def round_filling(nzp: int, nxp: int, nbatch: int, n_cu: int = 256, max_rounds: int = 4,
                  waves_per_cu: int = 20) -> list[tuple[int, int]]:
    """Launch shapes whose blocks fill whole rounds of the resident blocks.

    A launch of equal blocks takes about ceil(blocks / resident) block
    durations: a few blocks past a round boundary cost a whole extra round.
    For 256-, 128- and 192-lane blocks (resident: 20 waves per CU) and r = 1
    .. max_rounds rounds, the three shortest chunk heights (16 to 128 rows)
    whose block count stays within r rounds.  Measured on the 2000 x 6000
    shot (one launch of 0.96-1.0 rounds per step): 128 lanes, 42-row chunks
    (1.01 rounds) 249 us per step against 216 us with 43-row chunks (0.99);
    256 lanes, 43-row chunks 198 us against 205 us for the shape the
    candidates above chose."""
    out = []
    for v in (1, 2, 4):
        per_round = n_cu * (waves_per_cu // (threads(v) // 64))
        for r in range(1, max_rounds + 1):
            fits = [zc for zc in range(16, 129) if blocks(v, zc, nzp, nxp, nbatch) <= r * per_round]
            heights = sorted({math.ceil(nzp / math.ceil(nzp / zc)) for zc in fits})[:3]
            out += [(v, zc) for zc in heights if (v, zc) not in out]
    return out


# A run of at least this many cell-updates (cells x steps x shots) is worth a
# short tuning pass on its first batch: about 60 ms of steps against seconds.
TUNE_MIN_WORK = 2.0e10


def maybe_autotune(batch, work: float, n_cu: int = 256) -> tuple[int, int] | None:
    """Tune a fused batch's launch shape once per (engine, order, grid, batch
    size), when the run ahead is long enough to amortise it: 192- and 128-lane
    blocks with 16/24/32/48/64-row chunks, 6 steps each.  The batch's states are
    saved and restored around the timed steps, so the run is unaffected.  The
    fastest shape differs between MI355X boards (e.g. (2, 48) 646 us against
    (4, 48) 658 us on one, 714 us against 670 us on another)."""
    import torch

    c0 = batch.cores[0]
    if work < TUNE_MIN_WORK or not fused_supported(c0.engine, c0.order, c0.nmech):
        return None
    nzp, nxp = max(c.nzp for c in batch.cores), max(c.nxp for c in batch.cores)
    k = key(c0.engine, c0.order, nzp, nxp, batch.n)
    with _lock:
        if k in _cache:
            return None
    # 24- and 16-row chunks matter for single large grids, whose 48-row
    # chunks make about one round of blocks: the 2000 x 6000 shot migrates
    # 11 % faster with (4, 24) than with (4, 48)
    cands = [(v, zc) for v in (4, 2) for zc in (48, 32, 64, 24, 16)
             if blocks(v, zc, nzp, nxp, batch.n) >= 4 * n_cu]
    cands += [c for c in round_filling(nzp, nxp, batch.n, n_cu) if c not in cands]
    if len(cands) < 2:
        return None
    saved = [(c.state.clone(), c.cur) for c in batch.cores]
    best, _times = autotune(batch, steps=6, cands=cands)
    for c, (st, cur) in zip(batch.cores, saved):
        c.state.copy_(st)
        c.cur = cur
    torch.cuda.synchronize()
    del saved
    return best

