"""Replay segmentation of a checkpoint block (host logic, no GPU).

A block longer than its replay slots is replayed segment by segment, last
segment first, from sub-checkpoints saved by one pass from the block start.
These tests model that pass with the library's save rule (a state is saved
after step t when t >= origin and (t - origin) % interval == 0, into slot
(t - origin) / interval) and check that every later segment finds the
state just before its first step in the slot the replay restores from.
"""
from __future__ import annotations

import pytest

from ertm.imaging import replay_segments, sub_checkpoint_origin


def _saved_after(t_begin: int, t_end: int, origin: int, interval: int) -> dict[int, int]:
    """slot -> step after which the sub-checkpoint pass saved it."""
    out = {}
    for t in range(t_begin, t_end):
        if t >= origin and (t - origin) % interval == 0:
            out[(t - origin) // interval] = t
    return out


@pytest.mark.parametrize("s0,s1,slots", [(0, 852, 419), (852, 1704, 419), (0, 838, 419),
                                         (0, 200, 97), (400, 411, 5), (0, 7, 7), (0, 5, 9),
                                         (10, 11, 1), (0, 1000, 1), (3, 64, 8)])
def test_segments_cover_the_block_short_one_first(s0, s1, slots):
    segs = replay_segments(s0, s1, slots)
    assert segs[0][0] == s0 and segs[-1][1] == s1
    assert all(a < e for a, e in segs)
    assert all(segs[k][1] == segs[k + 1][0] for k in range(len(segs) - 1))
    assert all(e - a == slots for a, e in segs[1:])
    assert 1 <= segs[0][1] - segs[0][0] <= slots
    n = s1 - s0
    assert len(segs) == -(-n // slots)
    # the sub-checkpoint pass runs from s0 to the last segment's start
    assert segs[-1][0] - s0 == max(0, n - slots)
    if len(segs) > 1:
        origin = sub_checkpoint_origin(segs)
        saved = _saved_after(s0, segs[-1][0], origin, slots)
        # segment k >= 1 restores slot k - 1: the state after step a_k - 1
        for k in range(1, len(segs)):
            assert saved[k - 1] == segs[k][0] - 1
        assert len(saved) == len(segs) - 1


def test_short_segment_first_halves_the_overhead_of_a_long_block():
    """852-step blocks with 419 slots: the sub-checkpoint pass runs 433 steps
    instead of 838 with the remainder at the end."""
    segs = replay_segments(0, 852, 419)
    assert segs == [(0, 14), (14, 433), (433, 852)]
    assert segs[-1][0] == 433
