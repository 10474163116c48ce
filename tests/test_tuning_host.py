"""Launch-shape candidates of the tuner (host logic only)."""

from __future__ import annotations

import math

from ertm import tuning


def _per_round(v, n_cu=256):
    return n_cu * (20 // (tuning.threads(v) // 64))


def test_round_filling_shapes_fit_whole_rounds():
    """Every round-filling shape's blocks fit within some whole number of
    resident rounds, and one row less per chunk (more chunks) no longer fits
    that number for the shortest height listed per variant and round count."""
    for nzp, nxp, nb in ((2010, 6010, 1), (610, 1210, 64), (2050, 4050, 1), (300, 500, 8)):
        cands = tuning.round_filling(nzp, nxp, nb)
        assert cands, (nzp, nxp, nb)
        for v, zc in cands:
            assert v in (1, 2, 4) and 16 <= zc <= 128
            blocks = tuning.blocks(v, zc, nzp, nxp, nb)
            r = math.ceil(blocks / _per_round(v))
            assert 1 <= r <= 4, (v, zc, blocks)
        for v in (1, 2, 4):
            for r in range(1, 5):
                fit = [zc for zc in range(16, 129)
                       if tuning.blocks(v, zc, nzp, nxp, nb) <= r * _per_round(v)]
                if fit:
                    h = math.ceil(nzp / math.ceil(nzp / min(fit)))
                    assert (v, h) in cands, (nzp, nxp, nb, v, r, h)


def test_round_filling_includes_the_measured_large_shot_optimum():
    # 2000 x 6000 shot padded to 2010 x 6010: 256 lanes x 43 rows = 1222 blocks
    cands = tuning.round_filling(2010, 6010, 1)
    assert (1, 43) in cands and (2, 43) in cands
    assert tuning.blocks(1, 43, 2010, 6010, 1) == 1222
