"""The allocation-free post-processing entry point (ertm_postprocess_noalloc)
against ertm_postprocess with a caller-owned work buffer, through the C-ABI.

Both run the same kernels, so the images must be bit-identical.  The
allocation-free calls come two in a row, the second on a smaller image, so a
stale scratch left by the first call would show in the second.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(fn_name, num, ill, rill, opts, work):
    import torch

    from ertm import _lib as L

    nz, nx = num.shape
    out = torch.full((nz, nx), float("nan"), dtype=torch.float32, device="cuda")
    p = L.ErtmPost()
    p.nz, p.nx = nz, nx
    p.div_ill, p.div_rill = opts["div_ill"], opts["div_rill"]
    p.laplacian, p.lowpass = opts["laplacian"], 0
    p.npml, p.mask_top = opts["npml"], opts["mask_top"]
    p.dx, p.dz, p.eps, p.fraction = opts["dx"], opts["dz"], opts["eps"], 1.0
    p.num = L.ptr(num)
    p.ill = L.ptr(ill) if opts["div_ill"] else None
    p.rill = L.ptr(rill) if opts["div_rill"] else None
    p.work = L.ptr(work)
    p.out = L.ptr(out)
    L.check(getattr(L.lib(), fn_name)(C.byref(p), L.stream_handle()), fn_name)
    return out


def _inputs(rng, nz, nx):
    import torch

    num = torch.from_numpy(rng.standard_normal((nz, nx))).cuda()
    ill = torch.from_numpy(rng.uniform(0.0, 3.0, (nz, nx))).cuda()
    rill = torch.from_numpy(rng.uniform(0.0, 2.0, (nz, nx))).cuda()
    return num, ill, rill


OPTS = [
    dict(div_ill=1, div_rill=0, laplacian=1, npml=20, mask_top=0, dx=5.0, dz=5.0, eps=1e-8),
    dict(div_ill=0, div_rill=1, laplacian=0, npml=3, mask_top=1, dx=4.0, dz=2.5, eps=1e-3),
    dict(div_ill=1, div_rill=1, laplacian=1, npml=0, mask_top=1, dx=10.0, dz=7.0, eps=0.0),
]


@pytest.mark.parametrize("opts", OPTS, ids=["ill_lap", "rill_mask", "both_lap"])
def test_noalloc_postprocess_equals_postprocess(opts):
    import torch

    from ertm import _lib as L

    rng = np.random.default_rng(11)
    # the second call is smaller than the first; the first fills the scratch
    shapes = [(1024, L.NOALLOC_MAX_CELLS // 1024), (37, 53)]
    print("noalloc postprocess: shapes", shapes, "opts", opts, flush=True)
    cases = [_inputs(rng, nz, nx) for nz, nx in shapes]
    refs = []
    for num, ill, rill in cases:
        work = torch.empty(4 * num.numel(), dtype=torch.float64, device="cuda")
        refs.append(_run("ertm_postprocess", num, ill, rill, opts, work))
    torch.cuda.synchronize()
    got = [_run("ertm_postprocess_noalloc", num, ill, rill, opts, None)
           for num, ill, rill in cases]
    torch.cuda.synchronize()
    for (nz, nx), r, g in zip(shapes, refs, got):
        assert torch.isfinite(r).all()
        assert torch.equal(r, g), f"{nz}x{nx}: max diff {float((r - g).abs().max())}"


def test_noalloc_postprocess_refuses_what_it_cannot_do():
    import torch

    from ertm import _lib as L

    num, ill, rill = _inputs(np.random.default_rng(2), 8, 8)
    out = torch.empty((8, 8), dtype=torch.float32, device="cuda")
    p = L.ErtmPost()
    p.nz, p.nx, p.lowpass, p.num, p.out = 8, 8, 1, L.ptr(num), L.ptr(out)
    assert L.lib().ertm_postprocess_noalloc(C.byref(p), L.stream_handle()) == 1
    assert b"low-pass" in L.lib().ertm_last_error()
    p.lowpass, p.nz, p.nx = 0, 4096, 1024
    assert L.lib().ertm_postprocess_noalloc(C.byref(p), L.stream_handle()) == 1
    assert b"static scratch" in L.lib().ertm_last_error()
