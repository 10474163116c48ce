"""Host-side functions of the product on the CPU: source wavelets, sinc
resampling, the CFL guard and its bypass, and the float64 refusal.

The wavelet, resampling and CFL values are compared bit for bit with the
golden vectors written from the reference (tests/golden/make_golden.py).
"""

from __future__ import annotations

import numpy as np
import pytest

from golden_cases import load


def _model(vp=3000.0, nx=121, nz=91, dx=10.0, dz=10.0):
    """The reference tests' make_model (tests/test_rtm.py:64-82)."""
    from ertm.types import ElasticModel, GridSpec

    grid = GridSpec(nx=nx, nz=nz, dx=dx, dz=dz)
    full = lambda v: np.full((nz, nx), v, dtype=np.float32)  # noqa: E731
    return ElasticModel(grid=grid, vp=full(vp), vs=full(1700.0), rho=full(2400.0),
                        qp=full(100.0), qs=full(50.0))


def test_source_wavelets_match_reference():
    from ertm import numerics as num
    from ertm.propagation import source_wavelet
    from ertm.types import SourceSpec

    d = load("numerics")
    t = np.arange(400, dtype=np.float64) * 2.5e-4
    for k in range(3):
        f0, t0 = d[f"gauss{k}_params"]
        assert np.array_equal(num.gaussian_pulse(t, f0, t0), d[f"gauss{k}"])
    for k in range(3):
        f0, delay, amp, dt, nt = d[f"wavelet{k}_params"]
        src = SourceSpec(x=0.0, z=0.0, f0=f0, delay=delay, amplitude=amp,
                         kind=str(d[f"wavelet{k}_kind"]))
        assert np.array_equal(source_wavelet(src, dt, int(nt)), d[f"wavelet{k}"])


def test_resample_sinc_matches_reference():
    """ertm/numerics.py:337-392, the path _prepare_record takes when a shot's
    step differs from the solver's (ertm/imaging.py:410-411): bit for bit,
    several traces and one (1-D)."""
    from ertm import numerics as num

    d = load("numerics")
    for k in range(5):
        dt_in, dt_out = d[f"sinc{k}_params"]
        assert np.array_equal(num.resample_sinc(d[f"sinc{k}_in"], dt_in, dt_out),
                              d[f"sinc{k}_out"]), k
        assert np.array_equal(num.resample_sinc(d[f"sinc{k}_in1"], dt_in, dt_out),
                              d[f"sinc{k}_out1"]), k


def test_resample_sinc_preserves_a_sine():
    """tests/test_rtm.py:159-167 on the product's resampler."""
    from ertm import numerics as num

    dt_in, dt_out = 4.0e-4, 1.0e-4
    signal = np.sin(2 * np.pi * 35.0 * np.arange(500) * dt_in)
    out = num.resample_sinc(signal, dt_in, dt_out)
    analytic = np.sin(2 * np.pi * 35.0 * np.arange(out.size) * dt_out)
    assert np.max(np.abs(out[:-8] - analytic[:-8])) < 2e-2


def test_cfl_limits_match_reference():
    from ertm import numerics as num

    for vmax, dx, dz, order, safety, ref in load("numerics")["cfl_table"]:
        assert num.cfl_dt(vmax, dx, dz, int(order), safety=safety) == ref


def test_cfl_guard_rejects_unstable_step():
    """tests/test_rtm.py:169-175: dt = 2 ms at 4000 m/s on a 10 m grid is
    refused, 0.4 ms passes."""
    from ertm.exceptions import SolverError
    from ertm.propagation import PropagationSpec, check_stability

    model = _model(vp=4000.0)
    with pytest.raises(SolverError):
        check_stability(model, PropagationSpec(dt=2.0e-3, nt=10, order=8, npml=8))
    check_stability(model, PropagationSpec(dt=4.0e-4, nt=10, order=8, npml=8))


def test_guard_bypass_keeps_the_hard_stability_limit():
    """GpuConfig(strict_cfl=False) replaces the reference's guard (0.55 of the
    CFL bound) by the bound itself (check_hard_stability).  The plain-RTM
    workload's 0.5 ms at 3500 m/s on a 5 m grid is 1.158 x the guard and
    0.64 x the bound: refused by the guard, accepted by the bypass; a step
    above the bound is refused by both."""
    from ertm import numerics as num
    from ertm.config import GpuConfig, SolverConfig
    from ertm.exceptions import SolverError
    from ertm.propagation import PropagationSpec, check_hard_stability, check_stability
    from ertm.solver import build_solver

    model = _model(vp=3500.0, nx=300, nz=200, dx=5.0, dz=5.0)
    bound = num.cfl_dt(3500.0, 5.0, 5.0, 8, safety=1.0)
    assert 5e-4 / num.cfl_dt(3500.0, 5.0, 5.0, 8, safety=0.55) == pytest.approx(1.158, abs=1e-3)
    assert 5e-4 / bound == pytest.approx(0.637, abs=1e-3)
    spec = PropagationSpec(dt=5e-4, nt=10, order=8, npml=20)
    with pytest.raises(SolverError):
        check_stability(model, spec)
    check_hard_stability(model, spec)
    too_big = PropagationSpec(dt=1.001 * bound, nt=10, order=8, npml=20)
    for check in (check_stability, check_hard_stability):
        with pytest.raises(SolverError):
            check(model, too_big)

    cfg = SolverConfig(nt=10, dt=5e-4, order=8, boundary="cpml", engine="acoustic", npml=20)
    with pytest.raises(SolverError):
        build_solver(cfg).spec(model)
    assert build_solver(cfg, None, GpuConfig(strict_cfl=False)).spec(model).dt == 5e-4
    hot = SolverConfig(nt=10, dt=1.001 * bound, order=8, boundary="cpml", engine="acoustic",
                       npml=20)
    with pytest.raises(SolverError):
        build_solver(hot, None, GpuConfig(strict_cfl=False)).spec(model)


def test_imager_refuses_unstable_step_before_any_gpu_work():
    """The imager applies the same checks and raises ImagingError (no GPU is
    needed to refuse)."""
    from ertm.config import GpuConfig, ImagingConfig, SolverConfig
    from ertm.exceptions import ImagingError
    from ertm.imaging import RTMImager
    from ertm.types import ReceiverSpec, ShotRecord, SourceSpec

    model = _model(vp=3500.0, nx=60, nz=40, dx=5.0, dz=5.0)
    nt = 20
    rec = ReceiverSpec(x=np.arange(0.0, 295.0, 5.0), z=np.full(59, 10.0))
    shot = ShotRecord(data=np.zeros((nt, 59), np.float32), dt=2e-3, t=np.arange(nt) * 2e-3,
                      receivers=rec, source=SourceSpec(x=150.0, z=10.0, f0=25.0))
    cfg = SolverConfig(nt=nt, dt=2e-3, order=8, boundary="cpml", engine="acoustic", npml=8)
    for gpu in (GpuConfig(), GpuConfig(strict_cfl=False)):
        with pytest.raises(ImagingError, match="CFL"):
            RTMImager(ImagingConfig(), cfg, None, gpu).migrate(shot, model)


@pytest.mark.parametrize("engine", ["acoustic", "elastic", "viscoacoustic"])
def test_float64_spec_is_refused(engine):
    """A float64 spec (the reference's PropagationSpec.dtype,
    ertm/propagation.py:52, :207) is an error: the GPU engine computes in
    float32 only and does not narrow a request silently."""
    from ertm.exceptions import SolverError
    from ertm.propagation import PropagationSpec, QOptions, build_propagator

    spec = PropagationSpec(dt=4e-4, nt=10, order=8, npml=8, dtype="float64")
    q = QOptions(enabled=True) if engine == "viscoacoustic" else None
    with pytest.raises(SolverError, match="float32 only"):
        build_propagator(_model(), spec, engine=engine, q_options=q)
