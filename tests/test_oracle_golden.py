"""Pin the CPU oracle to the reference's own outputs (golden vectors).

Every comparison is bit-exact (``np.array_equal``): the oracle reproduces the
reference's float32 operation order, so any dropped term, swapped index or
reordered product shows up as a mismatch.  The fixtures were written by
``tests/golden/make_golden.py`` from the reference package.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from golden_cases import forward_case, load, migrate_case, tags
from oracle import engines as oeng
from oracle import gsls as ogsls
from oracle import numerics as onum
from oracle import rtm as ortm


def test_derivatives_match_reference():
    d = load("numerics")
    f = d["deriv_field"]
    for order in (2, 4, 6, 8):
        for axis in (0, 1):
            assert np.array_equal(onum.derivative(f, axis, order, 3.7, True),
                                  d[f"dplus_o{order}_a{axis}"])
            assert np.array_equal(onum.derivative(f, axis, order, 3.7, False),
                                  d[f"dminus_o{order}_a{axis}"])


def test_cpml_profiles_match_reference():
    d = load("numerics")
    for k in range(3):
        n, h, npml, dt, vmax, f0, lo, hi = d[f"cpml{k}_params"]
        a, b, _ = onum.cpml_profile(int(n), h, int(npml), dt, vmax, f0, bool(lo), bool(hi))
        assert np.array_equal(a, d[f"cpml{k}_a"])
        assert np.array_equal(b, d[f"cpml{k}_b"])


def test_source_wavelets_match_reference():
    """Gaussian pulse and the amplitude-scaled wavelets of source_wavelet
    (ertm/numerics.py:466-475, ertm/propagation.py:1004-1015), bit for bit."""
    d = load("numerics")
    t = np.arange(400, dtype=np.float64) * 2.5e-4
    for k in range(3):
        f0, t0 = d[f"gauss{k}_params"]
        assert np.array_equal(onum.gaussian(t, f0, t0), d[f"gauss{k}"])
    for k in range(3):
        f0, delay, amp, dt, nt = d[f"wavelet{k}_params"]
        w = ortm.wavelet(str(d[f"wavelet{k}_kind"]), f0, delay, amp, dt, int(nt))
        assert w.dtype == np.float32
        assert np.array_equal(w, d[f"wavelet{k}"])
    # the two kinds differ: a swapped branch would not pass the lines above
    assert not np.array_equal(d["wavelet0"], d["wavelet2"])


def test_resample_sinc_matches_reference():
    """Kaiser-windowed sinc resampling (ertm/numerics.py:337-392) on the
    36 us <-> 0.35 ms pairs of the borehole workload, the reference test's
    0.4 -> 0.1 ms, a non-divisible pair and the three-sample minimum.
    Several traces: bit for bit.  One trace: NumPy's einsum then reduces the
    16 taps with a vectorised kernel whose order the oracle does not follow,
    so within 1e-13 of the largest sample."""
    d = load("numerics")
    for k in range(5):
        dt_in, dt_out = d[f"sinc{k}_params"]
        x, ref = d[f"sinc{k}_in"], d[f"sinc{k}_out"]
        got = onum.resample_sinc(x, dt_in, dt_out)
        assert got.shape == ref.shape
        assert got.shape[0] == int(np.floor((x.shape[0] - 1) * dt_in / dt_out + 1e-9)) + 1
        if x.shape[1] > 1:
            assert np.array_equal(got, ref), k
        else:
            assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)), k
        one = onum.resample_sinc(d[f"sinc{k}_in1"], dt_in, dt_out)
        ref1 = d[f"sinc{k}_out1"]
        assert np.max(np.abs(one - ref1)) <= 1e-13 * np.max(np.abs(ref1)), k


def test_resample_sinc_preserves_a_sine():
    """The reference's own accuracy test (tests/test_rtm.py:159-167): a 35 Hz
    sine at 0.4 ms resampled to 0.1 ms within 2e-2 of the analytic sine, the
    last 8 samples (short of taps) excluded."""
    dt_in, dt_out = 4.0e-4, 1.0e-4
    signal = np.sin(2 * np.pi * 35.0 * np.arange(500) * dt_in)
    out = onum.resample_sinc(signal, dt_in, dt_out)
    analytic = np.sin(2 * np.pi * 35.0 * np.arange(out.size) * dt_out)
    assert out.size == 1997
    assert np.max(np.abs(out[:-8] - analytic[:-8])) < 2e-2


def test_cfl_limits_match_reference():
    """cfl_dt (ertm/numerics.py:314-326) at the workloads' media and steps."""
    for vmax, dx, dz, order, safety, ref in load("numerics")["cfl_table"]:
        assert onum.cfl_dt(vmax, dx, dz, int(order), safety) == ref


def test_gsls_fit_matches_reference():
    rows = load("gsls")["table"]
    for row in rows:
        q, f_ref, n, err = row[:4]
        tau, delta, e = ogsls.fit(q, f_ref, int(n), 1.5)
        k = int(n)
        np.testing.assert_allclose(tau, row[4 : 4 + k], rtol=1e-12)
        np.testing.assert_allclose(delta, row[7 : 7 + k], rtol=1e-12)
        assert abs(e - err) <= 1e-12


def _oracle_prop(c, q_override=None):
    m = ortm.Model(c["nz"], c["nx"], c["dx"], c["dz"], c["vp"], c["vs"], c["rho"], c["qp"],
                   c["x0"], c["z0"])
    q = None
    if c["engine"] == "viscoacoustic":
        q = ortm.QSpec(**(q_override or c["q"]))
    p = ortm.build(c["engine"], m, c["dt"], c["nt"], c["order"], c["npml"], c["boundary"], q=q)
    p.set_source(c["src"]["x"], c["src"]["z"], c["src"]["mechanism"])
    p.set_receivers(c["rx"], c["rz"])
    return p


@pytest.mark.parametrize("tag", tags("forward"))
def test_forward_records_and_states_match_reference(tag):
    c = forward_case(tag)
    p = _oracle_prop(c)
    s = c["src"]
    w = ortm.wavelet(s["kind"], s["f0"], s["delay"], s["amplitude"], c["dt"], c["nt"])
    rec = ortm.forward(p, w, c["components"])
    for comp in c["components"]:
        assert np.array_equal(rec[comp], c["rec"][comp]), comp
    for name, ref in c["state"].items():
        assert np.array_equal(getattr(p, name), ref), name
    assert set(p.psi) == set(c["psi"])
    for name, ref in c["psi"].items():
        assert np.array_equal(p.psi[name], ref), name
    if c["engine"] == "viscoacoustic":
        assert np.array_equal(p.a, c["visco_a"])
        assert np.array_equal(p.b, c["visco_b"])
        assert np.array_equal(p.kr, c["visco_kr"])
        assert [p.clipped, p.lowpass_steps] == list(c["visco_summary"])


@pytest.mark.parametrize("tag", ["ac", "el", "vi"])
def test_sample_and_inject_match_reference(tag):
    d = load("inject")
    case = {"ac": "ac_cpml", "el": "el_shear", "vi": "vi_fwd"}[tag]
    c = forward_case(case)
    p = _oracle_prop(c)
    p.set_receivers(d[f"{tag}/rx"], d[f"{tag}/rz"])
    for name in p.fields:
        getattr(p, name)[...] = d[f"{tag}/before_{name}"]
    comps = [str(v) for v in d[f"{tag}/components"]]
    for comp in comps:
        assert np.array_equal(p.sample((comp,))[comp], d[f"{tag}/sample_{comp}"]), comp
    for comp in comps:
        p.inject(comp, d[f"{tag}/values_{comp}"])
    for name in p.fields:
        assert np.array_equal(getattr(p, name), d[f"{tag}/after_{name}"]), name


def test_replay_blocks_match_reference():
    d = load("replay")
    c = forward_case("ac_cpml")
    p = _oracle_prop(c)
    s = c["src"]
    w = ortm.wavelet(s["kind"], s["f0"], s["delay"], s["amplitude"], c["dt"], c["nt"])
    st = ortm.record(p, w, 1.0)
    assert st.interval == int(d["interval"])
    assert [x for x, _ in st.states] == list(d["starts"])
    for b in (0, 3, len(st.states) - 1):
        fields = ortm.replay(p, w, st, b, "stress")
        assert np.array_equal(np.stack([f["P"] for f in fields]), d[f"block{b}"])


def migrate_oracle(c, dtype="float32"):
    m = ortm.Model(c["nz"], c["nx"], c["dx"], c["dz"], c["vp"], c["vs"], c["rho"], c["qp"],
                   c["x0"], c["z0"])
    cfg = ortm.Imaging(**c["imaging"])
    q = None
    q_enabled = False
    if c["engine"] == "viscoacoustic":
        qq = c["q"]
        q = ortm.QSpec(f_ref=qq["f_ref"], n=qq["n"], band_ratio=qq["band_ratio"],
                       alpha_c=qq["alpha_c"], amplitude_cap=qq["amplitude_cap"],
                       lowpass_interval=qq["lowpass_interval"], f_max=qq["f_max"],
                       gain_budget_db=qq["gain_budget_db"])
        q_enabled = qq["enabled"]
    if dtype == "float64":
        # same checkpoint interval as the float32 run: twice the budget for
        # states twice the size
        cfg = dataclasses.replace(cfg, checkpoint_mb=2.0 * cfg.checkpoint_mb)
    return ortm.migrate(c["engine"], m, c["dt"], c["nt"], c["order"], c["boundary"], cfg,
                        c["shot"], c["src"], c["rx"], c["rz"], q=q, q_enabled=q_enabled,
                        dtype=dtype, rec_dt=c["shot_dt"])


@pytest.mark.parametrize("tag", tags("migrate"))
def test_migrate_images_match_reference(tag):
    c = migrate_case(tag)
    img, info = migrate_oracle(c)
    assert (info["interval"], info["n_states"]) == c["checkpoints"]
    assert info["start_step"] == c["start_step"]
    assert np.array_equal(img, c["image"])


def test_forward_shot_records_feed_migrate_cases():
    """The shot each migrate case used is the oracle's own forward record
    (solver path: configured npml, ertm/solver.py:62-64)."""
    c = migrate_case("el_pp_divcurl")
    m = ortm.Model(c["nz"], c["nx"], c["dx"], c["dz"], c["vp"], c["vs"], c["rho"], c["qp"])
    p = ortm.build("elastic", m, c["dt"], c["nt"], c["order"], c["npml"], c["boundary"])
    p.set_source(c["src"]["x"], c["src"]["z"])
    p.set_receivers(c["rx"], c["rz"])
    s = c["src"]
    w = ortm.wavelet(s["kind"], s["f0"], s["delay"], s["amplitude"], c["dt"], c["nt"])
    rec = ortm.forward(p, w, c["components"])
    for comp in c["components"]:
        assert np.array_equal(rec[comp], c["shot"][comp])


@pytest.mark.parametrize("tag", ["ac_cpml", "el_force", "vi_fwd"])
def test_float64_mode_runs_in_float64(tag):
    """The oracle's float64 mode (the reference's PropagationSpec.dtype,
    ertm/propagation.py:52,207) keeps the wavefields in float64 while the CPML
    profiles and the records stay float32 (ertm/numerics.py:217-219,
    ertm/propagation.py:1045).  Its records must differ from the float32 run's
    (else the float64 path is not taken) by float32 rounding only."""
    c = forward_case(tag)
    m = ortm.Model(c["nz"], c["nx"], c["dx"], c["dz"], c["vp"], c["vs"], c["rho"], c["qp"],
                   c["x0"], c["z0"])
    q = None
    if c["engine"] == "viscoacoustic":
        qq = c["q"]
        q = ortm.QSpec(f_ref=qq["f_ref"], n=qq["n"], band_ratio=qq["band_ratio"],
                       compensation=False)
    s = c["src"]
    w = ortm.wavelet(s["kind"], s["f0"], s["delay"], s["amplitude"], c["dt"], c["nt"])
    out = {}
    for dtype in ("float32", "float64"):
        p = ortm.build(c["engine"], m, c["dt"], c["nt"], c["order"], c["npml"], c["boundary"],
                       dtype=dtype, q=q)
        assert p.vz.dtype == np.dtype(dtype)
        p.set_source(s["x"], s["z"], s.get("mechanism", "explosive"))
        p.set_receivers(c["rx"], c["rz"])
        out[dtype] = ortm.forward(p, w, c["components"])
    assert all(np.array_equal(out["float32"][k], c["rec"][k]) for k in c["components"])
    for k in c["components"]:
        a = out["float32"][k].astype(np.float64)
        b = out["float64"][k].astype(np.float64)
        if not np.any(b):
            continue
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert 0.0 < err < 1e-5, (k, err)
