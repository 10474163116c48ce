"""Formats: SEG-Y (ertm/segy.py) and the .npz store (ertm/data.py).

The SEG-Y tests are the reference's round trips (tests/test_io.py:48-111),
plus checks of the byte layout and of IBM floats against values fixed by the
format's definition; the .npz tests check round trips and the key layout of
the reference's files."""

from __future__ import annotations

import json
from pathlib import Path

STORE = Path(__file__).resolve().parent / "golden" / "store"

import numpy as np
import pytest

from ertm import segy
from ertm.data import NumpyDataStore, ensure_output_dirs
from ertm.exceptions import DataError
from ertm.types import ElasticModel, GridSpec, ImageResult, ReceiverSpec, ShotRecord, SourceSpec


def make_shot(source_x=100.0, nrec=11, nt=64, comps=("p",)):
    x = np.linspace(0.0, 200.0, nrec)
    rng = np.random.default_rng(int(source_x))
    data = {c: rng.normal(size=(nt, nrec)).astype(np.float32) for c in comps}
    return ShotRecord(data=data[comps[0]], dt=4.0e-4, t=np.arange(nt) * 4.0e-4,
                      receivers=ReceiverSpec(x=x, z=np.full(nrec, 20.0), component="p"),
                      source=SourceSpec(x=source_x, z=20.0, f0=20.0), components=data)


# ------------------------------------------------------------------ SEG-Y
def test_shot_roundtrip(tmp_path):
    shot = make_shot(120.0)
    segy.write_segy_shot(shot, tmp_path / "shot.sgy")
    (rec,) = segy.read_segy_shots(tmp_path / "shot.sgy")
    assert rec.receivers.nrec == shot.receivers.nrec
    assert rec.dt == pytest.approx(shot.dt, abs=1e-9)
    assert rec.source.x == pytest.approx(120.0, abs=1e-3)
    assert np.array_equal(rec.data, shot.data)          # IEEE samples: exact
    assert np.allclose(rec.receivers.x, shot.receivers.x, atol=1e-6)


def test_multi_shot_grouping(tmp_path):
    shots = [make_shot(100.0), make_shot(300.0)]
    segy.write_segy_shots(shots, tmp_path / "multi.sgy")
    grouped = segy.read_segy_shots(tmp_path / "multi.sgy", segy.SegYReadOptions(group_by="source"))
    single = segy.read_segy_shots(tmp_path / "multi.sgy", segy.SegYReadOptions(group_by="file"))
    by_record = segy.read_segy_shots(tmp_path / "multi.sgy",
                                     segy.SegYReadOptions(group_by="field_record"))
    assert len(grouped) == 2 and len(single) == 1 and len(by_record) == 2
    assert single[0].receivers.nrec == 2 * shots[0].receivers.nrec
    assert sorted(r.source.x for r in grouped) == [100.0, 300.0]
    for r, s in zip(sorted(grouped, key=lambda r: r.source.x), shots):
        assert np.array_equal(r.data, s.data)


def test_directory_and_glob_input(tmp_path):
    shots = [make_shot(100.0, nrec=5), make_shot(200.0, nrec=5)]
    d = tmp_path / "line1"
    d.mkdir()
    segy.write_segy_shot(shots[0], d / "shot_001.sgy")
    segy.write_segy_shot(shots[1], d / "shot_002.sgy")
    assert len(segy.read_segy_shots(d)) == 2
    assert len(segy.read_segy_shots(str(d / "*.sgy"))) == 2
    summary = segy.summarize_segy(d)
    assert (summary["n_files"], summary["n_shots"], summary["n_traces"]) == (2, 2, 10)
    assert summary["dt"] == pytest.approx(4e-4, abs=1e-9)
    assert summary["offset_range"] == [0.0, 200.0]
    with pytest.raises(DataError):
        segy.read_segy_shots(tmp_path / "missing.sgy")


def test_image_roundtrip(tmp_path):
    grid = GridSpec(nx=13, nz=17, dx=10.0, dz=5.0)
    image = ImageResult(image=np.random.default_rng(3).normal(size=grid.shape).astype(np.float32),
                        grid=grid, mode="pp")
    segy.write_segy_image(image, tmp_path / "image.sgy")
    back, dz, x = segy.read_segy_image(tmp_path / "image.sgy")
    assert back.shape == grid.shape and dz == pytest.approx(5.0)
    assert np.allclose(x, np.arange(grid.nx) * grid.dx) and np.array_equal(back, image.image)


def test_byte_layout(tmp_path):
    shot = make_shot(150.0, nrec=3, nt=10)
    p = segy.write_segy_shot(shot, tmp_path / "s.sgy")
    raw = p.read_bytes()
    assert len(raw) == 3600 + 3 * (240 + 10 * 4)
    assert raw[:3200].decode("cp500").startswith("C 1 ERTM shot 150.0/20.0")
    be = lambda off, n: int.from_bytes(raw[off:off + n], "big", signed=True)  # noqa: E731
    assert (be(3216, 2), be(3220, 2), be(3224, 2)) == (400, 10, 5)     # interval us, ns, format
    t1 = 3600 + 240 + 40                                                 # second trace header
    assert (be(t1 + 72, 4), be(t1 + 80, 4), be(t1 + 70, 2)) == (150, 100, 1)
    assert (be(t1 + 36, 4), be(t1 + 114, 2), be(t1 + 116, 2)) == (50, 10, 400)
    first = np.frombuffer(raw[3600 + 240:3600 + 240 + 40], dtype=">f4")
    assert np.array_equal(first, shot.data[:, 0])


def test_ibm_known_values():
    words = np.array([0x42640000, 0xC276A000, 0x41100000, 0x00000000, 0x3F800000], np.uint32)
    vals = segy.ibm_to_float32(words)
    assert vals.tolist() == [100.0, -118.625, 1.0, 0.0, 0.03125]
    assert segy.float32_to_ibm(vals).tolist() == words.tolist()
    # 0.1: 0x19999999A... rounds to 0x4019999A
    assert int(segy.float32_to_ibm(np.array([0.1], np.float64))[0]) == 0x4019999A


def test_ibm_roundtrip_and_precision(tmp_path):
    rng = np.random.default_rng(0)
    # normalised IBM words (leading hex digit nonzero) convert to float32 and back exactly
    expo = rng.integers(40, 90, 5000).astype(np.uint32)
    frac = rng.integers(0x100000, 0x1000000, 5000).astype(np.uint32)
    words = (rng.integers(0, 2, 5000).astype(np.uint32) << 31) | (expo << 24) | frac
    assert np.array_equal(segy.float32_to_ibm(segy.ibm_to_float32(words)), words)
    # float32 -> IBM keeps 21..24 bits: relative error <= 2^-21
    x = (rng.normal(size=5000) * 10.0 ** rng.uniform(-8, 8, 5000)).astype(np.float32)
    back = segy.ibm_to_float32(segy.float32_to_ibm(x))
    assert np.all(np.abs(back.astype(np.float64) - x) <= 2.0 ** -21 * np.abs(x.astype(np.float64)))
    shot = make_shot(80.0, nrec=4, nt=32)
    segy.write_segy_shot(shot, tmp_path / "ibm.sgy", fmt=1)
    (rec,) = segy.read_segy_shots(tmp_path / "ibm.sgy")
    assert np.array_equal(rec.data, segy.ibm_to_float32(segy.float32_to_ibm(shot.data)))


def test_components_and_scalars(tmp_path):
    shot = make_shot(60.0, nrec=6, nt=20, comps=("vx", "vz"))
    segy.write_segy_shot(shot, tmp_path / "vx.sgy", component="vx")
    segy.write_segy_shot(shot, tmp_path / "vz.sgy", component="vz")
    (m,) = segy.read_segy_components({"vx": tmp_path / "vx.sgy", "vz": tmp_path / "vz.sgy"})
    assert np.array_equal(m.components["vx"], shot.components["vx"])
    assert np.array_equal(m.components["vz"], shot.components["vz"])
    (r,) = segy.read_segy_shots(tmp_path / "vx.sgy",
                                segy.SegYReadOptions(coordinate_scalar=-10,
                                                     receiver_depth_from="group_y"))
    assert np.allclose(r.receivers.x, np.rint(shot.receivers.x) / 10.0)
    assert np.allclose(r.receivers.z, 2.0)
    assert np.allclose(segy._apply_scalar(np.array([7, 7, 7]), np.array([0, 3, -2])),
                       [7.0, 21.0, 3.5])


# ------------------------------------------------------------------ .npz
def _model():
    g = GridSpec(nx=9, nz=7, dx=5.0, dz=4.0, x0=10.0)
    vp = np.linspace(2000, 3000, 63, dtype=np.float32).reshape(7, 9)
    return ElasticModel(g, vp=vp, vs=vp / np.float32(1.8), rho=np.full_like(vp, 2200.0),
                        qp=np.full_like(vp, 90.0), qs=np.full_like(vp, 40.0), name="m1",
                        meta={"builder": "test"})


def test_npz_store_roundtrips(tmp_path):
    store = NumpyDataStore(tmp_path / "out")
    assert all((tmp_path / "out" / d).is_dir() for d in ("logs", "models", "shots", "images",
                                                        "reports"))
    m = _model()
    lm = store.load_model(store.save_model(m, "vp"))
    assert lm.grid == m.grid and lm.name == "m1" and lm.meta == {"builder": "test"}
    for f in ("vp", "vs", "rho", "qp", "qs"):
        assert np.array_equal(getattr(lm, f), getattr(m, f))
    s = make_shot(40.0, comps=("vx", "vz"))
    ls = store.load_shot(store.save_shot(s, "s1"))
    assert list(ls.components) == ["vx", "vz"] and ls.source == s.source
    assert np.array_equal(ls.components["vz"], s.components["vz"]) and ls.dt == s.dt
    im = ImageResult(np.arange(63, dtype=np.float32).reshape(7, 9), m.grid, "ps", {"k": 1})
    li = store.load_image(store.save_image(im, "i1"))
    assert li.mode == "ps" and li.meta == {"k": 1} and np.array_equal(li.image, im.image)
    assert json.loads(store.save_json({"a": 1}, "r").read_text()) == {"a": 1}


def test_npz_key_layout_matches_the_reference_files(tmp_path):
    store = NumpyDataStore(tmp_path)
    s = make_shot(40.0, comps=("p", "vz"))
    with np.load(store.save_shot(s, "s"), allow_pickle=False) as d:
        assert set(d.files) == {"data", "t", "rx", "rz", "meta", "component_0", "component_1",
                                "component_names"}
        meta = json.loads(str(d["meta"].item()))
        assert meta["kind"] == "shot_record" and meta["components"] == ["p", "vz"]
        assert set(meta["source"]) == {"x", "z", "f0", "kind", "amplitude", "delay", "mechanism"}
    with np.load(store.save_model(_model(), "m"), allow_pickle=False) as d:
        assert set(d.files) == {"vp", "vs", "rho", "qp", "qs", "meta"}
        assert json.loads(str(d["meta"].item()))["grid"] == {"nx": 9, "nz": 7, "dx": 5.0,
                                                              "dz": 4.0, "x0": 10.0, "z0": 0.0}
    # a file laid out by hand like the reference's loads
    g = {"nx": 3, "nz": 2, "dx": 1.0, "dz": 1.0}
    np.savez(tmp_path / "img.npz", image=np.ones((2, 3), np.float32),
             meta=np.array(json.dumps({"kind": "image_result", "mode": "pp", "grid": g})))
    assert store.load_image(tmp_path / "img.npz").grid.shape == (2, 3)
    with pytest.raises(DataError):
        store.load_model(tmp_path / "img.npz")
    assert ensure_output_dirs(tmp_path / "x")["reports"] == tmp_path / "x" / "reports"


def _npz(path):
    with np.load(path, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_files_written_by_the_reference_store_load(tmp_path):
    """model/shot/image .npz files saved by the reference's NumpyDataStore
    (ertm/data.py:84-114, 149-184, 231-253) load into the same objects: the
    arrays equal the inputs they were made from (stored separately with plain
    numpy), and grid, names, source, receivers and meta are those written."""
    store = NumpyDataStore(tmp_path)
    raw = _npz(STORE / "inputs.npz")
    m = store.load_model(STORE / "model.npz")
    assert m.grid == GridSpec(nx=9, nz=7, dx=5.0, dz=4.0, x0=10.0, z0=2.0)
    assert m.name == "layers" and m.meta == {"builder": "golden", "seed": 2024}
    for f in ("vp", "vs", "rho", "qp", "qs"):
        assert getattr(m, f).dtype == np.float32 and np.array_equal(getattr(m, f), raw[f]), f
    s = store.load_shot(STORE / "shot.npz")
    assert list(s.components) == ["p", "vx", "vz"]
    for c in ("p", "vx", "vz"):
        assert np.array_equal(s.components[c], raw[f"comp_{c}"]), c
    assert np.array_equal(s.data, raw["comp_vz"]) and s.dt == 2.5e-4
    assert np.array_equal(s.t, raw["t"])
    assert np.array_equal(s.receivers.x, raw["rx"]) and np.array_equal(s.receivers.z, raw["rz"])
    assert s.receivers.component == "vz" and s.meta == {"shot_index": 3}
    assert s.source == SourceSpec(x=30.0, z=8.0, f0=25.0, kind="ricker", amplitude=2.0,
                                  delay=0.04, mechanism="force_z")
    im = store.load_image(STORE / "image.npz")
    assert im.mode == "ps" and im.meta == {"condition": "normalized"} and im.grid == m.grid
    assert np.array_equal(im.image, raw["image"])


@pytest.mark.parametrize("kind", ["model", "shot", "image"])
def test_store_rewrites_reference_files_with_the_same_keys_and_meta(tmp_path, kind):
    """Loading a reference-written file and saving it again with this store
    gives the same keys, the same arrays (dtype included) and the same JSON
    meta as the reference's file."""
    store = NumpyDataStore(tmp_path)
    obj = getattr(store, f"load_{kind}")(STORE / f"{kind}.npz")
    ref, mine = _npz(STORE / f"{kind}.npz"), _npz(getattr(store, f"save_{kind}")(obj, "x"))
    assert set(mine) == set(ref)
    for k in ref:
        if k == "meta":
            assert json.loads(str(mine[k].item())) == json.loads(str(ref[k].item()))
        else:
            assert mine[k].dtype == ref[k].dtype and np.array_equal(mine[k], ref[k]), k
