""".npz persistence of models, shot records and images (ertm/data.py:1-297).

The file layout is the reference's, so files move between the two packages
in both directions:

* directory tree: ``root/{logs,models,shots,images,reports}``;
* every ``.npz`` holds a ``meta`` entry, a 0-d unicode array with a JSON
  object (grid, units, the object's own meta);
* model: ``vp``, ``vs``, ``rho``, ``qp``, ``qs`` + meta {kind
  "elastic_model", name, grid, model_meta, units};
* shot: ``data``, ``t``, ``rx``, ``rz``, ``component_<i>`` for each component
  in order, ``component_names`` + meta {kind "shot_record", dt, component,
  components, source {x, z, f0, kind, amplitude, delay, mechanism},
  shot_meta, units};
* image: ``image`` + meta {kind "image_result", mode, grid, image_meta,
  units}.

Files are written with ``np.savez_compressed`` and read with
``np.load(allow_pickle=False)``: nothing in a file is executed.
"""

from __future__ import annotations

import json
import logging
from pathlib import Path
from typing import Any

import numpy as np

from .exceptions import DataError
from .types import ElasticModel, GridSpec, ImageResult, ReceiverSpec, ShotRecord, SourceSpec

logger = logging.getLogger("ertm.data")

OUTPUT_SUBDIRS = ("logs", "models", "shots", "images", "reports")


def ensure_output_dirs(root: str | Path) -> dict[str, Path]:
    """Create root and its standard subdirectories (ertm/data.py:35-54)."""
    base = Path(root).expanduser()
    try:
        base.mkdir(parents=True, exist_ok=True)
        out = {"root": base}
        for name in OUTPUT_SUBDIRS:
            (base / name).mkdir(parents=True, exist_ok=True)
            out[name] = base / name
    except OSError as exc:
        raise DataError(f"cannot create output directories under {base}: {exc}") from exc
    if not base.is_dir():
        raise DataError(f"output root is not a directory: {base}")
    return out


def _meta_array(meta: dict[str, Any]) -> np.ndarray:
    try:
        return np.array(json.dumps(meta, ensure_ascii=False, default=str))
    except (TypeError, ValueError) as exc:
        raise DataError(f"metadata is not JSON-serialisable: {exc}") from exc


def _meta_dict(entry: np.ndarray) -> dict[str, Any]:
    try:
        value = json.loads(str(entry.item()))
    except (ValueError, TypeError) as exc:
        raise DataError(f"metadata is not valid JSON: {exc}") from exc
    if not isinstance(value, dict):
        raise DataError("metadata must be a JSON object")
    return value


def _grid_meta(g: GridSpec) -> dict[str, float]:
    return {"nx": g.nx, "nz": g.nz, "dx": g.dx, "dz": g.dz, "x0": g.x0, "z0": g.z0}


def _grid_of(m: dict[str, Any]) -> GridSpec:
    return GridSpec(nx=int(m["nx"]), nz=int(m["nz"]), dx=float(m["dx"]), dz=float(m["dz"]),
                    x0=float(m.get("x0", 0.0)), z0=float(m.get("z0", 0.0)))


class NumpyDataStore:
    """Save and load models, shots, images and JSON reports under one root
    (ertm/data.py:74-297)."""

    name = "numpy-data-store"

    def __init__(self, root: str | Path) -> None:
        self.dirs = ensure_output_dirs(root)
        self.root = self.dirs["root"]

    def _save(self, path: Path, what: str, **arrays) -> Path:
        try:
            np.savez_compressed(path, **arrays)
        except OSError as exc:
            raise DataError(f"cannot save {what} {path}: {exc}") from exc
        logger.info("%s saved: %s", what, path)
        return path

    @staticmethod
    def _checked(obj, path: Path, what: str):
        try:
            obj.validate()
        except ValueError as exc:
            raise DataError(f"invalid {what} in {path}: {exc}") from exc
        return obj

    # ------------------------------------------------------------- model
    def save_model(self, model: ElasticModel, name: str = "model") -> Path:
        model.validate()
        meta = {"kind": "elastic_model", "name": model.name, "grid": _grid_meta(model.grid),
                "model_meta": model.meta,
                "units": {"length": "m", "velocity": "m/s", "density": "kg/m3"}}
        return self._save(self.dirs["models"] / f"{name}.npz", "model", vp=model.vp, vs=model.vs,
                          rho=model.rho, qp=model.qp, qs=model.qs, meta=_meta_array(meta))

    def load_model(self, path: str | Path) -> ElasticModel:
        path = Path(path)
        try:
            with np.load(path, allow_pickle=False) as d:
                meta = _meta_dict(d["meta"])
                model = ElasticModel(grid=_grid_of(meta["grid"]), vp=np.asarray(d["vp"]),
                                     vs=np.asarray(d["vs"]), rho=np.asarray(d["rho"]),
                                     qp=np.asarray(d["qp"]), qs=np.asarray(d["qs"]),
                                     name=str(meta.get("name", path.stem)),
                                     meta=dict(meta.get("model_meta", {})))
        except (OSError, KeyError, ValueError) as exc:
            raise DataError(f"cannot load model {path}: {exc}") from exc
        return self._checked(model, path, "model")

    # -------------------------------------------------------------- shot
    def save_shot(self, shot: ShotRecord, name: str = "shot") -> Path:
        shot.validate()
        src = shot.source
        meta = {"kind": "shot_record", "dt": shot.dt, "component": shot.receivers.component,
                "components": list(shot.component_names),
                "source": {"x": src.x, "z": src.z, "f0": src.f0, "kind": src.kind,
                           "amplitude": src.amplitude, "delay": src.delay,
                           "mechanism": src.mechanism},
                "shot_meta": shot.meta, "units": {"time": "s", "amplitude": "a.u."}}
        arrays: dict[str, Any] = {"data": shot.data, "t": shot.t, "rx": shot.receivers.x,
                                  "rz": shot.receivers.z, "meta": _meta_array(meta)}
        for i, arr in enumerate(shot.components.values()):
            arrays[f"component_{i}"] = arr
        arrays["component_names"] = np.array(list(shot.components), dtype=str)
        return self._save(self.dirs["shots"] / f"{name}.npz", "shot record", **arrays)

    def load_shot(self, path: str | Path) -> ShotRecord:
        path = Path(path)
        try:
            with np.load(path, allow_pickle=False) as d:
                meta = _meta_dict(d["meta"])
                s = meta["source"]
                source = SourceSpec(x=float(s["x"]), z=float(s["z"]), f0=float(s["f0"]),
                                    kind=str(s["kind"]), amplitude=float(s["amplitude"]),
                                    delay=float(s["delay"]),
                                    mechanism=str(s.get("mechanism", "explosive")))
                receivers = ReceiverSpec(x=np.asarray(d["rx"], dtype=float),
                                         z=np.asarray(d["rz"], dtype=float),
                                         component=str(meta.get("component", "vz")))
                comps: dict[str, np.ndarray] = {}
                if "component_names" in d:
                    for i, nm in enumerate(str(v) for v in np.asarray(d["component_names"])):
                        if f"component_{i}" in d:
                            comps[nm] = np.asarray(d[f"component_{i}"], dtype=np.float32)
                shot = ShotRecord(data=np.asarray(d["data"]), dt=float(meta["dt"]),
                                  t=np.asarray(d["t"], dtype=float), receivers=receivers,
                                  source=source, meta=dict(meta.get("shot_meta", {})),
                                  components=comps)
        except (OSError, KeyError, ValueError) as exc:
            raise DataError(f"cannot load shot record {path}: {exc}") from exc
        return self._checked(shot, path, "shot record")

    # ------------------------------------------------------------- image
    def save_image(self, image: ImageResult, name: str = "image") -> Path:
        image.validate()
        meta = {"kind": "image_result", "mode": image.mode, "grid": _grid_meta(image.grid),
                "image_meta": image.meta, "units": {"length": "m", "amplitude": "a.u."}}
        return self._save(self.dirs["images"] / f"{name}.npz", "image", image=image.image,
                          meta=_meta_array(meta))

    def load_image(self, path: str | Path) -> ImageResult:
        path = Path(path)
        try:
            with np.load(path, allow_pickle=False) as d:
                meta = _meta_dict(d["meta"])
                image = ImageResult(image=np.asarray(d["image"]), grid=_grid_of(meta["grid"]),
                                    mode=str(meta.get("mode", "pp")),
                                    meta=dict(meta.get("image_meta", {})))
        except (OSError, KeyError, ValueError) as exc:
            raise DataError(f"cannot load image {path}: {exc}") from exc
        return self._checked(image, path, "image")

    # -------------------------------------------------------------- json
    def save_json(self, payload: Any, name: str, subdir: str = "reports") -> Path:
        """Any JSON-serialisable object under one of the subdirectories."""
        if subdir not in self.dirs:
            raise DataError(f"unknown output subdirectory {subdir!r}")
        path = self.dirs[subdir] / f"{name}.json"
        try:
            path.write_text(json.dumps(payload, ensure_ascii=False, indent=2, default=str) + "\n",
                            encoding="utf-8")
        except (OSError, TypeError, ValueError) as exc:
            raise DataError(f"cannot save JSON {path}: {exc}") from exc
        return path
