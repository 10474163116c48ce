"""Data contracts exchanged with callers (ertm/types.py:20-254).

Field names, array conventions ((nz, nx) grids, (nt, nrec) records) and
validation rules are those of the reference, so records, models and images
pass between the reference's pipeline and this engine unchanged.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np


@dataclass(frozen=True)
class GridSpec:
    """Regular 2-D grid (ertm/types.py:20-82)."""

    nx: int
    nz: int
    dx: float
    dz: float
    x0: float = 0.0
    z0: float = 0.0

    def __post_init__(self) -> None:
        if self.nx < 2 or self.nz < 2:
            raise ValueError(f"grid needs at least 2x2 cells, got {self.nx}x{self.nz}")
        if self.dx <= 0 or self.dz <= 0:
            raise ValueError(f"grid spacing must be positive, got dx={self.dx}, dz={self.dz}")

    @property
    def shape(self) -> tuple[int, int]:
        return (self.nz, self.nx)

    @property
    def size(self) -> int:
        return self.nx * self.nz

    @property
    def x_min(self) -> float:
        return float(self.x0)

    @property
    def x_max(self) -> float:
        return float(self.x0 + (self.nx - 1) * self.dx)

    @property
    def z_min(self) -> float:
        return float(self.z0)

    @property
    def z_max(self) -> float:
        return float(self.z0 + (self.nz - 1) * self.dz)

    @property
    def extent(self) -> tuple[float, float, float, float]:
        return (self.x_min, self.x_max, self.z_max, self.z_min)

    def x_at(self, ix: int) -> float:
        return float(self.x0 + ix * self.dx)

    def z_at(self, iz: int) -> float:
        return float(self.z0 + iz * self.dz)

    def index_of(self, x: float, z: float, *, clip: bool = True) -> tuple[int, int]:
        """Nearest node (ix, iz) with Python's round-half-to-even (ertm/types.py:73-82)."""
        ix = int(round((float(x) - self.x0) / self.dx))
        iz = int(round((float(z) - self.z0) / self.dz))
        if clip:
            ix = min(max(ix, 0), self.nx - 1)
            iz = min(max(iz, 0), self.nz - 1)
        elif not (0 <= ix < self.nx and 0 <= iz < self.nz):
            raise ValueError(f"point ({x}, {z}) lies outside the grid")
        return ix, iz


@dataclass
class ElasticModel:
    """Earth model on the grid (ertm/types.py:85-134)."""

    grid: GridSpec
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    qp: np.ndarray
    qs: np.ndarray
    name: str = "model"
    meta: dict[str, Any] = field(default_factory=dict)

    @property
    def shape(self) -> tuple[int, int]:
        return self.grid.shape

    def validate(self) -> "ElasticModel":
        """Shapes, finiteness, positivity and vs < vp (ertm/types.py:106-122)."""
        for key in ("vp", "vs", "rho", "qp", "qs"):
            value = getattr(self, key)
            if not isinstance(value, np.ndarray):
                raise ValueError(f"model field {key} must be a numpy array")
            if value.shape != self.shape:
                raise ValueError(f"model field {key} has shape {value.shape}, grid is {self.shape}")
            if not np.all(np.isfinite(value)):
                raise ValueError(f"model field {key} contains NaN/Inf")
            if np.any(value <= 0):
                raise ValueError(f"model field {key} must be strictly positive")
        if np.any(self.vs >= self.vp):
            raise ValueError("model must satisfy vs < vp")
        return self

    def summary(self) -> dict[str, dict[str, float]]:
        return {
            key: {"min": float(np.min(getattr(self, key))), "max": float(np.max(getattr(self, key))),
                  "mean": float(np.mean(getattr(self, key)))}
            for key in ("vp", "vs", "rho", "qp", "qs")
        }


@dataclass(frozen=True)
class SourceSpec:
    """Point source (ertm/types.py:137-159)."""

    x: float
    z: float
    f0: float = 20.0
    kind: str = "ricker"
    amplitude: float = 1.0
    delay: float = 0.0
    mechanism: str = "explosive"

    def __post_init__(self) -> None:
        if self.mechanism not in {"explosive", "shear", "force_z"}:
            raise ValueError(
                f"source.mechanism must be explosive/shear/force_z, got {self.mechanism!r}")


@dataclass(frozen=True)
class ReceiverSpec:
    """Receiver positions (ertm/types.py:162-182)."""

    x: np.ndarray
    z: np.ndarray
    component: str = "vz"

    def __post_init__(self) -> None:
        x = np.asarray(self.x, dtype=float).reshape(-1)
        z = np.asarray(self.z, dtype=float).reshape(-1)
        if x.size == 0 or z.size == 0:
            raise ValueError("receiver list is empty")
        if x.size != z.size:
            raise ValueError(f"receiver x/z counts differ: {x.size} vs {z.size}")
        object.__setattr__(self, "x", x)
        object.__setattr__(self, "z", z)

    @property
    def nrec(self) -> int:
        return int(self.x.size)


@dataclass
class ShotRecord:
    """One shot's record (ertm/types.py:185-236)."""

    data: np.ndarray
    dt: float
    t: np.ndarray
    receivers: ReceiverSpec
    source: SourceSpec
    meta: dict[str, Any] = field(default_factory=dict)
    components: dict[str, np.ndarray] = field(default_factory=dict)

    def validate(self) -> "ShotRecord":
        data = np.asarray(self.data)
        if data.ndim != 2:
            raise ValueError(f"shot record must be 2-D (nt, nrec), got {data.shape}")
        if data.shape[0] != self.t.size:
            raise ValueError(f"record has {data.shape[0]} samples but t has {self.t.size}")
        if data.shape[1] != self.receivers.nrec:
            raise ValueError(f"record has {data.shape[1]} traces, {self.receivers.nrec} receivers")
        if self.dt <= 0:
            raise ValueError("dt must be positive")
        if not np.all(np.isfinite(data)):
            raise ValueError("shot record contains NaN/Inf")
        for name, array in self.components.items():
            comp = np.asarray(array)
            if comp.shape != data.shape:
                raise ValueError(f"component {name!r} shape {comp.shape} != {data.shape}")
            if not np.all(np.isfinite(comp)):
                raise ValueError(f"component {name!r} contains NaN/Inf")
        return self

    def component(self, name: str) -> np.ndarray:
        if name in self.components:
            return self.components[name]
        if name == "primary":
            return self.data
        raise KeyError(f"no component {name!r}; available: {sorted(self.components) or ['primary']}")

    @property
    def component_names(self) -> tuple[str, ...]:
        return tuple(self.components) or ("primary",)


@dataclass
class ImageResult:
    """Migrated image (ertm/types.py:239-254)."""

    image: np.ndarray
    grid: GridSpec
    mode: str = "pp"
    meta: dict[str, Any] = field(default_factory=dict)

    def validate(self) -> "ImageResult":
        image = np.asarray(self.image)
        if image.shape != self.grid.shape:
            raise ValueError(f"image shape {image.shape} != grid {self.grid.shape}")
        if not np.all(np.isfinite(image)):
            raise ValueError("image contains NaN/Inf")
        return self
