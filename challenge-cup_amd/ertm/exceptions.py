"""Error hierarchy of the engine (same class names and ``to_dict`` shape as
ertm/exceptions.py:12-64 so callers' ``except`` clauses keep working)."""

from __future__ import annotations

from typing import Any


class ERTMError(Exception):
    """Base of every expected engine error; carries an optional context dict."""

    def __init__(self, message: str, *, context: dict[str, Any] | None = None) -> None:
        self.message = str(message)
        self.context = dict(context or {})
        super().__init__(self.message)

    def __str__(self) -> str:
        if not self.context:
            return self.message
        extra = ", ".join(f"{k}={v!r}" for k, v in self.context.items())
        return f"{self.message} ({extra})"

    def to_dict(self) -> dict[str, Any]:
        return {"type": type(self).__name__, "message": self.message, "context": self.context}


class ConfigError(ERTMError):
    """Invalid or out-of-range configuration."""


class ModelError(ERTMError):
    """Invalid earth model."""


class SolverError(ERTMError):
    """Forward / adjoint propagation failure (CFL, engine, component, device)."""


class ImagingError(ERTMError):
    """Invalid imaging input or imaging-condition failure."""


class DataError(ERTMError):
    """Record / data-format failure."""


class PipelineError(ERTMError):
    """Orchestration failure (e.g. stacking nothing)."""
