"""Which parameters are free, and writing a solution back (src/bmc_jax/fitting/params.py:38-112)."""
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from ..simulation.params import SimParams


@dataclass(frozen=True)
class FreeParam:
    """One free value: a SimParams field (and pool index), its scale, bounds and start (params.py:38-63)."""
    name: str
    field: str
    index: Optional[int]
    scale: float
    lo: float
    hi: float
    x0: float
    scope: str = "sample"
    samples: Optional[Tuple[int, ...]] = None
    mode: str = "value"

    def __post_init__(self):
        """Checks the combination of scope, mode and field (params.py:52-63); normalises `samples` to a
        sorted tuple of distinct ints."""
        allowed = {"scope": ("sample", "shared", "block"), "mode": ("value", "factor", "add")}
        for attr, choices in allowed.items():
            if getattr(self, attr) not in choices:
                raise ValueError(f"{self.name}: {attr} must be one of {choices}, not {getattr(self, attr)!r}")
        if self.field == "m_start":
            if self.index is not None or self.mode == "add":
                raise ValueError(f"{self.name}: m_start is a single number per sample (index None), fitted as a "
                                 "value or a factor, never as an offset")
        elif self.field == "b0_spread":
            if self.index is not None or self.mode != "value":
                raise ValueError(f"{self.name}: b0_spread is the spread width itself in Hz (index None, "
                                 "mode 'value')")
        if self.samples is not None:
            chosen = tuple(sorted(set(int(i) for i in self.samples)))
            if len(chosen) == 0:
                raise ValueError(f"{self.name}: an empty samples tuple frees the value in no sample; drop the "
                                 "slot instead")
            object.__setattr__(self, "samples", chosen)


@dataclass(frozen=True)
class FitSpec:
    """The free values of a fit (params.py:66-97)."""
    slots: List[FreeParam]

    @property
    def n(self):
        return len(self.slots)

    @property
    def scales(self):
        return np.array([s.scale for s in self.slots])

    @property
    def lo_phys(self):
        return np.array([s.lo for s in self.slots])

    @property
    def hi_phys(self):
        return np.array([s.hi for s in self.slots])

    @property
    def x0_phys(self):
        return np.array([s.x0 for s in self.slots])

    def to_scaled(self, x_phys):
        return np.asarray(x_phys) / np.array([s.scale for s in self.slots])

    def to_phys(self, x_scaled):
        return np.asarray(x_scaled) * np.array([s.scale for s in self.slots])

    def _needs_joint(self, s) -> bool:
        """A slot that fitting every sample on its own cannot express."""
        return (s.scope != "sample" or s.samples is not None or s.mode != "value"
                or s.field in ("m_start", "b0_spread"))

    @property
    def independent(self) -> bool:
        """True when every value is a plain per-sample value, free everywhere (what `fit` solves,
        params.py:80-86)."""
        return not any(self._needs_joint(s) for s in self.slots)

    def require_independent(self, caller: str) -> None:
        """ValueError naming the slots that need fit_joint (params.py:88-97)."""
        joint = [s.name for s in self.slots if self._needs_joint(s)]
        if joint:
            raise ValueError(f"{caller} fits every sample on its own with plain values; {joint} couple samples "
                             "(shared or per-block scope), are free in some samples only, scale or shift a "
                             "carried value, or are m_start / b0_spread: fit them with fit_joint")


def write_params(base: SimParams, x_scaled, spec: FitSpec) -> SimParams:
    """x_scaled -> physical (rounded to the params' dtype) -> written into the named slots (params.py:100-112).

    Works for one sample (x_scaled (P,)) or a batch (x_scaled (V, P), base with a leading axis).
    """
    dt = np.asarray(base.water_r1).dtype
    x_phys = (np.asarray(x_scaled, np.float64) * spec.scales).astype(dt)
    upd = {}
    for i, s in enumerate(spec.slots):
        val = x_phys[..., i]
        if s.index is None:
            upd[s.field] = val
        else:
            cur = np.array(upd.get(s.field, np.asarray(getattr(base, s.field))), copy=True)
            cur[..., s.index] = val
            upd[s.field] = cur
    return base._replace(**upd)
