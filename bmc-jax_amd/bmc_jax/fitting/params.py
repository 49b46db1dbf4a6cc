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
        if self.scope not in ("sample", "shared", "block"):
            raise ValueError(f"{self.name}: scope {self.scope!r}; use 'sample', 'shared' or 'block'")
        if self.mode not in ("value", "factor", "add"):
            raise ValueError(f"{self.name}: mode {self.mode!r}; use 'value', 'factor' or 'add'")
        if self.field == "m_start" and (self.index is not None or self.mode == "add"):
            raise ValueError(f"{self.name}: m_start is one number per sample (index None), as a value or a factor")
        if self.field == "b0_spread" and (self.index is not None or self.mode != "value"):
            raise ValueError(f"{self.name}: b0_spread is one width in Hz (index None, mode 'value')")
        if self.samples is not None:
            object.__setattr__(self, "samples", tuple(sorted({int(i) for i in self.samples})))
            if not self.samples:
                raise ValueError(f"{self.name}: samples=() frees it nowhere; leave it out of the spec")


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

    @property
    def independent(self) -> bool:
        """True when every value is a plain per-sample value, free everywhere (what `fit` solves)."""
        return all(s.scope == "sample" and s.samples is None and s.mode == "value"
                   and s.field not in ("m_start", "b0_spread") for s in self.slots)

    def require_independent(self, caller: str) -> None:
        if not self.independent:
            joint = [s.name for s in self.slots if s.scope != "sample" or s.samples is not None
                     or s.mode != "value" or s.field in ("m_start", "b0_spread")]
            raise ValueError(f"{caller}: {joint} are shared, per block, free in some samples only, a factor / an "
                             "offset, m_start or b0_spread (FreeParam scope / samples / mode / field), which "
                             "samples fitted alone with plain values cannot do; use fit_joint")


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
