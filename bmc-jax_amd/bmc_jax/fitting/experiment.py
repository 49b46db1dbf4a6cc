"""An experiment: a sample, a sequence, a starting state and where water's Mz sits
(src/bmc_jax/fitting/experiment.py:21-141)."""
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from ..simulation.events import SimEvent, pad_events, concat_events
from ..simulation.simulate import simulate_spectrum


@dataclass(frozen=True)
class Experiment:
    """Parameters, a sequence, a starting state, and the index of water's Mz."""
    params: object
    events: SimEvent
    m0: np.ndarray
    mz_loc: int

    @property
    def n_offsets(self) -> int:
        return int(np.shape(self.events.dt)[0])

    @property
    def n_events(self) -> int:
        return int(np.shape(self.events.dt)[1])

    def simulate(self):
        """The Z-spectrum: water Mz at every offset (experiment.py:58-61). One sample."""
        return simulate_spectrum(self.params, self.events, self.m0, return_full_trajectory=False)[:, self.mz_loc]

    @classmethod
    def from_parameters(cls, params, events) -> "Experiment":
        """From a config object: convert pools, build M0 with options.scale, find mz (experiment.py:63-99)."""
        from ..simulation.params import to_sim_params, initial_state
        jp = to_sim_params(params)
        scale = float(getattr(params.options, "scale", 1.0))
        return cls(jp, events, initial_state(jp, scale), params.mz_loc)

    @classmethod
    def stack(cls, experiments: Sequence["Experiment"]) -> "Experiment":
        """Join sequences along the offset axis, front-padding shorter ones (experiment.py:101-141)."""
        experiments = list(experiments)
        if len(experiments) == 0:
            raise ValueError("Experiment.stack was given no experiments; it needs at least one")
        water_z = sorted({p.mz_loc for p in experiments})
        if len(water_z) != 1:
            raise ValueError(f"the experiments read water Mz at different mz_loc {water_z}: stacked rungs must "
                             "share one pool model")
        width = max(p.n_events for p in experiments)
        events = concat_events([pad_events(p.events, width) for p in experiments])
        first = experiments[0]
        return cls(first.params, events, first.m0, first.mz_loc)
