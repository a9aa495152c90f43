"""Several simulators integrated by one persistent work-queue launch.

No counterpart in the reference, which compiles one right-hand side per
program and integrates one time span per launch (clode/cpp/CLODE.cpp:148,
clode/cpp/transient.cl:10).  The member simulators may differ in system,
dimensions, stepper and observer; they must share the precision and device
and be all feature simulators or all base simulators.  Each keeps its own
parameters, state, noise stream, per-member time spans and host semantics,
and its results are bit-identical to running it on its own.
"""
from __future__ import annotations

from typing import List, Optional, Sequence, Tuple

import numpy as np

from .cpp.clode_cpp_wrapper import BatchSimulatorBase
from .features import FeatureSimulator
from .solver import Simulator


class BatchSimulator:
    def __init__(self, simulators: Sequence[Simulator]):
        sims = list(simulators)
        if not sims:
            raise ValueError("a batch needs at least one simulator")
        self._sims: List[Simulator] = sims
        self._features = all(isinstance(s, FeatureSimulator) for s in sims)
        self._batch = BatchSimulatorBase([s._integrator for s in sims])

    def __len__(self) -> int:
        return len(self._sims)

    @property
    def simulators(self) -> List[Simulator]:
        return list(self._sims)

    def set_launch(self, block_size: int = 64, slice: int = 8) -> None:
        """Workgroup size (64..256, multiple of 64) and loop iterations a lane
        runs between work-queue checks."""
        self._batch.set_launch(block_size, slice)

    def get_launch(self) -> Tuple[int, int]:
        return tuple(self._batch.get_launch())

    def autotune(self, stride: int = 16, max_steps: int = 4096) -> Tuple[int, int]:
        """Pick the fastest (block size, slice) by timing the transient launch
        over every stride-th queued member, each capped at max_steps steps
        (0: no cap), three times per shape in interleaved rounds, and keeping
        the fastest median; member state is restored.  The choice is kept in
        the code-object cache directory and reused for the same program,
        device and ensemble sizes (CLODE_AUTOTUNE_FRESH=1 re-times)."""
        return tuple(self._batch.autotune(stride, max_steps))

    def autotune_table(self) -> Tuple[np.ndarray, bool]:
        """The last autotune's trials as rows (block, slice, median ms,
        fastest ms, slowest ms), and whether the shape came from the cache
        (then the table is empty)."""
        rows, cached = self._batch.autotune_table()
        return np.asarray(rows), bool(cached)

    def _invalidate(self, x0: bool = False) -> None:
        for s in self._sims:
            s._device_final_state = s._device_dt = s._device_tf = None
            if x0:
                s._device_initial_state = None

    def transient(self, update_x0: bool = True) -> None:
        self._batch.transient()
        self._invalidate()
        if update_x0:
            self.shift_x0()

    def initialize_observer(self) -> None:
        self._batch.initialize_observer()

    def features(self, initialize_observer: Optional[bool] = None, update_x0: bool = True) -> None:
        """features() of every member simulator (observer continuation rules
        of each simulator apply); results through each simulator's
        get_observer_results()."""
        if not self._features:
            raise ValueError("features needs feature simulators")
        self._batch.features(-1 if initialize_observer is None else int(bool(initialize_observer)))
        self._invalidate()
        if update_x0:
            self.shift_x0()

    def shift_x0(self) -> None:
        self._batch.shift_x0()
        self._invalidate(x0=True)

    def sweep(self, rounds: int, graph: bool = True) -> None:
        """rounds x (transient, x0 <- xf, features, x0 <- xf); with graph=True
        the whole sequence is one captured graph launched once."""
        self._batch.sweep(rounds, graph)
        self._invalidate(x0=True)

    def last_kernel_ms(self) -> float:
        """Device time of the last batch operation, milliseconds."""
        return self._batch.last_kernel_ms()
