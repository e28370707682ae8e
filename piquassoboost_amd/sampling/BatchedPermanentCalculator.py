"""Batched front-end of one interferometer's permanents
(BatchedPermanentCalculator.h:52-112, .cpp:109-296).

The reference accumulates (input state, output state) pairs of one
interferometer and computes their permanents in one call -- on the CPU one
after another (GlynnRep: GlynnPermanentCalculatorRepeatedLongDouble on the
compact effective matrix, ChinHuh: CChinHuhPermanentCalculator), the batching
only paying off for the DFE.  Here the accumulated pairs go to the engine as
one batched launch (pqb_perm_batch_states: the interferometer once plus the
state arrays; the engine drops the unoccupied rows and columns itself) in
extended precision for GlynnRep and in double for ChinHuh, as the reference's
calculators.
"""

from __future__ import annotations

import numpy as np

from .. import _capi as C

GlynnRep = 0
ChinHuh = 1
GlynnRepSingleDFE = 2
GlynnRepDualDFE = 3
GlynnRepMultiSingleDFE = 4
GlynnRepMultiDualDFE = 5


class BatchednPermanentCalculator:
    def __init__(self, interferometer_matrix=None):
        self.interferometer_matrix = (None if interferometer_matrix is None
                                      else np.ascontiguousarray(interferometer_matrix, dtype=np.complex128))
        self.input_states = []
        self.output_states = []

    def reserve_space(self, num_of_permanents: int) -> None:
        """:64-71: the containers are sized up front; add() writes by index."""
        n = int(num_of_permanents)
        self.input_states = [None] * n
        self.output_states = [None] * n

    def add(self, input_state, output_state, index: int) -> None:
        """:78-85"""
        self.input_states[index] = np.asarray(input_state, dtype=np.int64)
        self.output_states[index] = np.asarray(output_state, dtype=np.int64)

    def clear(self) -> None:
        self.input_states = []
        self.output_states = []

    def calculate(self, lib: int = 0) -> np.ndarray:
        """Permanents of the accumulated pairs as a 1 x count complex matrix (:109-296)."""
        if self.interferometer_matrix is None or self.interferometer_matrix.size == 0:
            raise Exception("BatchednPermanentCalculator::calculate:  interferometer matrix is zero")
        if lib in (GlynnRepSingleDFE, GlynnRepDualDFE):
            raise Exception("BatchednPermanentCalculator::calculate:  repeated DFE not implemented yet")
        if lib not in (GlynnRep, ChinHuh):
            raise Exception("Wrong value set for permanent library.")
        U = self.interferometer_matrix
        if not self.input_states:
            return np.zeros((1, 0), dtype=np.complex128)
        try:  # one (count, modes) array per side: the engine plans every pair from U itself
            ins = np.stack([np.asarray(s, dtype=np.int64) for s in self.input_states])
            outs = np.stack([np.asarray(s, dtype=np.int64) for s in self.output_states])
        except (ValueError, TypeError):
            raise Exception("BatchednPermanentCalculator::calculate:  every state needs one entry per mode") from None
        algo, prec = (C.ALGO_GLYNN, C.PREC_EXTENDED) if lib == GlynnRep else (C.ALGO_BBFG, C.PREC_DOUBLE)
        return C.permanent_batch_states(U, ins, outs, algo=algo, precision=prec).reshape(1, -1)
