"""Permanent calculator classes with the reference's names and ``lib`` codes.

Mirrors piquassoboost/sampling/Boson_Sampling_Utilities.py:30-335 (the
hafnian classes of that module are outside this engine).  Each class fixes the
``lib`` code of its base type exactly as the reference does; the DFE classes
keep their codes and therefore raise "Wrong value set for permanent library."
as a reference build without DFE support does.
"""

from __future__ import annotations

import numpy as np

from .Boson_Sampling_Utilities_wrapper import (
    ChinHuhPermanentCalculator_wrapper,
    GlynnPermanentCalculator_wrapper,
)


class RepeatedPermanentCalculator(ChinHuhPermanentCalculator_wrapper):
    """Permanent of the effective scattering matrix of a boson sampling instance
    (Boson_Sampling_Utilities.py:30-90)."""

    def __init__(self, lib, matrix, input_state, output_state):
        if not (type(input_state) is np.ndarray):
            if not isinstance(input_state, list):
                input_state = np.array(input_state, dtype=np.int64)
            else:
                for item in input_state:
                    if not (type(item) is np.ndarray) and not isinstance(item, list):
                        raise Exception("The input state should be a numpy array, or a list of numpy arrays")
        if not (type(output_state) is np.ndarray):
            if not isinstance(output_state, list):
                output_state = np.array(output_state, dtype=np.int64)
            else:
                for item in output_state:
                    if not (type(item) is np.ndarray) and not isinstance(item, list):
                        raise Exception("The output state should be a numpy array, or a list of numpy arrays")
        super().__init__(lib=lib, matrix=matrix, input_state=input_state, output_state=output_state)

    def calculate(self):
        return super().calculate()


class ChinHuhPermanentCalculator(RepeatedPermanentCalculator):
    def __init__(self, matrix, input_state, output_state):
        super().__init__(0, matrix, input_state, output_state)


class GlynnRepeatedPermanentCalculator(RepeatedPermanentCalculator):
    """BBFG formula with repeated rows/columns, long double in the reference."""

    def __init__(self, matrix, input_state, output_state):
        super().__init__(1, matrix, input_state, output_state)


class GlynnRepeatedInfPermanentCalculator(RepeatedPermanentCalculator):
    def __init__(self, matrix, input_state, output_state):
        super().__init__(9, matrix, input_state, output_state)


class GlynnRepeatedSingleDFEPermanentCalculator(RepeatedPermanentCalculator):
    def __init__(self, matrix, input_state, output_state):
        super().__init__(2, matrix, input_state, output_state)


class GlynnRepeatedDualDFEPermanentCalculator(RepeatedPermanentCalculator):
    def __init__(self, matrix, input_state, output_state):
        super().__init__(3, matrix, input_state, output_state)


class GlynnRepeatedMultiSingleDFEPermanentCalculator(RepeatedPermanentCalculator):
    def __init__(self, matrix, input_state, output_state):
        super().__init__(4, matrix, input_state, output_state)


class GlynnRepeatedMultiDualDFEPermanentCalculator(RepeatedPermanentCalculator):
    def __init__(self, matrix, input_state, output_state):
        super().__init__(5, matrix, input_state, output_state)


class GlynnRepeatedPermanentCalculatorDouble(RepeatedPermanentCalculator):
    def __init__(self, matrix, input_state, output_state):
        super().__init__(6, matrix, input_state, output_state)


class BBFGRepeatedPermanentCalculatorDouble(RepeatedPermanentCalculator):
    def __init__(self, matrix, input_state, output_state):
        super().__init__(7, matrix, input_state, output_state)


class BBFGRepeatedPermanentCalculatorLongDouble(RepeatedPermanentCalculator):
    def __init__(self, matrix, input_state, output_state):
        super().__init__(8, matrix, input_state, output_state)


class GlynnPermanent(GlynnPermanentCalculator_wrapper):
    def __init__(self, matrix):
        super().__init__(matrix, 0)


class GlynnPermanentInf(GlynnPermanentCalculator_wrapper):
    def __init__(self, matrix):
        super().__init__(matrix, 1)


class GlynnPermanentSingleDFE(GlynnPermanentCalculator_wrapper):
    def __init__(self, matrix):
        super().__init__(matrix, 2)


class GlynnPermanentDualDFE(GlynnPermanentCalculator_wrapper):
    def __init__(self, matrix):
        super().__init__(matrix, 3)


class GlynnPermanentDouble(GlynnPermanentCalculator_wrapper):
    def __init__(self, matrix):
        super().__init__(matrix, 6)


class BBFGPermanentDouble(GlynnPermanentCalculator_wrapper):
    def __init__(self, matrix):
        super().__init__(matrix, 7)


class BBFGPermanentLongDouble(GlynnPermanentCalculator_wrapper):
    def __init__(self, matrix):
        super().__init__(matrix, 8)


class GlynnPermanentSimpleDouble(GlynnPermanentCalculator_wrapper):
    def __init__(self, matrix):
        super().__init__(matrix, 10)


class GlynnPermanentSimpleLongDouble(GlynnPermanentCalculator_wrapper):
    def __init__(self, matrix):
        super().__init__(matrix, 11)
