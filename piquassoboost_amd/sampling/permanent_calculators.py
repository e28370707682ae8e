"""Module-level permanent functions of the reference's ``permanent_calculators``
extension (piquassoboost/sampling/permanent_calculators_wrapper.cpp:28-190)."""

from __future__ import annotations

from .. import _capi as C
from .Boson_Sampling_Utilities_wrapper import _run


def permanent_CPU_repeated_double(matrix, input_state, output_state):
    """GlynnPermanentCalculatorRepeatedDouble semantics (…wrapper.cpp:101-168)."""
    return _run(C.permanent_repeated, C.as_matrix(matrix), C.as_state(output_state), C.as_state(input_state),
                C.ALGO_GLYNN, C.PREC_DOUBLE)


def permanent_CPU_repeated_long_double(matrix, input_state, output_state):
    """GlynnPermanentCalculatorRepeatedLongDouble semantics (…wrapper.cpp:28-95)."""
    return _run(C.permanent_repeated, C.as_matrix(matrix), C.as_state(output_state), C.as_state(input_state),
                C.ALGO_GLYNN, C.PREC_EXTENDED)
