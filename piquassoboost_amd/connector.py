"""Permanent hook for piquasso (piquassoboost/connector.py:23-39).

``cpp_permanent_function(matrix, input, output)`` is what BoostConnector
installs as ``self.permanent``; piquasso itself is not a dependency here, so
only the function is provided."""

from __future__ import annotations

from .sampling.Boson_Sampling_Utilities import BBFGRepeatedPermanentCalculatorDouble


def cpp_permanent_function(matrix, input, output):
    result = BBFGRepeatedPermanentCalculatorDouble(matrix, input, output).calculate()
    if isinstance(result, list):
        # connector.py:27-29: an empty list result stands for 1.0
        result = 1.0
    return result
