"""Python mirror of the reference's CPython calculator types.

``GlynnPermanentCalculator_wrapper`` follows
piquassoboost/sampling/GlynnPermanentCalculator_Wrapper.hpp:34-504 and
``ChinHuhPermanentCalculator_wrapper`` follows
piquassoboost/sampling/ChinHuhPermanentCalculator_Wrapper.hpp:36-610: same
keyword names, same ``lib`` codes and defaults (7), a ``matrix`` property (and
``input_state`` / ``output_state`` for the repeated form), ``calculate()``
returning a Python ``complex`` -- or nested lists for the list-batched
surfaces -- and ``Exception`` with the reference's message on the reference's
error conditions.  Every value is computed by the HIP engine.

Intentional differences from the reference (each the documented fix of a
reference defect):
* inputs are converted with ``np.ascontiguousarray(x, complex128 / int64)``;
  the reference reinterprets the buffer of a contiguous float64 or complex64
  array as complex128 (GlynnPermanentCalculator_Wrapper.hpp:151-157);
* state lengths must match the matrix shape (the reference reads out of
  bounds otherwise);
* the output-list batching writes each converted array back at its own index
  (the reference writes index ``i`` instead of ``j``,
  ChinHuhPermanentCalculator_Wrapper.hpp:263);
* the MPFR codes (plain lib 1, repeated lib 9) run the extended-precision
  engine path instead of requiring an MPFR build.
"""

from __future__ import annotations

import numpy as np

from .. import _capi as C

WRONG_LIB = "Wrong value set for permanent library."

# plain calculator codes, GlynnPermanentCalculator_Wrapper.hpp:34-44
GlynnLongDouble = 0
GlynnInf = 1
GlynnSingleDFE = 2
GlynnDualDFE = 3
GlynnSingleDFEF = 4
GlynnDualDFEF = 5
GlynnDouble = 6
BBFGPermanentCalculatorDouble = 7
BBFGPermanentCalculatorLongDouble = 8
GlynnSimpleDouble = 10
GlynnSimpleLongDouble = 11

# lib code -> (algorithm, precision) of the engine
PLAIN_LIBS = {
    GlynnLongDouble: (C.ALGO_GLYNN, C.PREC_EXTENDED),
    GlynnInf: (C.ALGO_BBFG, C.PREC_INF),
    GlynnDouble: (C.ALGO_GLYNN, C.PREC_DOUBLE),
    BBFGPermanentCalculatorDouble: (C.ALGO_BBFG, C.PREC_DOUBLE),
    BBFGPermanentCalculatorLongDouble: (C.ALGO_BBFG, C.PREC_EXTENDED),
    GlynnSimpleDouble: (C.ALGO_SIMPLE, C.PREC_DOUBLE),
    GlynnSimpleLongDouble: (C.ALGO_SIMPLE, C.PREC_EXTENDED),
}

# repeated calculator codes, ChinHuhPermanentCalculator_Wrapper.hpp:36-47
ChinHuhDouble = 0
GlynnRep = 1
GlynnRepSingleDFE = 2
GlynnRepDualDFE = 3
GlynnRepMultiSingleDFE = 4
GlynnRepMultiDualDFE = 5
GlynnRepCPUDouble = 6
BBFGPermanentCalculatorRepeatedDouble = 7
BBFGPermanentCalculatorRepeatedLongDouble = 8
GlynnRepInf = 9
GlynnRepSingleDFEF = 10
GlynnRepDualDFEF = 11

REPEATED_LIBS = {
    ChinHuhDouble: (C.ALGO_BBFG, C.PREC_DOUBLE),
    GlynnRep: (C.ALGO_GLYNN, C.PREC_EXTENDED),
    GlynnRepCPUDouble: (C.ALGO_GLYNN, C.PREC_DOUBLE),
    BBFGPermanentCalculatorRepeatedDouble: (C.ALGO_BBFG, C.PREC_DOUBLE),
    BBFGPermanentCalculatorRepeatedLongDouble: (C.ALGO_BBFG, C.PREC_EXTENDED),
    GlynnRepInf: (C.ALGO_BBFG, C.PREC_INF),
}


def _run(fn, *args):
    try:
        return fn(*args)
    except C.EngineError as e:  # the reference raises a bare Exception with the C++ message
        raise Exception(str(e)) from None


def _keep_matrix(m):
    # GlynnPermanentCalculator_Wrapper.hpp:151-157: lists are kept as lists
    if isinstance(m, list):
        return m
    return C.as_matrix(m)


def _keep_state(s):
    if isinstance(s, list):
        return s
    return C.as_state(s)


class GlynnPermanentCalculator_wrapper:
    """Plain permanent calculator (GlynnPermanentCalculator_Wrapper.hpp:133-190)."""

    def __init__(self, matrix=None, lib=BBFGPermanentCalculatorDouble):
        if matrix is None:
            raise TypeError("GlynnPermanentCalculator_wrapper: matrix is required")
        self.lib = int(lib)
        if self.lib not in PLAIN_LIBS:
            raise Exception(WRONG_LIB)
        self._matrix = _keep_matrix(matrix)

    @property
    def matrix(self):
        return self._matrix

    @matrix.setter
    def matrix(self, value):
        self._matrix = _keep_matrix(value)

    def calculate(self):
        """Permanent of ``matrix`` (a list of matrices gives a list of permanents,
        GlynnPermanentCalculator_Wrapper.hpp:204-293)."""
        algo, prec = PLAIN_LIBS[self.lib]
        if isinstance(self._matrix, list):
            # one call for the whole list (the reference loops over the matrices,
            # GlynnPermanentCalculator_Wrapper.hpp:210-237): one launch per shape
            return _run(C.permanent_list, self._matrix, algo, prec)
        return _run(C.permanent, self._matrix, algo, prec)


class ChinHuhPermanentCalculator_wrapper:
    """Repeated-row calculator (ChinHuhPermanentCalculator_Wrapper.hpp:143-216).

    ``output_state`` gives the row multiplicities and ``input_state`` the
    column multiplicities of ``matrix`` (ChinHuhPermanentCalculator_Wrapper.hpp:
    328-369 passes (matrix, input, output) to calculate(mtx, col_mult, row_mult)).
    """

    def __init__(self, lib=BBFGPermanentCalculatorRepeatedDouble, matrix=None, input_state=None,
                 output_state=None):
        if matrix is None or input_state is None or output_state is None:
            raise TypeError("ChinHuhPermanentCalculator_wrapper: matrix, input_state and output_state are required")
        self.lib = int(lib)
        if self.lib not in REPEATED_LIBS:
            raise Exception(WRONG_LIB)
        self._matrix = C.as_matrix(matrix)
        self._input_state = _keep_state(input_state)
        self._output_state = _keep_state(output_state)

    @property
    def matrix(self):
        return self._matrix

    @matrix.setter
    def matrix(self, value):
        self._matrix = C.as_matrix(value)

    @property
    def input_state(self):
        return self._input_state

    @input_state.setter
    def input_state(self, value):
        self._input_state = _keep_state(value)

    @property
    def output_state(self):
        return self._output_state

    @output_state.setter
    def output_state(self, value):
        self._output_state = _keep_state(value)

    def _one(self, ins, outs):
        algo, prec = REPEATED_LIBS[self.lib]
        return _run(C.permanent_repeated, self._matrix, C.as_state(outs), C.as_state(ins), algo, prec)

    def calculate(self):
        ins, outs = self._input_state, self._output_state
        if isinstance(ins, list) and isinstance(outs, list):
            # ChinHuhPermanentCalculator_Wrapper.hpp:233-323.  Without
            # multiInput: ins[i] is one input state, outs[i] a list of output
            # states.  With multiInput (ins[0] is a list): outs[i] is one output
            # state and ins[i] a list of input states.
            multi = len(ins) != 0 and isinstance(ins[0], list)
            pairs = []  # (i, input state, output state)
            for i in range(len(ins)):
                if multi:
                    pairs += [(i, C.as_state(x), C.as_state(outs[i])) for x in ins[i]]
                else:
                    pairs += [(i, C.as_state(ins[i]), C.as_state(y)) for y in outs[i]]
            algo, prec = REPEATED_LIBS[self.lib]
            # all pairs in one batched launch (the reference loops over them,
            # ChinHuhPermanentCalculator_Wrapper.hpp:283-308)
            vals = _run(C.permanent_batch, [(self._matrix, o, x) for _, x, o in pairs], algo, prec)
            res = [[] for _ in range(len(ins))]
            for (i, _, _), v in zip(pairs, vals):
                res[i].append(v)
            return res
        return self._one(ins, outs)
