"""piquassoboost.sampling.BosonSamplingSimulator (BosonSamplingSimulator.py:24-30)."""

from __future__ import annotations


class BosonSamplingSimulator:
    def __init__(self, simulation_strategy) -> None:
        self.simulation_strategy = simulation_strategy

    def get_classical_simulation_results(self, input_state, samples_number: int = 1):
        return self.simulation_strategy.simulate(input_state, samples_number)
