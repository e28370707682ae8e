"""Boson-sampling simulation strategies (piquassoboost.sampling.simulation_strategies)."""
