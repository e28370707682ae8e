"""Permanent calculators with the reference's class names (piquassoboost.sampling)."""
