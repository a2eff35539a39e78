"""Periodogram methods: the registry and the BLS method on the HIP engine."""

from cuperiod_amd.methods import bls as _bls  # noqa: F401  (registers "BLS")
