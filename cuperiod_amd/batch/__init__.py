"""Batched BLS: many light curves per launch sequence, spectra reduced on the device."""

from cuperiod_amd.batch.engine import BatchResult, LinspaceGrid, bls_batch

__all__ = ["BatchResult", "LinspaceGrid", "bls_batch"]
