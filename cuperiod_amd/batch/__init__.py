"""Batched BLS: many light curves per launch sequence, spectra reduced on the device."""

from cuperiod_amd.batch.engine import BatchResult, LinspaceGrid, bls_batch
from cuperiod_amd.batch.io import periodogram_to_row
from cuperiod_amd.batch.runner import BatchSummary, batch_periodograms

__all__ = ["BatchResult", "BatchSummary", "LinspaceGrid", "batch_periodograms", "bls_batch",
           "periodogram_to_row"]
