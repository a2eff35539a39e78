"""Learned candidate proposer: a 1-D convolutional network on BF16 MFMA that nominates
frequency windows for a fine BLS search, with the plain full-grid search as fallback."""

from cuperiod_amd.learned.network import ProposerNet, init_parameters, load_parameters
from cuperiod_amd.learned.proposer import LearnedProposer, LearnedResult

__all__ = ["LearnedProposer", "LearnedResult", "ProposerNet", "init_parameters",
           "load_parameters"]
