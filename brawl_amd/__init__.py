"""brawl_amd -- MI355X (gfx950) implementation of BraWl's lattice-sampling hot
path: the Bragg-Williams energy and pair-swap dE, the Metropolis atom-swap
sweep, and the Wang-Landau / nested-sampling walker moves.

The compute runs in hand-written HIP kernels behind the C-ABI of
include/brawl_hip.h (libbrawl_hip.so, built in-tree).  This package is the
host-side mirror of the reference's procedure interface over that ABI.
"""
from .lattice import (BCC, FCC, K_B_IN_RY, RY_TO_EV, AnnealResult, BrawlError, Lattice,
                      anneal_batch, beta_of, warren_cowley)

__all__ = ["Lattice", "BrawlError", "AnnealResult", "anneal_batch", "beta_of", "warren_cowley",
           "FCC", "BCC", "K_B_IN_RY", "RY_TO_EV"]
