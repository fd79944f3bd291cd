"""MI355X-native (gfx950, HIP, FP64) qTIP4P/F path-integral timestep.

A drop-in for the hot path of rayseashyn/qtip4pf_water: forces (LJ, Ewald
real + reciprocal on the M-site, intramolecular), Verlet lists, the bead
normal-mode transform, Suzuki-Chin propagation and PILE-G, behind the C ABI of
include/qtip4pf.h (libqtip4pf.so) and mirrored here under the reference's names.
"""
from .api import (Batch, HostRng, Surrogate, Water, initialise_fcc,  # noqa: F401
                  nn_descriptor_params, philox4x32_10, read_positions, write_positions)
from ._lib import ListOverflow, QtpError  # noqa: F401
