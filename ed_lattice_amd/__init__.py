"""MI355X engine for the on-the-fly H x v + Lanczos path of aamaricci/ED_Lattice.

The compute lives in libedlat_hip.so (HIP, gfx950) behind the C-ABI of
include/edlat_gpu.h; ``solver.EDLattice`` is the host binding, ``lattice``
builds the hopping matrices of the benchmark lattices.
"""
from . import lattice  # noqa: F401
from ._lib import LIB_PATH, load  # noqa: F401
from .solver import EDError, EDLattice, eval_gf, lanczos_poles, optical_conductivity, sector_index  # noqa: F401
