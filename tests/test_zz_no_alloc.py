"""edl_hxv_noalloc through the C-ABI (GPU): the allocation-free operator
(intermediate in the library's file-scope device scratch) against edl_hxv
on the same device buffers and against the oracle, in fp64 and fp32, two
calls in a row with the second on a smaller sector; refusals of an unbuilt
sector and of a sector whose intermediate exceeds the scratch."""
import ctypes

import numpy as np
import pytest

from ed_lattice_amd import EDLattice, lattice
from ed_lattice_amd.solver import PREC

pytestmark = pytest.mark.gpu


def _relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_noalloc_equals_edl_hxv_and_oracle(orc):
    H, U = lattice.square(3, 4), [6.0]
    m = orc.Model(H, U)
    with EDLattice(0) as ed:
        ed.set_hamiltonian(H, U)
        for prec, dt, tol in (("fp64", np.float64, 1e-13), ("fp32", np.float32, 2e-6)):
            # (6,6): dim 853 776, then the smaller (7,6): dim 731 808
            for sec in ((6, 6), (7, 6)):
                isec = ed.sector_index(*sec)
                dim, _, _ = ed.build_sector(isec)
                print(f"noalloc {prec} sector {sec} dim {dim}", flush=True)
                nb = dim * np.dtype(dt).itemsize
                v, h1, h2 = ed.alloc(nb), ed.alloc(nb), ed.alloc(nb)
                ed.vec_random(prec, v, dim, 31 + sec[0])
                ed.hxv_device(isec, prec, v, h1)
                rc = ed.lib.edl_hxv_noalloc(ed.h, isec, PREC[prec], ctypes.c_void_p(v), ctypes.c_void_p(h2))
                assert rc == 0, ed.lib.edl_last_error(ed.h).decode()
                a = ed.copy_to_host(h1, np.empty(dim, dt)).astype(np.float64)
                b = ed.copy_to_host(h2, np.empty(dim, dt)).astype(np.float64)
                # same kernels and schedule as edl_hxv: the same numbers
                assert _relerr(b, a) <= tol
                vh = ed.copy_to_host(v, np.empty(dim, dt)).astype(np.float64)
                ref = orc.matvec(m, *sec, vh.astype(np.complex128), nthreads=8).real
                assert _relerr(b, ref) < tol
        # a sector that is not built yet is refused, not built on the fly
        isec = ed.sector_index(5, 5)
        p = ed.alloc(8)
        assert ed.lib.edl_hxv_noalloc(ed.h, isec, PREC["fp64"], ctypes.c_void_p(p), ctypes.c_void_p(p)) != 0
        assert "not built" in ed.lib.edl_last_error(ed.h).decode()
    # 4x4 (8,8): an fp32 intermediate of 663 MB exceeds the 256 MiB scratch
    with EDLattice(0) as ed:
        ed.set_hamiltonian(lattice.square(4, 4), [8.0])
        isec = ed.sector_index(8, 8)
        ed.build_sector(isec)
        p = ed.alloc(8)
        assert ed.lib.edl_hxv_noalloc(ed.h, isec, PREC["fp32"], ctypes.c_void_p(p), ctypes.c_void_p(p)) != 0
        assert "exceeds" in ed.lib.edl_last_error(ed.h).decode()
