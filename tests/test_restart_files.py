"""The restart files the Fortran host writes (write_gfmatrix_all3 /
write_gfmatrix_all2 layouts, VARS_GLOBAL.f90:440-452,587-606), read back on
the CPU with the record layout of read_formatted_gfmatrix (:611-632): the
fixtures tests/golden/ring6_*.restart come from the product's host on the
6-site ring, U = 3, sector (3,3), 20 Lanczos steps per seed
(tests/golden/make_restart_fixtures.py).  Every record is checked against
the oracle: the poles and weights of a 20-step FP64 Lanczos from the dense
ground state's seeds (add_to_lanczos_gf_normal, GF_ELECTRONS.f90:417-457;
add_to_lanczos_spinChi, CHI_SPIN_ELECTRONS.f90:305-366)."""
import os

import numpy as np
import pytest

from ed_lattice_amd import lattice
from fortran_records import FortranRecords, read_gfmatrix

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NS, NLANC = 6, 20


@pytest.fixture(scope="module")
def ring6(orc):
    m = orc.Model(lattice.ring(NS), [3.0])
    w, V = np.linalg.eigh(orc.hamiltonian_csr(m, 3, 3).toarray().real)
    return m, w[0], V[:, 0]


def _poles(orc, m, e0, seed, sector, isign):
    a, b, n = orc.lanczos_tridiag(orc.hamiltonian_csr(m, *sector), seed, NLANC)
    assert n == NLANC
    p, w = orc.gf_poles(a, b, seed @ seed, e0, isign)
    k = np.argsort(p)
    return p[k], w[k]


def _same(got, ref):
    """Poles to 1e-7 (the GPU chains start from a Lanczos ground state, the
    oracle's from the dense one: a 20-step Ritz value of small weight moves
    by up to 4e-8), weights to 1e-9."""
    (pg, wg), (pr, wr) = got, ref
    k = np.argsort(pg)
    np.testing.assert_allclose(pg[k], pr, rtol=0, atol=1e-7)
    np.testing.assert_allclose(wg[k], wr, rtol=0, atol=1e-9)


def test_gfmatrix_restart_records(orc, ring6):
    m, e0, gs = ring6
    rd = FortranRecords(os.path.join(GOLDEN, "ring6_gfmatrix.restart"))
    # write_gfmatrix_all3: Nspin x Ns x Ns records, i slowest after the spin
    rec = [[read_gfmatrix(rd) for _ in range(NS)] for _ in range(NS)]
    assert rd.exhausted()
    seeds = {}
    for i in range(1, NS + 1):
        seeds[i] = (orc.seed_c_cdg(NS, 2, i, 1, (3, 3), (4, 3), gs).real,
                    orc.seed_c_cdg(NS, 1, i, 1, (3, 3), (2, 3), gs).real)
    for i in range(NS):
        for j in range(NS):
            r = rec[i][j]
            assert r is not None and len(r) == 1 and len(r[0]) == 2, (i, j)
            for ch in range(2):
                assert r[0][ch][0].size == NLANC
                np.testing.assert_array_equal(r[0][ch][0], rec[j][i][0][ch][0])
            sp = seeds[i + 1][0] + (seeds[j + 1][0] if j != i else 0.0)
            sh = seeds[i + 1][1] + (seeds[j + 1][1] if j != i else 0.0)
            # channel 1: particle, isign +1 in (4,3); channel 2: hole, isign -1 in (2,3)
            _same(r[0][0], _poles(orc, m, e0, sp, (4, 3), 1))
            _same(r[0][1], _poles(orc, m, e0, sh, (2, 3), -1))


def test_chispin_restart_records(orc, ring6):
    m, e0, gs = ring6
    rd = FortranRecords(os.path.join(GOLDEN, "ring6_ChiSpinMatrix.restart"))
    # write_gfmatrix_all2 over (Ns, Ns): F where nothing was built (off the diagonal)
    for i in range(NS):
        for j in range(NS):
            r = read_gfmatrix(rd)
            if i != j:
                assert r is None
                continue
            assert r is not None and len(r) == 1 and len(r[0]) == 1
            s = orc.seed_sz(NS, i + 1, (3, 3), gs).real
            _same(r[0][0], _poles(orc, m, e0, s, (3, 3), 1))
    assert rd.exhausted()
