"""The committed workload fixtures (tests/golden, written by
tests/golden/make_fixtures.py): the Hij input files equal the lattice
builders bit for bit; the oracle reproduces the expected outputs; the GPU
engine matches them (ground-state energies, observables, leading Lanczos
coefficients of GF seeds)."""
import json
import os

import numpy as np
import pytest

from ed_lattice_amd import lattice

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name,make,Ns", [
    ("ring8", lambda: lattice.ring(8), 8), ("sq4x4", lambda: lattice.square(4, 4), 16),
    ("rect3x4", lambda: lattice.square(3, 4), 12), ("obc3x6_disorder", lambda: lattice.disordered_rect(3, 6), 18)])
def test_hij_fixtures_equal_builders(name, make, Ns):
    H = lattice.read_hij(os.path.join(GOLDEN, f"hij_{name}.dat"), Ns)
    np.testing.assert_array_equal(H, make())


def test_oracle_reproduces_ring8(orc):
    g = _golden()["ring8"]
    m = orc.Model(lattice.ring(8), [g["U"]])
    w, Z = np.linalg.eigh(orc.hamiltonian_csr(m, 4, 4).toarray())
    assert abs(w[0] - g["e0"]) < 1e-12 * abs(g["e0"])
    assert w[1] - w[0] > 0.1  # a non-degenerate ground state: its observables are well defined
    np.testing.assert_allclose(orc.observables(8, 4, 4, Z[:, 0])["docc"], g["docc"], atol=1e-12)
    for op, d in (("cdg", 1), ("c", -1)):
        s = orc.seed_c_cdg(8, 2 if op == "cdg" else 1, 1, 1, (4, 4), (4 + d, 4), Z[:, 0]).real
        a, b, _ = orc.lanczos_tridiag(orc.hamiltonian_csr(m, 4 + d, 4), s, 20)
        ref = g["seeds"][op]
        assert abs(s @ s - ref["norm2"]) < 1e-12
        np.testing.assert_allclose(a[:10], ref["alpha"][:10], rtol=0, atol=1e-9)
        np.testing.assert_allclose(b[:10], ref["beta"][:10], rtol=0, atol=1e-9)


@pytest.mark.gpu
def test_gpu_matches_golden():
    from ed_lattice_amd import EDLattice

    with EDLattice(0) as ed:
        _gpu_checks(ed, _golden())


def _gpu_checks(ed, g):
    r8 = g["ring8"]
    ed.set_hamiltonian(lattice.read_hij(os.path.join(GOLDEN, "hij_ring8.dat"), 8), [r8["U"]])
    isec = ed.sector_index(4, 4)
    # mixed first: the seeds below are built from the state installed last,
    # and the FP64 run's eigenvector is the converged one (the mixed run's
    # refinement stops on the energy, its vector is good to ~1e-7)
    for prec, tol in (("mixed", 1e-12), ("fp64", 1e-12)):
        r = ed.lanczos_eigh(isec, prec, niter=200, want_vector=True)
        print(f"golden ring8 {prec}: E0 {r['e0']!r} vs {r8['e0']!r}", flush=True)
        assert abs(r["e0"] - r8["e0"]) < tol * abs(r8["e0"])
        np.testing.assert_allclose(ed.observables()["docc"], r8["docc"], atol=1e-9)
    rb = ed.tridiag_batched([("cdg", 1, 1, 0), ("c", 1, 1, 0)], 20, "fp64")
    for k, op in enumerate(("cdg", "c")):
        ref = r8["seeds"][op]
        assert abs(rb["norm2"][k] - ref["norm2"]) < 1e-10
        np.testing.assert_allclose(rb["alpha"][k][:10], ref["alpha"][:10], rtol=0, atol=1e-8)
        np.testing.assert_allclose(rb["beta"][k][:10], ref["beta"][:10], rtol=0, atol=1e-8)
    r34 = g["rect3x4"]
    ed.set_hamiltonian(lattice.read_hij(os.path.join(GOLDEN, "hij_rect3x4.dat"), 12), [r34["U"]])
    r = ed.lanczos_eigh(ed.sector_index(6, 6), "fp64", niter=300, want_vector=False)
    print(f"golden rect3x4: E0 {r['e0']!r} vs {r34['e0_lanczos300']!r}", flush=True)
    assert abs(r["e0"] - r34["e0_lanczos300"]) < 1e-10 * abs(r34["e0_lanczos300"])
