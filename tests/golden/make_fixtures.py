"""Writes the workload fixtures of tests/golden (test infrastructure).

Inputs, in the reference's Hij_read format (GRAPH_MATRIX.f90:280-301, one
link `i j a b spin re im` per line): the 8-site ring, the 4x4 and 3x4
periodic squares and the 3x6 open cluster with SplitMix64 disorder (seed
20261202: bonds U[0.8, 1.2] in bond order, then onsite U[-0.5, 0.5]).

Expected outputs (golden.json), from the oracle's restatement of the
reference, each computed by a route independent of the engine:
  ring8 (4,4), U = 4: E0 and the site double occupancies of the dense
    ground state (scipy eigh of the oracle's CSR matrix), and the first 20
    Lanczos coefficients of the c+_1 / c_1 spin-up seeds of that state
    (three-term recurrence on the oracle's matrices of the N+-1 sectors);
  rect3x4 (6,6), U = 6: E0 of a 300-step FP64 Lanczos from the SplitMix64
    start vector (seed 0x5EED) over the oracle's matrix.
Run from the repository root:  python3 tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from ed_lattice_amd import lattice  # noqa: E402
from oracle import oracle as orc  # noqa: E402

WORKLOADS = {
    "ring8": lambda: lattice.ring(8),
    "sq4x4": lambda: lattice.square(4, 4),
    "rect3x4": lambda: lattice.square(3, 4),
    "obc3x6_disorder": lambda: lattice.disordered_rect(3, 6),
}


def main():
    for name, make in WORKLOADS.items():
        lattice.write_hij(os.path.join(HERE, f"hij_{name}.dat"), make())
    golden = {}
    # ring8 (4,4), U = 4: dense ground state, observables, seed coefficients
    H = lattice.ring(8)
    m = orc.Model(H, [4.0])
    Hc = orc.hamiltonian_csr(m, 4, 4).toarray()
    w, Z = np.linalg.eigh(Hc)
    gs = Z[:, 0]
    obs = orc.observables(8, 4, 4, gs)
    ring = {"U": 4.0, "sector": [4, 4], "e0": float(w[0]), "e1": float(w[1]),
            "docc": [float(x) for x in obs["docc"]], "seeds": {}}
    for op, d in (("cdg", 1), ("c", -1)):
        s = orc.seed_c_cdg(8, 2 if op == "cdg" else 1, 1, 1, (4, 4), (4 + d, 4), gs).real
        a, b, n = orc.lanczos_tridiag(orc.hamiltonian_csr(m, 4 + d, 4), s, 20)
        ring["seeds"][op] = {"norm2": float(s @ s), "alpha": [float(x) for x in a],
                             "beta": [float(x) for x in b], "nsteps": int(n)}
    golden["ring8"] = ring
    # rect3x4 (6,6), U = 6: FP64 Lanczos from the seeded start vector
    m = orc.Model(lattice.square(3, 4), [6.0])
    Hc = orc.hamiltonian_csr(m, 6, 6)
    e0, _, _ = orc.lanczos_ground_state(Hc, lattice.splitmix_vector(Hc.shape[0], 0x5EED), 300, want_vector=False)
    golden["rect3x4"] = {"U": 6.0, "sector": [6, 6], "e0_lanczos300": float(e0)}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(golden, f, indent=1)


if __name__ == "__main__":
    main()
