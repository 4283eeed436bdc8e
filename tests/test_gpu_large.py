"""The 3x6 open cluster at full size (GPU).

18 sites, Nup = Ndw = 9: Dim = 48620^2 = 2 363 904 400 > 2^31 - 1, beyond the
reference's default integer, so no reference run exists for it.  Checked
through properties that do not depend on size: <u,Hv> = <Hu,v> on random
vectors, and sampled rows of the fp32 product against the oracle's 64-bit
row evaluator (the restated per-element formulas)."""
import numpy as np
import pytest

from ed_lattice_amd import EDLattice, lattice

pytestmark = pytest.mark.gpu


def test_obc3x6_full_size(orc):
    H = lattice.disordered_rect(3, 6)
    with EDLattice(0) as ed:
        ed.set_hamiltonian(H, [4.0])
        isec = ed.sector_index(9, 9)
        dim, du, dd = ed.build_sector(isec)
        assert (dim, du, dd) == (2363904400, 48620, 48620)
        nb = dim * 4
        u, v, hu, hv = (ed.alloc(nb) for _ in range(4))
        ed.vec_random("fp32", u, dim, 11)
        ed.vec_random("fp32", v, dim, 12)
        ms = ed.hxv_device(isec, "fp32", v, hv)
        ed.hxv_device(isec, "fp32", u, hu)
        print(f"obc3x6 fp32 H x v: {ms:.1f} ms", flush=True)
        a = ed.vec_dot("fp32", u, hv, dim)
        b = ed.vec_dot("fp32", hu, v, dim)
        nu, nv = ed.vec_dot("fp32", u, u, dim), ed.vec_dot("fp32", v, v, dim)
        nhv = ed.vec_dot("fp32", hv, hv, dim)
        assert abs(a - b) < 1e-5 * np.sqrt(nu * nhv)
        # sampled rows against the oracle
        vh = ed.copy_to_host(v, np.empty(dim, np.float32))
        hvh = ed.copy_to_host(hv, np.empty(dim, np.float32))
        rng = np.random.default_rng(5)
        rows = np.concatenate([[0, dim - 1, 2**31 - 1, 2**31 + 12345], rng.integers(0, dim, 200)])
        m = orc.Model(H, [4.0])
        ref = orc.hxv_rows_real64(m, 9, 9, rows, vh)
        scale = np.max(np.abs(hvh[rows]))
        assert np.max(np.abs(hvh[rows] - ref)) < 5e-6 * scale


def test_obc3x6_chain_layout_first_step():
    """The chain's own layout (padded column stride: segmented up pass and the
    down pass the tuner picks for it, Lanczos terms fused) against the
    natural-layout operator at full size: for the start vector v (the fp64
    random vector of the same seed), alpha_0 = <v,Hv>/<v,v> and
    beta_1^2 = <Hv,Hv>/<v,v> - alpha_0^2."""
    H = lattice.disordered_rect(3, 6)
    with EDLattice(0) as ed:
        ed.set_hamiltonian(H, [4.0])
        isec = ed.sector_index(9, 9)
        dim, _, _ = ed.build_sector(isec)
        nb = dim * 8
        v, hv = ed.alloc(nb), ed.alloc(nb)
        ed.vec_random("fp64", v, dim, 0x5EED)
        ed.hxv_device(isec, "fp64", v, hv)
        vv, vhv, hh = ed.vec_dot("fp64", v, v, dim), ed.vec_dot("fp64", v, hv, dim), ed.vec_dot("fp64", hv, hv, dim)
        a0 = vhv / vv
        b1 = np.sqrt(hh / vv - a0 * a0)
        r = ed.lanczos_eigh(isec, "fp64", niter=3, seed=0x5EED)
        print(f"obc3x6 chain alpha0 {r['alpha'][0]:.15f} ref {a0:.15f} beta1 {r['beta'][1]:.15f} ref {b1:.15f}",
              flush=True)
        assert abs(r["alpha"][0] - a0) < 1e-11 * np.sqrt(hh / vv)
        assert abs(r["beta"][1] - b1) < 1e-9 * b1
