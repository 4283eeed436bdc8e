"""Parity of the HIP engine against the oracle (GPU).

Integer work (maps, ranks, hop tables) must be bit-exact; FP64 H x v within
1e-13 of the oracle relative to max|Hv|; FP32 H x v within 2e-6; FP64 E0
within 1e-10 relative; GF/chi continued fractions within 1e-8 on a
frequency grid."""
import numpy as np
import pytest

from ed_lattice_amd import EDLattice, eval_gf, lattice

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(2024)


@pytest.fixture(scope="module")
def ed():
    e = EDLattice(0)
    yield e
    e.close()


def _use(ed, orc, H, U, **kw):
    ed.set_hamiltonian(H, U, **kw)
    return orc.Model(H, U, **kw)


def _relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("Ns,nup,ndw", [(8, 4, 4), (12, 6, 6), (12, 7, 6), (12, 5, 6), (16, 8, 8), (18, 9, 9),
                                        (6, 0, 6), (5, 5, 0)])
def test_maps_bit_exact(ed, orc, Ns, nup, ndw):
    print(f"maps Ns={Ns} ({nup},{ndw})", flush=True)
    ed.set_hamiltonian(lattice.ring(Ns), [1.0])
    mu, md = ed.sector_maps(ed.sector_index(nup, ndw))
    assert np.array_equal(mu, orc.build_map(Ns, nup))
    assert np.array_equal(md, orc.build_map(Ns, ndw))


def test_rank_equals_binary_search(ed, orc):
    Ns = 12
    ed.set_hamiltonian(lattice.square(3, 4), [6.0])
    isec = ed.sector_index(7, 5)
    states = RNG.integers(0, 1 << Ns, 5000).astype(np.int32)
    mu = orc.build_map(Ns, 7)
    states[:100] = mu[RNG.integers(0, mu.size, 100)]
    pos = ed.sector_rank(isec, 1, states)
    ref = np.array([orc.binary_search(mu, int(s)) for s in states])
    assert np.array_equal(pos, ref)


@pytest.mark.parametrize("H,Ns,N", [(lattice.ring(8), 8, 4), (lattice.square(3, 4), 12, 7),
                                    (lattice.disordered_rect(2, 4, seed=11), 8, 3)])
def test_hop_tables_reproduce_c_cdg(ed, orc, H, Ns, N):
    """Every table entry of output string s is a source string s' with
    c^dg_io c_jo s' = sg1*sg2 s (c/cdg of AUX_FUNX.f90) and coefficient
    Hij(io,jo)*sg1*sg2; and every such source is listed."""
    ed.set_hamiltonian(H, [2.0])
    isec = ed.sector_index(N, N)
    nbr, coef = ed.sector_hops(isec, 1)
    m = orc.build_map(Ns, N)
    for a, s in enumerate(m):
        got = {}
        for k in range(nbr.shape[0]):
            if coef[k, a] != 0.0:
                got[int(nbr[k, a])] = coef[k, a]
        want = {}
        for io in range(1, Ns + 1):
            for jo in range(1, Ns + 1):
                if io == jo or H[0, io - 1, jo - 1] == 0.0:
                    continue
                src = s ^ (1 << (io - 1)) ^ (1 << (jo - 1))
                if not ((s >> (io - 1)) & 1) or ((s >> (jo - 1)) & 1):
                    continue
                k1, g1 = orc.c_op(jo, int(src))
                k2, g2 = orc.cdg_op(io, k1)
                assert k2 == s
                want[orc.binary_search(m, int(src)) - 1] = H[0, io - 1, jo - 1] * g1 * g2
        assert got.keys() == want.keys()
        for j in want:
            assert got[j] == want[j]


CASES = [
    ("ring8_U4", lambda: (lattice.ring(8), [4.0], {}), (4, 4)),
    ("rect3x4_U6_76", lambda: (lattice.square(3, 4), [6.0], {}), (7, 6)),
    ("rect2x4_dis_mu", lambda: (lattice.disordered_rect(2, 4, seed=3), [3.5], dict(xmu=0.4)), (3, 5)),
    ("ring7_nohf_fill", lambda: (lattice.ring(7), [2.0], dict(hfmode=False, xmu=0.3, ed_filling=7)), (4, 3)),
    ("nspin2", lambda: (np.stack([lattice.disordered_rect(2, 3, seed=5)[0], lattice.disordered_rect(2, 3, seed=6)[0]]),
                        [2.5], {}), (2, 4)),
    ("two_orbital", lambda: (_two_orb(), [3.0, 2.0], dict(Nsites=[2, 2], Ust=1.5, Jh=0.4, xmu=0.1)), (2, 2)),
    # spin exchange and pair hopping (HxV_non_local.f90), sectors with and
    # without the moves available to every state
    ("two_orbital_kanamori", lambda: (_two_orb(), [3.0, 2.0], dict(Nsites=[2, 2], Ust=1.5, Jh=0.4, Jx=0.4, Jp=0.35)),
     (2, 2)),
    ("three_orbital_kanamori", lambda: (_three_orb(), [4.0, 3.0, 2.5],
                                        dict(Nsites=[2, 2, 2], Ust=2.0, Jh=0.6, Jx=0.6, Jp=0.5, xmu=0.2)), (3, 2)),
    # DimUp = C(17,8) = 24310: an fp64 column (194 KB) exceeds LDS, so fp64
    # takes the segmented up pass (two row segments split by site 17) and
    # fp32 the two-pass kernels
    ("ring17_wide_columns", lambda: (lattice.ring(17), [3.0], {}), (8, 1)),
]


def _three_orb():
    # 3 orbitals x 2 sites: intra-orbital hops between the sites, weak
    # inter-orbital hybridisation on a site, crystal field
    Ns = 6
    H = np.zeros((1, Ns, Ns))
    pairs = [(0, 1, -1.0), (2, 3, -0.8), (4, 5, -0.6), (0, 2, -0.2), (1, 3, -0.2), (2, 4, -0.1), (3, 5, -0.1)]
    for a, b, t in pairs:
        H[0, a, b] = H[0, b, a] = t
    for k, e in enumerate([0.3, 0.3, 0.0, 0.0, -0.2, -0.2]):
        H[0, k, k] = e
    return H


def _two_orb_ring4():
    # 2 orbitals x 4 sites: rings of 4 per orbital (orbital 2 is sites 5-8),
    # on-site hybridisation, crystal field
    Ns = 8
    H = np.zeros((1, Ns, Ns))
    for a in range(4):
        b = (a + 1) % 4
        H[0, a, b] = H[0, b, a] = -1.0
        H[0, 4 + a, 4 + b] = H[0, 4 + b, 4 + a] = -0.7
        H[0, a, 4 + a] = H[0, 4 + a, a] = -0.2
    for k in range(4):
        H[0, k, k], H[0, 4 + k, 4 + k] = 0.15, -0.15
    return H


def test_kanamori_on_two_pass_chains_and_batched(ed, orc, monkeypatch):
    """The non-local Kanamori block (Jx, Jp; HxV_non_local.f90) as the third
    term of the two-pass operator (T += N q between the passes), in the
    ground-state chains and in both batched seed forms: 2 orbitals x 4
    sites, sector (4,4), dim 4900, against the oracle and dense eigh."""
    H = _two_orb_ring4()
    U = [3.0, 2.5]
    kw = dict(Nsites=[4, 4], Ust=1.5, Jh=0.5, Jx=0.5, Jp=0.4, xmu=0.1)
    m = _use(ed, orc, H, U, **kw)
    isec = ed.sector_index(4, 4)
    dim, _, _ = ed.build_sector(isec)
    for prec in ("fp64", "fp32"):
        assert ed.bench_passes(isec, prec, 1)[2]["two_pass"] == 1
    for _ in range(2):
        v = RNG.standard_normal(dim)
        ref = orc.matvec(m, 4, 4, v, nthreads=8).real
        assert _relerr(ed.direct_matvec(isec, v, "fp64"), ref) < 1e-13
        assert _relerr(ed.direct_matvec(isec, v, "fp32"), ref) < 2e-6
    Hd = orc.hamiltonian_csr(m, 4, 4).toarray().real
    w, V = np.linalg.eigh(Hd)
    assert w[1] - w[0] > 1e-6
    r = ed.lanczos_eigh(isec, "mixed", niter=300, want_vector=True)
    assert abs(r["e0"] - w[0]) < 1e-10 * abs(w[0])
    assert 1 - abs(r["gs"] @ V[:, 0]) < 1e-9
    seeds = [("cdg", 1, 1, 0), ("c", 1, 5, 0), ("c", 1, 2, 6), ("sz", 1, 3, 0)]
    targets = [(5, 4), (3, 4), (3, 4), (4, 4)]
    refs = []
    for (op, ispin, io, jo), tg in zip(seeds, targets):
        if op == "sz":
            s = orc.seed_sz(8, io, (4, 4), r["gs"]).real
        else:
            opc = 2 if op == "cdg" else 1
            s = orc.seed_c_cdg(8, opc, io, 1, (4, 4), tg, r["gs"]).real
            if jo:
                s = s + orc.seed_c_cdg(8, opc, jo, 1, (4, 4), tg, r["gs"]).real
        a, b, _ = orc.lanczos_tridiag(orc.hamiltonian_csr(m, *tg), s, 100)
        refs.append((s @ s, a, b))
    for form in ("tiles", "persist"):
        monkeypatch.setenv("EDL_BATCHED_KERNELS", form)
        rb = ed.tridiag_batched(seeds, 100, "fp64")
        print(f"kanamori seeds, {form}: {rb['ms']:.2f} ms", flush=True)
        for k, (n2, a, b) in enumerate(refs):
            assert abs(rb["norm2"][k] - n2) < 1e-11
            _assert_leading(rb["alpha"][k], rb["beta"][k], a, b, 10, 1e-9)


def _two_orb():
    H = np.zeros((1, 4, 4))
    for a, b, t in [(0, 1, -1.0), (2, 3, -0.6), (0, 2, -0.3), (1, 3, -0.3)]:
        H[0, a, b] = H[0, b, a] = t
    H[0, 0, 0], H[0, 3, 3] = 0.2, -0.1
    return H


@pytest.mark.parametrize("name,mk,sec", CASES, ids=[c[0] for c in CASES])
def test_hxv_fp64_and_fp32_vs_oracle(ed, orc, name, mk, sec):
    print(f"hxv case {name} sector {sec}", flush=True)
    H, U, kw = mk()
    m = _use(ed, orc, H, U, **kw)
    isec = ed.sector_index(*sec)
    dim, _, _ = ed.build_sector(isec)
    for _ in range(3):
        v = RNG.standard_normal(dim)
        ref = orc.matvec(m, *sec, v, nthreads=8).real
        assert _relerr(ed.direct_matvec(isec, v, "fp64"), ref) < 1e-13
        assert _relerr(ed.direct_matvec(isec, v, "fp32"), ref) < 2e-6
    vc = RNG.standard_normal(dim) + 1j * RNG.standard_normal(dim)
    refc = orc.matvec(m, *sec, vc, nthreads=8)
    assert _relerr(ed.direct_matvec(isec, vc, "fp64"), refc) < 1e-13


@pytest.mark.parametrize("prec,dt,tol", [("fp64", np.float64, 1e-13), ("fp32", np.float32, 2e-6)])
def test_hxv_device_buffers_natural_layout(ed, orc, prec, dt, tol):
    """edl_hxv on caller buffers (natural layout i = iup + idw*DimUp, no
    padding) against the oracle on the same vector."""
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    isec = ed.sector_index(7, 6)
    dim, _, _ = ed.build_sector(isec)
    nb = dim * np.dtype(dt).itemsize
    v, hv = ed.alloc(nb), ed.alloc(nb)
    ed.vec_random(prec, v, dim, 77)
    ed.hxv_device(isec, prec, v, hv)
    vh = ed.copy_to_host(v, np.empty(dim, dt)).astype(np.float64)
    got = ed.copy_to_host(hv, np.empty(dim, dt)).astype(np.float64)
    ref = orc.matvec(m, 7, 6, vh.astype(np.complex128), nthreads=8).real
    assert _relerr(got, ref) < tol


@pytest.mark.parametrize("rank,cols", [(0, 0), (1, 0), (1, 8)])
@pytest.mark.parametrize("cfg,sec,rows", [("ring8", (4, 4), 16), ("rect3x4", (7, 6), 100), ("rect3x4", (6, 6), 300),
                                          ("rect3x4", (3, 6), 40), ("rect3x4", (11, 6), 4)])
def test_hxv_segmented_up_pass_vs_oracle(ed, orc, cfg, sec, rows, rank, cols, monkeypatch):
    """The segmented up pass (columns cut into row ranges that share their
    top sites' occupation; cross-segment neighbours from global memory),
    forced on small sectors by capping the rows per segment, with hop
    targets from the hop table (rank 0) and by rank arithmetic (rank 1:
    split rank tables, bonds below / above / across the split, local and
    leaving the segment, with and without sites between their ends), the
    panel width tuned or pinned to 8 columns (fp32; 4 in fp64)."""
    monkeypatch.setenv("EDL_UP_SEGMENT_ROWS", str(rows))
    monkeypatch.setenv("EDL_UP_SEG_RANK", str(rank))
    monkeypatch.setenv("EDL_UP_SEG_COLS", str(cols) if cols else "")
    c = lattice.CONFIGS[cfg]
    m = _use(ed, orc, c["H"](), [c["U"]])
    isec = ed.sector_index(*sec)
    dim, du, _ = ed.build_sector(isec)
    print(f"segmented up pass {cfg} {sec} DimUp {du} rows/segment {rows} rank {rank}", flush=True)
    up, dw, sched = ed.bench_passes(isec, "fp64", 1)
    assert sched["two_pass"] == 2
    assert sched["up_bank_scheduled"] == rank
    if cols:
        assert sched["up_cols"] == 4 and ed.bench_passes(isec, "fp32", 1)[2]["up_cols"] == cols
    for _ in range(2):
        v = RNG.standard_normal(dim)
        ref = orc.matvec(m, *sec, v, nthreads=8).real
        assert _relerr(ed.direct_matvec(isec, v, "fp64"), ref) < 1e-13
        assert _relerr(ed.direct_matvec(isec, v, "fp32"), ref) < 2e-6


@pytest.mark.parametrize("rpt,kch,R,dn,dk,vec,col,C", [
    (1, 8, 64, 2, 8, 0, 0, 3), (2, 4, 32, 1, 16, 0, 0, 3), (2, 4, 64, 2, 4, 1, 0, 3), (1, 8, 64, 1, 2, 1, 0, 3),
    (1, 16, 32, 2, 1, 1, 0, 3), (1, 8, 16, 2, 2, 1, 0, 3), (2, 8, 128, 1, 4, 1, 0, 3), (2, 4, 256, 2, 2, 1, 0, 3),
    (1, 8, 128, 2, 2, 1, 1, 3), (2, 4, 64, 1, 4, 1, 1, 3),
    (1, 8, 128, 2, 2, 1, 2, 3), (2, 4, 64, 1, 4, 1, 2, 3), (1, 16, 64, 2, 8, 0, 2, 1), (2, 8, 32, 1, 16, 0, 2, 1),
    (1, 8, 128, 2, 2, 1, 3, 3), (2, 4, 64, 1, 4, 1, 3, 2), (1, 16, 64, 2, 8, 0, 3, 1),
    (1, 8, 128, 2, 2, 1, 4, 3), (2, 4, 64, 1, 4, 1, 4, 1), (2, 8, 32, 2, 8, 0, 4, 3),
    (1, 16, 128, 2, 2, 1, 5, 3), (2, 4, 64, 1, 4, 1, 5, 1), (1, 16, 64, 2, 2, 1, 6, 3), (1, 8, 32, 1, 4, 1, 6, 1)])
def test_hxv_pinned_schedules_vs_oracle(ed, orc, rpt, kch, R, dn, dk, vec, col, C, monkeypatch):
    """Schedules pinned through the override, among them the vector-gather
    down pass (16-byte lanes, panels of 16 to 256 rows; the 8-site ring's
    DimUp 70 leaves a quad that straddles DimUp) and the bank-scheduled up
    pass (col = 1, fp32, and col = 2, its form with two hops to a table word
    and zero rows for the unused slots; the last 64-row block is partial on
    every sector here), the table-order two-hop table (col = 3, fp32 and
    fp64), the bank-scheduled two-hop words split by sign (col = 4, one
    coefficient per word; fp32 and fp64), and forms 2 and 4 with each
    block's next panel loaded into registers before the barrier that ends
    the current one (col = 5, 6; several panels per block on these grids):
    H x v on three sectors and a mixed-precision ground state (the
    chain-mode passes) against the oracle."""
    tail = f"{R},256,32,{dn},{dk},{rpt},{kch},{vec}"
    tail64 = f"{min(R, 128)},256,32,{dn},{dk},{rpt},{kch},{vec}"  # 256-row panels are fp32 only
    monkeypatch.setenv("EDL_TWO_PASS_CFG_F32", f"{C},1024,256," + tail + f",{col}")
    # fp64 (one column): the two-hop tables, bank-scheduled over 16 bank pairs (2) or in table order (3)
    monkeypatch.setenv("EDL_TWO_PASS_CFG_F64", "1,512,512," + tail64 + (f",{col}" if col >= 2 else ""))
    for cfg, sec in [("ring8", (4, 4)), ("rect3x4", (7, 6)), ("rect3x4", (6, 5))]:
        c = lattice.CONFIGS[cfg]
        m = _use(ed, orc, c["H"](), [c["U"]])
        isec = ed.sector_index(*sec)
        dim, du, _ = ed.build_sector(isec)
        print(f"pinned schedule {tail}: {cfg} {sec} DimUp {du}", flush=True)
        for prec in ("fp32", "fp64"):
            sched = ed.bench_passes(isec, prec, 1)[2]
            rows = 128 if (R == 256 and prec == "fp64") else R
            assert (sched["dw_rows"], sched["dw_vector_gather"], sched["up_rows_per_thread"]) == (rows, vec, rpt)
            assert sched["up_bank_scheduled"] == (col if prec == "fp32" or col >= 2 else 0)
        v = RNG.standard_normal(dim)
        ref = orc.matvec(m, *sec, v, nthreads=8).real
        assert _relerr(ed.direct_matvec(isec, v, "fp64"), ref) < 1e-13
        assert _relerr(ed.direct_matvec(isec, v, "fp32"), ref) < 2e-6
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    isec = ed.sector_index(6, 6)
    Hc = orc.hamiltonian_csr(m, 6, 6)
    e0, _, _ = orc.lanczos_ground_state(Hc, lattice.splitmix_vector(Hc.shape[0], 0x5EED), 200, want_vector=False)
    rmx = ed.lanczos_eigh(isec, "mixed", niter=200, refine_tol=1e-13)
    assert abs(rmx["e0"] - e0) < 1e-10 * abs(e0)


@pytest.mark.parametrize("R,b,th", [(32, 2, 256), (64, 3, 1024), (32, 4, 512)])
def test_hxv_range_blocked_down_pass_vs_oracle(ed, orc, R, b, th, monkeypatch):
    """The range-blocked down pass (row slabs of the column ranges that share
    the occupation of the top b sites staged in LDS, hops inside a range
    from LDS, the others from global memory), pinned: H x v on sectors of
    the 3x4 cluster against the oracle in fp64 and fp32 (natural and padded
    layouts), and a mixed ground state whose chains run the same form with
    the Lanczos terms fused (MODE 1)."""
    monkeypatch.setenv("EDL_TWO_PASS_CFG_F32", f"3,1024,256,{R},{th},32,2,8,1,8,0,1,{b}")
    monkeypatch.setenv("EDL_TWO_PASS_CFG_F64", f"1,512,512,{R},{th},32,2,8,2,8,0,0,{b}")
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    for sec in ((7, 6), (6, 6), (5, 6), (6, 4)):
        isec = ed.sector_index(*sec)
        dim, du, dd = ed.build_sector(isec)
        print(f"range-blocked R {R} b {b}: sector {sec} DimUp {du} DimDw {dd}", flush=True)
        for prec in ("fp32", "fp64"):
            assert ed.bench_passes(isec, prec, 1)[2]["dw_range_sites"] == b
        v = RNG.standard_normal(dim)
        ref = orc.matvec(m, *sec, v, nthreads=8).real
        assert _relerr(ed.direct_matvec(isec, v, "fp64"), ref) < 1e-13
        assert _relerr(ed.direct_matvec(isec, v, "fp32"), ref) < 2e-6
    isec = ed.sector_index(6, 6)
    Hc = orc.hamiltonian_csr(m, 6, 6)
    e0, _, _ = orc.lanczos_ground_state(Hc, lattice.splitmix_vector(Hc.shape[0], 0x5EED), 200, want_vector=False)
    rmx = ed.lanczos_eigh(isec, "mixed", niter=200, refine_tol=1e-13)
    assert abs(rmx["e0"] - e0) < 1e-10 * abs(e0)


@pytest.mark.parametrize("R,th,nch,dn,dk", [(128, 1024, 29, 2, 2), (64, 512, 32, 1, 4), (32, 256, 7, 2, 1)])
def test_hxv_ordered_down_pass_vs_oracle(ed, orc, R, th, nch, dn, dk, monkeypatch):
    """The vector-gather down pass taking its column chunks in the down
    spin's locality order (Cuthill-McKee over the hop graph; the vectors keep
    their layout), pinned: H x v on four sectors of the 3x4 cluster against
    the oracle in fp64 and fp32 (the engine's padded layout), bit-identical
    to the same shape in the natural order of columns (a column's hops are
    summed in the same order in both), and a mixed ground state whose chains
    run the same form with the Lanczos terms fused (the pin carries over to
    the chain's down pass)."""
    pins = {o: (f"3,1024,256,{R},{th},{nch},{dn},{dk},1,8,1,1,0,{o}", f"1,512,512,{R},{th},{nch},{dn},{dk},2,8,1,0,0,{o}")
            for o in (0, 1)}

    def pin(o):
        monkeypatch.setenv("EDL_TWO_PASS_CFG_F32", pins[o][0])
        monkeypatch.setenv("EDL_TWO_PASS_CFG_F64", pins[o][1])

    pin(1)
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    for sec in ((7, 6), (6, 6), (5, 6), (6, 4)):
        isec = ed.sector_index(*sec)
        dim, du, dd = ed.build_sector(isec)
        print(f"ordered down pass R {R} chunks {nch}: sector {sec} DimUp {du} DimDw {dd}", flush=True)
        for prec in ("fp32", "fp64"):
            assert ed.bench_passes(isec, prec, 1)[2]["dw_vector_gather"] == 2
        v = RNG.standard_normal(dim)
        ref = orc.matvec(m, *sec, v, nthreads=8).real
        h1 = {p: ed.direct_matvec(isec, v, p) for p in ("fp64", "fp32")}
        assert _relerr(h1["fp64"], ref) < 1e-13
        assert _relerr(h1["fp32"], ref) < 2e-6
        pin(0)
        for prec in ("fp32", "fp64"):
            assert ed.bench_passes(isec, prec, 1)[2]["dw_vector_gather"] == 1
            np.testing.assert_array_equal(ed.direct_matvec(isec, v, prec), h1[prec])
        pin(1)
    isec = ed.sector_index(6, 6)
    Hc = orc.hamiltonian_csr(m, 6, 6)
    e0, _, _ = orc.lanczos_ground_state(Hc, lattice.splitmix_vector(Hc.shape[0], 0x5EED), 200, want_vector=False)
    rmx = ed.lanczos_eigh(isec, "mixed", niter=200, refine_tol=1e-13)
    assert abs(rmx["e0"] - e0) < 1e-10 * abs(e0)


@pytest.mark.parametrize("R", [64, 128])
def test_hxv_vector_gather_degenerate_sectors(ed, orc, R, monkeypatch):
    """The vector-gather down pass on sectors where one spin has a single
    string (DimUp or DimDw = 1, and dim 1) or an odd DimUp (35: quads that
    straddle DimUp in fp32 and fp64), pinned with one column chunk."""
    monkeypatch.setenv("EDL_TWO_PASS_CFG_F32", f"3,1024,256,{R},256,1,2,2,1,8,1")
    monkeypatch.setenv("EDL_TWO_PASS_CFG_F64", f"1,512,512,{R},256,1,2,2,1,8,1")
    for H, U, sec in [(lattice.ring(8), 4.0, (0, 4)), (lattice.ring(8), 4.0, (4, 8)), (lattice.ring(8), 4.0, (8, 0)),
                      (lattice.ring(8), 4.0, (1, 7)), (lattice.ring(7), 3.0, (3, 3))]:
        m = _use(ed, orc, H, [U])
        isec = ed.sector_index(*sec)
        dim, du, dd = ed.build_sector(isec)
        print(f"vector gather R {R}: sector {sec} DimUp {du} DimDw {dd}", flush=True)
        assert ed.bench_passes(isec, "fp32", 1)[2]["dw_vector_gather"] == 1
        v = RNG.standard_normal(dim)
        ref = orc.matvec(m, *sec, v, nthreads=8).real
        assert _relerr(ed.direct_matvec(isec, v, "fp64"), ref) < 1e-13
        assert _relerr(ed.direct_matvec(isec, v, "fp32"), ref) < 2e-6


def test_lanczos_segmented_rect3x4_vs_oracle(ed, orc, monkeypatch):
    """Ground state through the segmented step (phase B, beta, segmented up,
    down, alpha): FP64 E0 against the oracle's Lanczos from the same start
    vector, the mixed-precision E0 against FP64, and the densities of the
    replayed eigenvector against the unsegmented run."""
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    isec = ed.sector_index(6, 6)
    ref_run = ed.lanczos_eigh(isec, "fp64", niter=200)
    obs_ref = ed.observables()
    monkeypatch.setenv("EDL_UP_SEGMENT_ROWS", "300")
    monkeypatch.setenv("EDL_UP_SEG_RANK", "1")  # the other form: the chain-mode test of the batched routes
    r = ed.lanczos_eigh(isec, "fp64", niter=200)
    obs = ed.observables()
    Hc = orc.hamiltonian_csr(m, 6, 6)
    e0, _, _ = orc.lanczos_ground_state(Hc, lattice.splitmix_vector(Hc.shape[0], 0x5EED), 200, want_vector=False)
    assert abs(r["e0"] - e0) < 1e-10 * abs(e0)
    np.testing.assert_allclose(r["alpha"][:30], ref_run["alpha"][:30], rtol=1e-10, atol=1e-10)
    for k in ("dens_up", "dens_dw", "docc"):
        np.testing.assert_allclose(obs[k], obs_ref[k], atol=1e-10)
    rmx = ed.lanczos_eigh(isec, "mixed", niter=200, refine_tol=1e-13)
    assert abs(rmx["e0"] - e0) < 1e-10 * abs(e0)


def test_hxv_rect3x4_gs_sector(ed, orc):
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    isec = ed.sector_index(6, 6)
    dim, _, _ = ed.build_sector(isec)
    v = RNG.standard_normal(dim)
    ref = orc.matvec(m, 6, 6, v, nthreads=8).real
    assert _relerr(ed.direct_matvec(isec, v, "fp64"), ref) < 1e-13


def test_hxv_sq4x4_full_size_rows_and_hermiticity(ed, orc):
    """4x4 PBC (8,8), dim 165 636 900: sampled rows against the oracle's row
    evaluator, and <u,Hv> = <Hu,v>."""
    m = _use(ed, orc, lattice.square(4, 4), [8.0])
    isec = ed.sector_index(8, 8)
    dim, _, _ = ed.build_sector(isec)
    assert dim == 165636900
    v = RNG.standard_normal(dim)
    hv = ed.direct_matvec(isec, v, "fp64")
    rows = np.concatenate([[0, dim - 1], RNG.integers(0, dim, 300)])
    ref = orc.hxv_rows(m, 8, 8, rows, v).real
    assert np.max(np.abs(hv[rows] - ref)) < 1e-13 * np.max(np.abs(hv))
    u = RNG.standard_normal(dim)
    hu = ed.direct_matvec(isec, u, "fp64")
    a, b = u @ hv, hu @ v
    assert abs(a - b) < 1e-12 * np.sqrt((u @ u) * (v @ v)) * 40


def test_lanczos_gs_ring8_vs_dense(ed, orc):
    m = _use(ed, orc, lattice.ring(8), [4.0])
    isec = ed.sector_index(4, 4)
    Hd = orc.hamiltonian_csr(m, 4, 4).toarray().real
    w, V = np.linalg.eigh(Hd)
    r = ed.lanczos_eigh(isec, "fp64", niter=200, want_vector=True)
    assert abs(r["e0"] - w[0]) < 1e-10 * abs(w[0])
    assert 1 - abs(r["gs"] @ V[:, 0]) < 1e-9
    # alpha/beta of the first steps equal the oracle's recurrence
    Hc = orc.hamiltonian_csr(m, 4, 4)
    a, b, _ = orc.lanczos_tridiag(Hc, lattice.splitmix_vector(Hd.shape[0], 0x5EED), 20)
    np.testing.assert_allclose(r["alpha"][:20], a, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(r["beta"][1:20], b[1:], rtol=1e-10, atol=1e-10)
    obs = ed.observables()
    ref = orc.observables(8, 4, 4, r["gs"])
    for k in ("dens_up", "dens_dw", "docc"):
        np.testing.assert_allclose(obs[k], ref[k], atol=1e-12)
    # <Sz_i Sz_j>, <n_i n_j> (OBSERVABLES.f90:142-150) from the bits of the
    # basis strings, element by element
    mu, md = orc.build_map(8, 4), orc.build_map(8, 4)
    w = (r["gs"] ** 2).reshape(md.size, mu.size)          # [idw, iup]
    bu = (mu[:, None] >> np.arange(8)) & 1                 # [iup, site]
    bd = (md[:, None] >> np.arange(8)) & 1
    n = bu[None, :, :] + bd[:, None, :]                    # [idw, iup, site]
    sz = 0.5 * (bu[None, :, :] - bd[:, None, :])
    n2 = np.einsum("ab,abi,abj->ij", w, n, n)
    sz2 = np.einsum("ab,abi,abj->ij", w, sz, sz)
    cor = ed.correlations()
    np.testing.assert_allclose(cor["n2"], n2, atol=1e-12)
    np.testing.assert_allclose(cor["sz2"], sz2, atol=1e-12)
    np.testing.assert_allclose(np.diag(cor["n2"]), obs["dens_up"] + obs["dens_dw"] + 2 * obs["docc"], atol=1e-12)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "mixed"])
def test_ground_state_vector_stored_basis_equals_replay(ed, orc, prec, monkeypatch):
    """The eigenvector from the stored Lanczos basis (one combination pass)
    equals the one from the second run of the chain (EDL_LANCZOS_BASIS=0),
    and both equal the dense ground state."""
    m = _use(ed, orc, lattice.ring(8), [4.0])
    isec = ed.sector_index(4, 4)
    w, V = np.linalg.eigh(orc.hamiltonian_csr(m, 4, 4).toarray().real)
    r1 = ed.lanczos_eigh(isec, prec, niter=120, want_vector=True)
    monkeypatch.setenv("EDL_LANCZOS_BASIS", "0")
    r0 = ed.lanczos_eigh(isec, prec, niter=120, want_vector=True)
    assert r1["e0"] == r0["e0"]
    assert np.max(np.abs(r1["gs"] - r0["gs"])) < 1e-12
    tol = 1e-9 if prec != "fp32" else 1e-5
    assert 1 - abs(r1["gs"] @ V[:, 0]) < tol
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    isec = ed.sector_index(6, 6)
    monkeypatch.delenv("EDL_LANCZOS_BASIS")
    ed.lanczos_eigh(isec, prec, niter=150)
    o1 = ed.observables()
    monkeypatch.setenv("EDL_LANCZOS_BASIS", "0")
    ed.lanczos_eigh(isec, prec, niter=150)
    o0 = ed.observables()
    for k in ("dens_up", "dens_dw", "docc"):
        np.testing.assert_allclose(o1[k], o0[k], atol=1e-12)


@pytest.mark.parametrize("cfg", ["ring8", "rect3x4"])
def test_mixed_precision_e0(ed, orc, cfg):
    c = lattice.CONFIGS[cfg]
    H = c["H"]()
    m = _use(ed, orc, H, [c["U"]])
    isec = ed.sector_index(c["nup"], c["ndw"])
    r64 = ed.lanczos_eigh(isec, "fp64", niter=300, want_vector=True)
    d64 = ed.observables()["docc"].mean()
    rmx = ed.lanczos_eigh(isec, "mixed", niter=300, want_vector=True, refine_tol=1e-13)
    dmx = ed.observables()["docc"].mean()
    print(f"{cfg}: E0 fp64 {r64['e0']:.14f} mixed {rmx['e0']:.14f} fp32-only {rmx['e0_fp32']:.14f} "
          f"steps fp32 {rmx['steps_fp32']} fp64 {rmx['steps_fp64']} docc {d64:.12f} {dmx:.12f}", flush=True)
    assert abs(rmx["e0"] - r64["e0"]) < 1e-10 * abs(r64["e0"])
    assert abs(dmx - d64) < 1e-7


def test_stored_path_route_e0(ed, orc):
    """ED_SPARSE_H=T routed to the on-the-fly kernels: with ED_FILLING = N
    the stored path keeps the -xmu*N - U/2*N diagonal terms the direct path
    drops, so E0 moves by exactly that constant; both runs agree with dense
    eigh of the oracle's matching operator."""
    H, U, xmu, sec = lattice.ring(8), 4.0, 0.3, (4, 4)
    N = sum(sec)
    e = {}
    for sparse in (False, True):
        ed.set_hamiltonian(H, [U], xmu=xmu, ed_filling=N, sparse_h=sparse)
        isec = ed.sector_index(*sec)
        e[sparse] = ed.lanczos_eigh(isec, "fp64", niter=200)["e0"]
        m = orc.Model(H, [U], xmu=xmu, ed_filling=N, sparse_h=sparse)
        w0 = np.linalg.eigvalsh(orc.hamiltonian_csr(m, *sec).toarray().real)[0]
        assert abs(e[sparse] - w0) < 1e-10 * abs(w0)
    assert abs((e[True] - e[False]) - (-(xmu + U / 2) * N)) < 1e-9


def test_kanamori_ground_state_vs_dense(ed, orc):
    """Ground state of a three-orbital model with spin exchange and pair
    hopping (the non-local block on the single-pass kernels): E0 and the
    double occupancies against dense eigh of the oracle's Hamiltonian,
    FP64 and mixed precision."""
    kw = dict(Nsites=[2, 2, 2], Ust=2.0, Jh=0.6, Jx=0.6, Jp=0.5, xmu=0.2)
    m = _use(ed, orc, _three_orb(), [4.0, 3.0, 2.5], **kw)
    isec = ed.sector_index(3, 3)
    w, V = np.linalg.eigh(orc.hamiltonian_csr(m, 3, 3).toarray().real)
    r = ed.lanczos_eigh(isec, "fp64", niter=300, want_vector=True)
    assert abs(r["e0"] - w[0]) < 1e-10 * abs(w[0])
    ref = orc.observables(6, 3, 3, V[:, 0])
    np.testing.assert_allclose(ed.observables()["docc"], ref["docc"], atol=1e-9)
    rmx = ed.lanczos_eigh(isec, "mixed", niter=300, refine_tol=1e-13)
    assert abs(rmx["e0"] - w[0]) < 1e-10 * abs(w[0])


def test_rect3x4_e0_vs_oracle(ed, orc):
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    isec = ed.sector_index(6, 6)
    r = ed.lanczos_eigh(isec, "fp64", niter=200)
    Hc = orc.hamiltonian_csr(m, 6, 6)
    e0, _, _ = orc.lanczos_ground_state(Hc, lattice.splitmix_vector(Hc.shape[0], 0x5EED), 200, want_vector=False)
    assert abs(r["e0"] - e0) < 1e-10 * abs(e0)


def _gz(poles, weights, z):
    return (weights[None, :] / (z[:, None] - poles[None, :])).sum(1)


def _ritz(a, b, n2, e0, isign):
    """Poles, weights and the last eigenvector component of each Ritz pair."""
    T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
    w, Z = np.linalg.eigh(T)
    return isign * (w - e0), n2 * Z[0] ** 2, np.abs(Z[-1])


def _assert_leading(a, b, ao, bo, k, tol):
    """The first k Lanczos coefficients fix the first 2k spectral moments;
    past them three-term Lanczos without reorthogonalisation amplifies
    rounding differences (ghost eigenvalues), so two correct
    implementations agree only on the leading part."""
    k = min(k, len(a), len(ao))
    np.testing.assert_allclose(a[:k], ao[:k], rtol=tol, atol=tol)
    np.testing.assert_allclose(b[1:k], bo[1:k], rtol=tol, atol=tol)


def _assert_converged_poles_match(gpu, ora, tol, window=1e-6, conv_tol=1e-10):
    """Every converged pole of `gpu` (weight > 1e-10 and last Ritz-vector
    component < 1e-10, i.e. residual ~ 0) matches the oracle: the weights
    summed over a window around it agree within tol and the weighted mean
    pole positions within tol.  Ghost copies of one eigenvalue, which
    three-term Lanczos without reorthogonalisation produces, fall in the
    same window and are summed."""
    p, w, zl = gpu
    po, wo, _ = ora
    conv = (w > 1e-10) & (zl < conv_tol)
    assert conv.sum() >= 1
    for pj in p[conv]:
        sg = np.abs(p - pj) < window
        so = np.abs(po - pj) < window
        wg, wo_ = w[sg].sum(), wo[so].sum()
        assert abs(wg - wo_) < tol, (pj, wg, wo_)
        if wg > 1e-6:
            assert abs((p[sg] * w[sg]).sum() / wg - (po[so] * wo[so]).sum() / wo_) < tol * max(1.0, 1e-6 / wg)


@pytest.fixture(params=["two_pass", "tiles", "persist", "single", "chain", "chain_segmented", "chain_segmented_rank"])
def batched_kernels(request, monkeypatch):
    """Kernels of the batched chains: the batched two-pass LDS kernels
    (default where the sectors fit; the faster of the tile launches and the
    persistent launch, or either one forced), the single-pass gather
    kernels, or one chain at a time through the two-pass step (the route of
    sectors large enough to fill the GPU alone), also with the segmented up
    pass (hop table, or hop targets by rank arithmetic)."""
    if request.param in ("tiles", "persist"):
        monkeypatch.setenv("EDL_BATCHED_KERNELS", request.param)
    elif request.param == "single":
        monkeypatch.setenv("EDL_BATCHED_KERNELS", "single")
    elif request.param.startswith("chain"):
        monkeypatch.setenv("EDL_BATCHED_KERNELS", "chain")
        if request.param.startswith("chain_segmented"):
            monkeypatch.setenv("EDL_UP_SEGMENT_ROWS", "24")
            monkeypatch.setenv("EDL_UP_SEG_RANK", "1" if request.param.endswith("rank") else "0")
    else:
        monkeypatch.delenv("EDL_BATCHED_KERNELS", raising=False)
    print(f"batched kernels: {request.param}", flush=True)
    return request.param


def test_tridiag_seed_vs_oracle(ed, orc, batched_kernels):
    m = _use(ed, orc, lattice.ring(8), [4.0])
    isec = ed.sector_index(5, 4)
    dim, _, _ = ed.build_sector(isec)
    seed = RNG.standard_normal(dim)
    a, b, n2 = ed.tridiag_Hv_sector(isec, seed, 100, "fp64")
    assert abs(n2 - seed @ seed) < 1e-10 * n2
    Hc = orc.hamiltonian_csr(m, 5, 4)
    ao, bo, _ = orc.lanczos_tridiag(Hc, seed, 100)
    # A random seed spreads over the whole spectrum: past ~40 steps the two
    # recurrences lose orthogonality differently (rounding-order chaos), so
    # the coefficients are compared where they are determined, and G(z) at a
    # broadening that only the leading moments control.
    np.testing.assert_allclose(a[:40], ao[:40], rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(b[1:40], bo[1:40], rtol=1e-8, atol=1e-8)
    z = np.linspace(-8, 8, 81) + 1.0j
    p, w = ed.lanczos_poles(a, b, n2, -3.0, 1)
    po, wo = orc.gf_poles(ao, bo, n2, -3.0, 1)
    np.testing.assert_allclose(_gz(p, w, z), _gz(po, wo, z), atol=1e-8 * n2)


def test_tridiag_one_up_string_sectors_vs_oracle(ed, orc, batched_kernels):
    """Sectors with a few up strings, whose padded column stride (32 rows)
    is many times DimUp: 3x4 (1,6) (DimUp 12) through tridiag_Hv_sector, and
    a batched group from the ring8 (2,4) ground state whose c seeds land in
    (1,4) (DimUp 8) and c^dg seeds in (3,4)."""
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    isec = ed.sector_index(1, 6)
    dim, du, _ = ed.build_sector(isec)
    assert du == 12
    seed = RNG.standard_normal(dim)
    a, b, n2 = ed.tridiag_Hv_sector(isec, seed, 60, "fp64")
    assert abs(n2 - seed @ seed) < 1e-10 * n2
    ao, bo, _ = orc.lanczos_tridiag(orc.hamiltonian_csr(m, 1, 6), seed, 60)
    np.testing.assert_allclose(a[:30], ao[:30], rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(b[1:30], bo[1:30], rtol=1e-8, atol=1e-8)
    m = _use(ed, orc, lattice.ring(8), [4.0])
    isec = ed.sector_index(2, 4)
    r = ed.lanczos_eigh(isec, "fp64", niter=200, want_vector=True)
    seeds = [("c", 1, 1, 0), ("c", 1, 3, 5), ("cdg", 1, 2, 0), ("c", 2, 4, 0)]
    rb = ed.tridiag_batched(seeds, 80, "fp64")
    assert list(rb["isector"]) == [ed.sector_index(1, 4)] * 2 + [ed.sector_index(3, 4), ed.sector_index(2, 3)]
    for k, (op, ispin, io, jo) in enumerate(seeds):
        opc = 2 if op == "cdg" else 1
        d = 1 if opc == 2 else -1
        nJ = (2 + d, 4) if ispin == 1 else (2, 4 + d)
        s = orc.seed_c_cdg(8, opc, io, ispin, (2, 4), nJ, r["gs"]).real
        if jo:
            s = s + orc.seed_c_cdg(8, opc, jo, ispin, (2, 4), nJ, r["gs"]).real
        assert abs(rb["norm2"][k] - s @ s) < 1e-10
        ao, bo, _ = orc.lanczos_tridiag(orc.hamiltonian_csr(m, *nJ), s, 80)
        _assert_leading(rb["alpha"][k], rb["beta"][k], ao, bo, 10, 1e-9)


@pytest.mark.parametrize("prec,tol,conv_tol,lead,nlead,window", [("fp64", 1e-8, 1e-10, 1e-9, 10, 1e-6),
                                                                 ("fp32", 2e-4, 1e-5, 1e-4, 6, 1e-4)])
def test_gf_batched_ring8_vs_oracle(ed, orc, prec, tol, conv_tol, lead, nlead, window, batched_kernels):
    """Config ring8: all diagonal + off-diagonal GF seeds of spin up in one
    batched launch, continued fractions vs the oracle's seeds + Lanczos.

    The GF seeds of this small cluster put their weight on a few poles that
    converge within a few steps; after that, loss of orthogonality grows the
    fp32 rounding by about 10x per step (measured: alpha agrees to 1e-6 up to
    step 6, 2e-3 at step 9), so fp32 is held to the first 6 coefficients,
    the converged poles and G(z) at broadening 1.  fp32 pole positions carry
    ~1e-6 absolute error (eps_fp32 * |H|), so their match window is 1e-4: a
    numpy fp32 Lanczos on the same seeds misses the oracle's weight by 0.3
    with a 1e-6 window and by 2e-6 with windows 1e-5 .. 1e-3."""
    m = _use(ed, orc, lattice.ring(8), [4.0])
    isec = ed.sector_index(4, 4)
    r = ed.lanczos_eigh(isec, "fp64", niter=300, want_vector=True)
    gs, e0 = r["gs"], r["e0"]
    sites = [1, 2, 5]
    gf, raw = ed.build_gf_electrons(e0, sites, offdiag=True, spins=(1, 2), nlanc=200, prec=prec)
    z = np.linspace(-6, 6, 121) + 1.0j
    for (ispin, io, jo), chans in gf.items():
        for ch, (op, isign) in enumerate([(2, 1), (1, -1)]):
            d = 1 if op == 2 else -1
            sec_j = (4 + d, 4) if ispin == 1 else (4, 4 + d)
            s = orc.seed_c_cdg(8, op, io, ispin, (4, 4), sec_j, gs).real
            if jo != io:
                s = s + orc.seed_c_cdg(8, op, jo, ispin, (4, 4), sec_j, gs).real
            Hj = orc.hamiltonian_csr(m, *sec_j)
            ao, bo, nst = orc.lanczos_tridiag(Hj, s, 200)
            po, wo = orc.gf_poles(ao, bo, s @ s, e0, isign)
            p, w = chans[ch]
            np.testing.assert_allclose(_gz(p, w, z), _gz(po, wo, z), atol=tol)
            assert abs(w.sum() - s @ s) < 1e-10
            k = [kk for kk, key in enumerate(gf) if key == (ispin, io, jo)][0]
            sidx = 2 * k + ch
            _assert_leading(raw["alpha"][sidx], raw["beta"][sidx], ao, bo, nlead, lead)
            g = _ritz(raw["alpha"][sidx], raw["beta"][sidx], raw["norm2"][sidx], e0, isign)
            _assert_converged_poles_match(g, _ritz(ao, bo, s @ s, e0, isign), tol, window=window,
                                          conv_tol=conv_tol)


@pytest.mark.parametrize("h16", ["0", "1"])
def test_batched_up_tile_table_forms_vs_oracle(ed, orc, h16, monkeypatch):
    """The batched up tiles with the up-spin hops one to a word (LDS
    coefficients) and two to a word (+-h from a register, zero rows for the
    odd last hop): both spins of the 3x4 seeds against the oracle's chains."""
    monkeypatch.setenv("EDL_BATCHED_KERNELS", "tiles")
    monkeypatch.setenv("EDL_BATCHED_UP_H16", h16)
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    r = ed.lanczos_eigh(ed.sector_index(6, 6), "fp64", niter=300, want_vector=True)
    seeds = [("cdg", 1, 1, 0), ("c", 1, 2, 7), ("cdg", 2, 3, 0), ("c", 2, 4, 9)]
    # fp32 chains: the leading coefficients to single-precision rounding
    # amplified over a few steps
    for prec, nk, tol in (("fp64", 8, 1e-9), ("fp32", 5, 1e-3)):
        rb = ed.tridiag_batched(seeds, 30, prec)
        for k, (op, sp, io, jo) in enumerate(seeds):
            opc = 2 if op == "cdg" else 1
            d = 1 if opc == 2 else -1
            nJ = (6 + d, 6) if sp == 1 else (6, 6 + d)
            s = orc.seed_c_cdg(12, opc, io, sp, (6, 6), nJ, r["gs"]).real
            if jo:
                s = s + orc.seed_c_cdg(12, opc, jo, sp, (6, 6), nJ, r["gs"]).real
            ao, bo, _ = orc.lanczos_tridiag(orc.hamiltonian_csr(m, *nJ), s, 30)
            _assert_leading(rb["alpha"][k], rb["beta"][k], ao, bo, nk, tol)


def test_stream_parts_equal_one_stream(ed, orc, monkeypatch):
    """The batched tile path splits a chain group over two to four streams
    (each part replayed from its own graph, the parts' launches
    interleaved): bit-identical alpha, beta and norm2 to one stream, on 16
    chains over both spins of the 3x4 seeds (four target sectors of two
    shapes; three parts of 5, 5 and 6 chains)."""
    monkeypatch.setenv("EDL_BATCHED_KERNELS", "tiles")
    _use(ed, orc, lattice.square(3, 4), [6.0])
    ed.lanczos_eigh(ed.sector_index(6, 6), "fp64", niter=200, want_vector=False)
    seeds = [(op, sp, i, j) for sp in (1, 2) for i, j in ((1, 0), (2, 7), (5, 0), (3, 12)) for op in ("cdg", "c")]
    res = {}
    for n in ("1", "2", "3", "4"):
        monkeypatch.setenv("EDL_BATCHED_STREAMS", n)
        for prec in ("fp64", "fp32"):
            rb = ed.tridiag_batched(seeds, 40, prec)
            res[(n, prec)] = (rb["alpha"], rb["beta"], rb["norm2"])
    for n in ("2", "3", "4"):
        for prec in ("fp64", "fp32"):
            for a, b in zip(res[("1", prec)], res[(n, prec)]):
                np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("route", ["tiles", "single", "chain"])
def test_direct_launches_equal_graph_replay(ed, orc, route, monkeypatch):
    """EDL_GRAPHS=0 launches each step's kernels directly instead of
    replaying the captured graph: the same kernels on the same data give
    bit-identical alpha, beta and ground-state energy, on every step runner
    (batched tiles, single-pass chains, one chain through the two-pass
    step)."""
    monkeypatch.setenv("EDL_BATCHED_KERNELS", route)
    _use(ed, orc, lattice.square(3, 4), [6.0])
    isec = ed.sector_index(6, 6)
    seeds = [("cdg", 1, 1, 0), ("c", 1, 2, 5), ("cdg", 2, 3, 0), ("c", 2, 4, 0)]
    res = {}
    for g in ("1", "0"):
        monkeypatch.setenv("EDL_GRAPHS", g)
        r = ed.lanczos_eigh(isec, "fp64", niter=120, want_vector=False)
        rb = ed.tridiag_batched(seeds, 60, "fp64")
        res[g] = (r["e0"], rb["alpha"], rb["beta"], rb["norm2"])
    assert res["0"][0] == res["1"][0]
    for a, b in zip(res["0"][1:], res["1"][1:]):
        np.testing.assert_array_equal(np.asarray(a), np.asarray(b))


@pytest.mark.parametrize("form", ["tiles", "persist"])
def test_gf_batched_equals_sequential_rect3x4(ed, orc, form, monkeypatch):
    """Config rect3x4: 2 x 78 seeds (i <= j pairs, particle + hole) of spin up;
    one batched launch (tile launches per step, or one persistent launch) vs
    one seed at a time, and a few seeds vs the oracle."""
    monkeypatch.setenv("EDL_BATCHED_KERNELS", form)
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    isec = ed.sector_index(6, 6)
    r = ed.lanczos_eigh(isec, "fp64", niter=300, want_vector=True)
    seeds = []
    for i in range(1, 13):
        for j in range(i, 13):
            jo = 0 if i == j else j
            seeds += [("cdg", 1, i, jo), ("c", 1, i, jo)]
    assert len(seeds) == 156
    rb = ed.tridiag_batched(seeds, 150, "fp64")
    rs = ed.tridiag_batched(seeds[:12], 150, "fp64", group=1)
    for k in range(12):
        _assert_leading(rb["alpha"][k], rb["beta"][k], rs["alpha"][k], rs["beta"][k], 10, 1e-9)
    print(f"rect3x4 156 seeds: batched {rb['ms']:.1f} ms, 12 sequential {rs['ms']:.1f} ms", flush=True)
    z = np.linspace(-10, 10, 101) + 1.0j
    csr = {}
    for k in (0, 1, 5):
        op, ispin, io, jo = seeds[k]
        opc = 2 if op == "cdg" else 1
        d = 1 if opc == 2 else -1
        s = orc.seed_c_cdg(12, opc, io, 1, (6, 6), (6 + d, 6), r["gs"]).real
        if jo:
            s = s + orc.seed_c_cdg(12, opc, jo, 1, (6, 6), (6 + d, 6), r["gs"]).real
        assert abs(rb["norm2"][k] - s @ s) < 1e-10
        if d not in csr:
            csr[d] = orc.hamiltonian_csr(m, 6 + d, 6)
        ao, bo, _ = orc.lanczos_tridiag(csr[d], s, 150)
        isign = 1 if opc == 2 else -1
        po, wo = orc.gf_poles(ao, bo, s @ s, r["e0"], isign)
        p, w = ed.lanczos_poles(rb["alpha"][k], rb["beta"][k], rb["norm2"][k], r["e0"], isign)
        np.testing.assert_allclose(_gz(p, w, z), _gz(po, wo, z), atol=1e-8)
        _assert_leading(rb["alpha"][k], rb["beta"][k], ao, bo, 10, 1e-9)
        _assert_converged_poles_match(_ritz(rb["alpha"][k], rb["beta"][k], rb["norm2"][k], r["e0"], isign),
                                      _ritz(ao, bo, s @ s, r["e0"], isign), 1e-8)


def test_chi_batched_vs_oracle(ed, orc, batched_kernels):
    m = _use(ed, orc, lattice.ring(8), [4.0])
    isec = ed.sector_index(4, 4)
    r = ed.lanczos_eigh(isec, "fp64", niter=300, want_vector=True)
    chi, raw = ed.build_chi_spin(r["e0"], [1, 3], offdiag=True, nlanc=150)
    Hc = orc.hamiltonian_csr(m, 4, 4)
    z = np.linspace(0, 8, 81) + 1.0j
    for (io, jo), (p, w) in chi.items():
        s = orc.seed_sz(8, io, (4, 4), r["gs"]).real
        if jo != io:
            s = s + orc.seed_sz(8, jo, (4, 4), r["gs"]).real
        ao, bo, _ = orc.lanczos_tridiag(Hc, s, 150)
        po, wo = orc.gf_poles(ao, bo, s @ s, r["e0"], 1)
        np.testing.assert_allclose(_gz(p, w, z), _gz(po, wo, z), atol=1e-8)
        k = list(chi).index((io, jo))
        _assert_leading(raw["alpha"][k], raw["beta"][k], ao, bo, 10, 1e-9)
        _assert_converged_poles_match(_ritz(raw["alpha"][k], raw["beta"][k], raw["norm2"][k], r["e0"], 1),
                                      _ritz(ao, bo, s @ s, r["e0"], 1), 1e-8)


def test_optical_conductivity_ring8_vs_oracle(ed, orc, batched_kernels):
    """Current seed of lanc_ed_build_oc on the ring8 ground state: norm2 and
    the first alpha/beta against the oracle's scatter seed and Lanczos, and
    the broadened optical conductivity and Drude weight of lanc_ed_eval_oc
    against the exact Lehmann sum of dense diagonalisation."""
    from ed_lattice_amd import optical_conductivity

    m = _use(ed, orc, lattice.ring(8), [4.0])
    isec = ed.sector_index(4, 4)
    r = ed.lanczos_eigh(isec, "fp64", niter=200, want_vector=True)
    gs, e0 = r["gs"], r["e0"]
    res = ed.tridiag_batched([("j", 0, 1, 0)], 200, "fp64")
    assert res["isector"][0] == isec
    a, b, n2 = res["alpha"][0], res["beta"][0], res["norm2"][0]
    seed = orc.seed_current(m, 4, 4, 1, gs).real
    assert abs(n2 - seed @ seed) < 1e-10 * (seed @ seed)
    # the seed lies in a small symmetry subspace: its Ritz values converge
    # within a few steps, after which two runs without reorthogonalisation
    # drift apart in rounding; the recurrence is compared over its first
    # steps and the spectrum through sigma(w) and the Drude weight below
    Hc = orc.hamiltonian_csr(m, 4, 4)
    ao, bo, _ = orc.lanczos_tridiag(Hc, seed, 10)
    np.testing.assert_allclose(a[:10], ao, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(b[1:10], bo[1:], rtol=1e-9, atol=1e-9)
    p, w = ed.lanczos_poles(a[:res["nsteps"][0]], b[:res["nsteps"][0]], n2, e0, +1)
    Hd = Hc.toarray().real
    pl, wl = orc.oc_lehmann(Hd, seed, e0)
    vr = np.linspace(0.0, 12.0, 400)
    dg, sg = optical_conductivity(p, w, vr, 0.2)
    dl, sl = optical_conductivity(pl, wl, vr, 0.2)
    print(f"OC ring8: norm2 {n2:.10f} Drude {dg:.10f} exact {dl:.10f}", flush=True)
    assert abs(dg - dl) < 1e-7 * abs(dl)
    np.testing.assert_allclose(sg, sl, atol=1e-7 * np.max(np.abs(sl)))


@pytest.mark.parametrize("cfg", ["ring8_mu", "rect3x4", "three_orbital", "three_orbital_filling"])
def test_local_energy_vs_oracle(ed, orc, cfg):
    """lanc_energy of the installed ground state: E0 = <K> + <Eloc> + <Hi> -
    xmu*N at ED_FILLING = 0 (<Hi> holds the Hartree terms that the diagonal
    carries then), <K> + <Eloc> + <Hi> - <Ehf> otherwise; <K>, Dse and Dph as
    expectation values of the oracle's operators built from only that term;
    <Eloc> from the basis; U*docc for one orbital."""
    if cfg == "ring8_mu":
        H, U, kw, sec = lattice.ring(8), [4.0], dict(xmu=0.3), (4, 4)
    elif cfg == "rect3x4":
        H, U, kw, sec = lattice.square(3, 4), [6.0], {}, (6, 6)
    else:
        H, U, sec = _three_orb(), [4.0, 3.0, 2.5], (3, 3)
        kw = dict(Nsites=[2, 2, 2], Ust=2.0, Jh=0.6, Jx=0.6, Jp=0.5, xmu=0.2)
        if cfg == "three_orbital_filling":
            kw["ed_filling"] = 6
    m = _use(ed, orc, H, U, **kw)
    isec = ed.sector_index(*sec)
    r = ed.lanczos_eigh(isec, "fp64", niter=300, want_vector=True)
    gs, e0 = r["gs"], r["e0"]
    e = ed.local_energy()
    N = sum(sec)
    if kw.get("ed_filling", 0) == 0:
        total = e["Ekin"] + e["Eknot"] + e["Epot"] - kw.get("xmu", 0.0) * N
    else:
        total = e["Ekin"] + e["Eknot"] + e["Epot"] - e["Ehartree"]
    assert abs(total - e0) < 1e-10 * abs(e0)
    Hs = np.array(H if np.ndim(H) == 3 else H[None], float).copy()
    Ns = Hs.shape[1]
    hop = Hs.copy()
    for i in range(Ns):
        hop[:, i, i] = 0.0
    nsites = kw.get("Nsites")
    zeros = [0.0] * len(U)
    mk = orc.Model(hop, zeros, Nsites=nsites, hfmode=False)
    assert abs(e["Ekin"] - gs @ orc.matvec(mk, *sec, gs).real) < 1e-10 * max(1.0, abs(e["Ekin"]))
    mu, md = orc.build_map(Ns, sec[0]), orc.build_map(Ns, sec[1])
    occ_u = (mu[:, None] >> np.arange(Ns)) & 1
    occ_d = (md[:, None] >> np.arange(Ns)) & 1
    w = (gs ** 2).reshape(len(md), len(mu))
    nu_site = (w.sum(0)[:, None] * occ_u).sum(0)
    nd_site = (w.sum(1)[:, None] * occ_d).sum(0)
    eknot = np.diag(Hs[0]) @ nu_site + np.diag(Hs[-1]) @ nd_site
    assert abs(e["Eknot"] - eknot) < 1e-10 * max(1.0, abs(eknot))
    if len(U) == 1:
        docc = ed.observables()["docc"].sum()
        assert abs(e["Eint"] - U[0] * docc) < 1e-10 * max(1.0, abs(e["Eint"]))
        assert e["Dse"] == 0.0 and e["Dph"] == 0.0
    else:
        zero_h = np.zeros_like(Hs)
        for key, jx, jp in (("Dse", 1.0, 0.0), ("Dph", 0.0, 1.0)):
            mj = orc.Model(zero_h, zeros, Nsites=nsites, Jx=jx, Jp=jp, hfmode=False)
            ref = gs @ orc.matvec(mj, *sec, gs).real
            assert abs(e[key] - ref) < 1e-10 * max(1.0, abs(ref)), key


@pytest.mark.parametrize("route", ["two_pass", "segmented", "fp32"])
def test_lanczos_states_ring8_vs_dense(ed, orc, route, monkeypatch):
    """Lowest eigenpairs by deflated Lanczos against dense eigh of the
    oracle's ring8 (4,4) Hamiltonian, degenerate levels included: energies,
    residuals |Hx - Ex| and orthonormality of the vectors."""
    if route == "segmented":
        monkeypatch.setenv("EDL_UP_SEGMENT_ROWS", "24")
        monkeypatch.setenv("EDL_UP_SEG_RANK", "1")
    m = _use(ed, orc, lattice.ring(8), [4.0])
    isec = ed.sector_index(4, 4)
    Hd = orc.hamiltonian_csr(m, 4, 4).toarray().real
    w = np.linalg.eigvalsh(Hd)
    k = 6
    prec = "fp32" if route == "fp32" else "fp64"
    r = ed.lanczos_states(isec, k, prec=prec, niter=400, tol=1e-13)
    e, X = r["energies"], r["vecs"]
    print(f"states {route}: {e} dense {w[:k]}", flush=True)
    tol = 2e-5 if prec == "fp32" else 1e-9
    np.testing.assert_allclose(np.sort(e), w[:k], rtol=tol, atol=tol)
    np.testing.assert_allclose(X @ X.T, np.eye(k), atol=1e-9)
    for j in range(k):
        res = np.linalg.norm(Hd @ X[j] - e[j] * X[j])
        assert res < (2e-3 if prec == "fp32" else 1e-6) * abs(e[j]), (j, res)


def test_lanczos_states_rect3x4(ed, orc):
    """3x4 (6,6), dim 853 776: the first state equals the ground state of
    the single-state solver, the others lie above it, and every vector is an
    eigenvector of the oracle's operator."""
    m = _use(ed, orc, lattice.square(3, 4), [6.0])
    isec = ed.sector_index(6, 6)
    e0 = ed.lanczos_eigh(isec, "fp64", niter=300, tol=1e-13)["e0"]
    r = ed.lanczos_states(isec, 3, prec="fp64", niter=400, tol=1e-13)
    e, X = r["energies"], r["vecs"]
    print(f"rect3x4 states {e}", flush=True)
    assert abs(e[0] - e0) < 1e-10 * abs(e0)
    assert np.all(e[1:] >= e[0] - 1e-9)
    np.testing.assert_allclose(X @ X.T, np.eye(3), atol=1e-9)
    for j in range(3):
        hx = orc.matvec(m, 6, 6, X[j], nthreads=8).real
        assert np.linalg.norm(hx - e[j] * X[j]) < 1e-5 * abs(e[j]), j


def test_new_model_drops_installed_state(ed, orc):
    """A new Hamiltonian invalidates the installed ground state: the calls
    that read it refuse until a state of the new model is installed."""
    from ed_lattice_amd.solver import EDError

    _use(ed, orc, lattice.ring(8), [4.0])
    isec = ed.sector_index(4, 4)
    ed.lanczos_eigh(isec, "fp64", niter=100)
    ed.observables()
    _use(ed, orc, lattice.ring(6), [4.0])
    with pytest.raises(EDError):
        ed.observables()
    with pytest.raises(EDError):
        ed.tridiag_batched([("c", 1, 1, 0)], 20, "fp64")
    r = ed.lanczos_eigh(ed.sector_index(3, 3), "fp64", niter=100)
    assert np.isfinite(ed.observables()["docc"]).all() and r["e0"] < 0
