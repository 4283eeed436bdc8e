"""Host-side binding of the HIP engine with the reference's operator and driver names.

Every numerical step runs in libedlat_hip.so on the GPU; this module only
moves host arrays across the C-ABI and assembles the reference's result
structures (alpha/beta, GF/chi pole-weight lists, observables).

Reference entry points mirrored here:
  build_sector            SECTOR.f90:69-115          -> EDLattice.build_sector
  directMatVec            HAMILTONIAN_DIRECT_HxV.f90:19-59 -> EDLattice.direct_matvec
  sp_lanc_eigh (DIAG)     DIAG.f90:206-218           -> EDLattice.lanczos_eigh
  tridiag_Hv_sector       HAMILTONIAN.f90:242-275    -> EDLattice.tridiag_Hv_sector
  lanc_build_gf_electrons GF_ELECTRONS.f90:202-405   -> EDLattice.build_gf_electrons
  add_to_lanczos_gf_normal GF_ELECTRONS.f90:417-457  -> EDLattice.lanczos_poles
  lanc_eval_gf_electrons  GF_ELECTRONS.f90:472-512   -> eval_gf
  lanc_build_spinChi      CHI_SPIN_ELECTRONS.f90:189-293 -> EDLattice.build_chi_spin
  lanc_observables        OBSERVABLES.f90:78-192     -> EDLattice.observables
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import EDL_FP32, EDL_FP64, EDL_MIXED, EDL_OP_C, EDL_OP_CDG, EDL_OP_J, EDL_OP_SZ, SeedDesc

PREC = {"fp64": EDL_FP64, "fp32": EDL_FP32, "mixed": EDL_MIXED}


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _i32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


class EDError(RuntimeError):
    pass


def sector_index(Ns, nup, ndw):
    """get_Sector (SECTOR.f90:370-381): isector = 1 + Nup*(Ns+1) + Ndw."""
    return _lib.load().edl_sector_index(Ns, nup, ndw)


def lanczos_poles(alpha, beta, norm2, Ei, isign):
    """add_to_lanczos_gf_normal (GF_ELECTRONS.f90:417-457): poles, weights."""
    lib = _lib.load()
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    n = a.size
    p = np.zeros(n)
    w = np.zeros(n)
    if lib.edl_lanczos_poles(n, _dp(a), _dp(b), float(norm2), float(Ei), int(isign), _dp(p), _dp(w)):
        raise EDError(lib.edl_last_error(None).decode())
    return p, w


def optical_conductivity(poles, weights, vr, eps, cutoff=1e-9, pesoBZ=1.0):
    """lanc_ed_eval_oc (OC_ELECTRONS.f90:285-315) for one state's current
    channel: Drude weight sum peso/dE and OptCond_w(vr) = sum peso/dE *
    eps / ((vr - dE)^2 + eps^2), peso = pesoBZ * weight, skipping poles with
    |dE| < cutoff or |peso/dE| < cutoff (and non-finite ratios)."""
    p = np.asarray(poles, float)
    w = pesoBZ * np.asarray(weights, float)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = w / p
    keep = np.isfinite(r) & (np.abs(p) >= cutoff) & (np.abs(r) >= cutoff)
    r, p = r[keep], p[keep]
    vr = np.asarray(vr, float)
    sigma = (r[None, :] * eps / ((vr[:, None] - p[None, :]) ** 2 + eps ** 2)).sum(1)
    return float(r.sum()), sigma


def eval_gf(channels, wm, wr, eps, pesoBZ=1.0):
    """lanc_eval_gf_electrons (GF_ELECTRONS.f90:472-512) for one (ispin, io, jo):
    G(iw) = sum pesoBZ*w/(iw - p), G(w) = sum pesoBZ*w/(w + i eps - p)."""
    gm = np.zeros(len(wm), complex)
    gr = np.zeros(len(wr), complex)
    zm = 1j * np.asarray(wm)
    zr = np.asarray(wr) + 1j * eps
    for poles, weights in channels:
        for p, w in zip(poles, weights):
            gm += pesoBZ * w / (zm - p)
            gr += pesoBZ * w / (zr - p)
    return gm, gr


class EDLattice:
    """One GPU context holding a Hamiltonian, its sectors and the ground state."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        if self.lib.edl_ctx_create(ctypes.byref(h), device):
            raise EDError(self.lib.edl_last_error(None).decode())
        self.h = h
        self.Ns = None

    def _ck(self, rc):
        if rc:
            raise EDError(self.lib.edl_last_error(self.h).decode())

    def close(self):
        if self.h:
            self.lib.edl_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ model
    def set_hamiltonian(self, Hij, Uloc, Nsites=None, Ust=0.0, Jh=0.0, Jx=0.0, Jp=0.0, xmu=0.0, hfmode=True,
                        ed_filling=0, sparse_h=False):
        """sparse_h: the operator of the reference's stored-H path
        (ED_SPARSE_H=T, its default).  That path conjugates Hij on the hops
        (stored/H_up.f90:19; the same for the real Hij accepted here) and
        applies the -xmu*N and hfmode -U/2*N diagonal terms whatever
        ED_FILLING is (stored/H_local.f90:61-90, where the direct path has
        them only for ED_FILLING=0, direct/HxV_local.f90:63-92), so it is the
        direct operator with ED_FILLING=0 and runs on the same kernels."""
        H = np.asarray(Hij)
        if H.ndim == 2:
            H = H[None]
        nspin, Ns, _ = H.shape
        Uloc = np.atleast_1d(np.asarray(Uloc, dtype=np.float64))
        Norb = Uloc.size
        Nsites = np.asarray(Nsites if Nsites is not None else [Ns // Norb] * Norb, dtype=np.int32)
        re = np.asfortranarray(np.real(H)).ravel(order="F").copy()
        im = np.asfortranarray(np.imag(H)).ravel(order="F").copy() if np.iscomplexobj(H) else None
        self._ck(self.lib.edl_set_hamiltonian(self.h, Ns, nspin, Norb, _ip(Nsites), _dp(re),
                                              _dp(im) if im is not None else None, _dp(Uloc), Ust, Jh, Jx, Jp,
                                              xmu, int(bool(hfmode)), 0 if sparse_h else ed_filling))
        self.Ns = Ns

    # ----------------------------------------------------------- sectors
    def sector_index(self, nup, ndw):
        return sector_index(self.Ns, nup, ndw)

    def build_sector(self, isector):
        d = ctypes.c_int64()
        du = ctypes.c_int32()
        dd = ctypes.c_int32()
        self._ck(self.lib.edl_sector_build(self.h, isector, ctypes.byref(d), ctypes.byref(du), ctypes.byref(dd)))
        return d.value, du.value, dd.value

    def sector_maps(self, isector):
        _, du, dd = self.build_sector(isector)
        mu = np.zeros(du, np.int32)
        md = np.zeros(dd, np.int32)
        self._ck(self.lib.edl_sector_maps(self.h, isector, _i32(mu), _i32(md)))
        return mu, md

    def sector_hops(self, isector, ispin):
        _, du, dd = self.build_sector(isector)
        dim = du if ispin == 1 else dd
        k = ctypes.c_int32()
        self._ck(self.lib.edl_sector_hops(self.h, isector, ispin, ctypes.byref(k), None, None))
        nbr = np.zeros(max(k.value * dim, 1), np.int32)
        coef = np.zeros(max(k.value * dim, 1))
        self._ck(self.lib.edl_sector_hops(self.h, isector, ispin, ctypes.byref(k), _i32(nbr), _dp(coef)))
        n = k.value * dim
        return nbr[:n].reshape(k.value, dim), coef[:n].reshape(k.value, dim)

    def sector_rank(self, isector, ispin, states):
        s = np.ascontiguousarray(states, dtype=np.int32)
        out = np.zeros(s.size, np.int32)
        self._ck(self.lib.edl_sector_rank(self.h, isector, ispin, s.size, _i32(s), _i32(out)))
        return out

    # ---------------------------------------------------------- operator
    def direct_matvec(self, isector, v, prec="fp64"):
        """H v on the GPU for a host vector (real or complex)."""
        if np.iscomplexobj(v):
            vin = np.ascontiguousarray(v, dtype=np.complex128)
            out = np.zeros_like(vin)
            self._ck(self.lib.edl_hxv_host_complex(self.h, isector, PREC[prec], _dp(vin.view(np.float64)),
                                                   _dp(out.view(np.float64))))
            return out
        vin = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros_like(vin)
        self._ck(self.lib.edl_hxv_host(self.h, isector, PREC[prec], _dp(vin), _dp(out)))
        return out

    # ------------------------------------------------- device-side vectors
    def alloc(self, nbytes):
        """Device buffer owned by the context (freed when it is closed)."""
        p = ctypes.c_void_p()
        self._ck(self.lib.edl_device_alloc(self.h, int(nbytes), ctypes.byref(p)))
        return p.value

    def vec_random(self, prec, ptr, n, seed):
        self._ck(self.lib.edl_vec_random(self.h, PREC[prec], ptr, int(n), seed))

    def vec_dot(self, prec, a, b, n):
        out = ctypes.c_double()
        self._ck(self.lib.edl_vec_dot(self.h, PREC[prec], a, b, int(n), ctypes.byref(out)))
        return out.value

    def hxv_device(self, isector, prec, v_ptr, hv_ptr):
        """H v with both vectors already in HBM (edl_hxv)."""
        self._ck(self.lib.edl_hxv(self.h, isector, PREC[prec], v_ptr, hv_ptr))
        return self.last_timing()[0]

    def copy_to_host(self, ptr, out):
        self._ck(self.lib.edl_copy_d2h(self.h, out.ctypes.data, ptr, out.nbytes))
        return out

    def bench_hxv(self, isector, prec="fp32", reps=10):
        ms = ctypes.c_double()
        self._ck(self.lib.edl_bench_hxv(self.h, isector, PREC[prec], reps, ctypes.byref(ms)))
        return ms.value

    def bench_passes(self, isector, prec="fp32", reps=10):
        """Mean ms of each kernel launch of the operator and its schedule."""
        ms = np.zeros(2)
        cfg = np.zeros(16, np.int32)
        self._ck(self.lib.edl_bench_passes(self.h, isector, PREC[prec], reps, ms.ctypes.data_as(_lib._dp),
                                           cfg.ctypes.data_as(_lib._ip)))
        keys = ("two_pass", "ld", "up_cols", "up_threads", "up_grid", "dw_rows", "dw_threads", "dw_chunks",
                "dw_cols_per_lane", "dw_hop_chunk", "up_rows_per_thread", "up_hop_chunk", "dw_vector_gather",
                "up_coef_registers", "up_bank_scheduled", "dw_range_sites")
        return float(ms[0]), float(ms[1]), dict(zip(keys, (int(x) for x in cfg)))

    def last_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        self._ck(self.lib.edl_last_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    # ------------------------------------------------------ ground state
    def lanczos_eigh(self, isector, prec="fp64", niter=512, tol=0.0, seed=0x5EED, refine_tol=1e-13,
                     want_vector=False):
        dim, _, _ = self.build_sector(isector)
        e0 = ctypes.c_double()
        ns = ctypes.c_int()
        a = np.zeros(niter)
        b = np.zeros(niter)
        gs = np.zeros(dim) if want_vector else None
        self._ck(self.lib.edl_lanczos_gs(self.h, isector, PREC[prec], niter, tol, seed, refine_tol,
                                         ctypes.byref(e0), ctypes.byref(ns), _dp(a), _dp(b),
                                         _dp(gs) if gs is not None else None))
        lo = ctypes.c_int()
        hi = ctypes.c_int()
        elo = ctypes.c_double()
        self._ck(self.lib.edl_lanczos_stats(self.h, ctypes.byref(lo), ctypes.byref(hi), ctypes.byref(elo)))
        ms, nh = self.last_timing()
        return dict(e0=e0.value, nsteps=ns.value, alpha=a[:ns.value], beta=b[:ns.value], gs=gs,
                    steps_fp32=lo.value, steps_fp64=hi.value, e0_fp32=elo.value, ms=ms, hxv=nh)

    def lanczos_states(self, isector, nstates, prec="fp64", niter=300, tol=1e-12, seed=0x5EED, want_vectors=True):
        """Lowest nstates eigenpairs by deflated Lanczos (sp_eigh role for
        lanc_nstates_sector > 1): energies and, if asked, vectors (rows)."""
        dim, _, _ = self.build_sector(isector)
        e = np.zeros(nstates)
        v = np.zeros((nstates, dim)) if want_vectors else None
        self._ck(self.lib.edl_lanczos_states(self.h, isector, PREC[prec], nstates, niter, tol, seed, _dp(e),
                                             _dp(v) if v is not None else None))
        ms, nh = self.last_timing()
        return dict(energies=e, vecs=v, ms=ms, hxv=nh)

    def set_state(self, isector, energy, vec):
        v = np.ascontiguousarray(vec, dtype=np.float64)
        self._ck(self.lib.edl_set_state(self.h, isector, energy, _dp(v)))

    def get_state(self, dim):
        v = np.zeros(dim)
        self._ck(self.lib.edl_get_state(self.h, _dp(v)))
        return v

    def correlations(self):
        """<Sz_i Sz_j> and <n_i n_j> (Ns x Ns) of the installed state."""
        sz2 = np.zeros((self.Ns, self.Ns), order="F")
        n2 = np.zeros((self.Ns, self.Ns), order="F")
        self._ck(self.lib.edl_correlations(self.h, sz2.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                           n2.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return dict(sz2=sz2, n2=n2)

    def observables(self):
        """dens_up, dens_dw, docc per site of the installed ground state."""
        n = self.Ns
        u, d, o = np.zeros(n), np.zeros(n), np.zeros(n)
        self._ck(self.lib.edl_observables(self.h, _dp(u), _dp(d), _dp(o)))
        return dict(dens_up=u, dens_dw=d, dens=u + d, docc=o)

    ENERGY_KEYS = ("Ekin", "Epot", "Eint", "Eknot", "Ehartree", "Dust", "Dund", "Dse", "Dph", "Dkxy", "Dkz")

    def local_energy(self):
        """lanc_energy (ENERGY.f90:70-353) of the installed ground state, in
        the column order of energy_last.ed: <K>, <Hi>, <V> = <Hi> - <Ehf>,
        <Eloc>, <Ehf>, Dust, Dund, Dse, Dph, Dkxy, Dkz."""
        e = np.zeros(11)
        self._ck(self.lib.edl_local_energy(self.h, _dp(e)))
        return dict(zip(self.ENERGY_KEYS, e))

    # ------------------------------------------------------------ seeds
    def tridiag_Hv_sector(self, isector, vvinit, nlanc, prec="fp64"):
        v = np.ascontiguousarray(vvinit, dtype=np.float64)
        a = np.zeros(nlanc)
        b = np.zeros(nlanc)
        n2 = ctypes.c_double()
        ns = ctypes.c_int()
        self._ck(self.lib.edl_tridiag(self.h, isector, PREC[prec], _dp(v), nlanc, _dp(a), _dp(b),
                                      ctypes.byref(n2), ctypes.byref(ns)))
        return a, b, n2.value

    def tridiag_batched(self, seeds, nlanc, prec="fp64", group=0):
        """seeds: list of (op, ispin, io, jo) with op in {'c','cdg','sz','j'};
        'j' is the optical-conductivity current of orbital io (ispin and jo
        unused, give 0)."""
        opmap = {"c": EDL_OP_C, "cdg": EDL_OP_CDG, "sz": EDL_OP_SZ, "j": EDL_OP_J}
        n = len(seeds)
        arr = (SeedDesc * max(n, 1))()
        for k, (op, ispin, io, jo) in enumerate(seeds):
            arr[k] = SeedDesc(opmap[op], ispin, io, jo)
        a = np.zeros(n * nlanc)
        b = np.zeros(n * nlanc)
        n2 = np.zeros(n)
        ns = np.zeros(n, np.int32)
        js = np.zeros(n, np.int32)
        self._ck(self.lib.edl_tridiag_batched(self.h, n, arr, nlanc, PREC[prec], group, _dp(a), _dp(b), _dp(n2),
                                              _ip(ns), _ip(js)))
        ms, nh = self.last_timing()
        return dict(alpha=a.reshape(n, nlanc), beta=b.reshape(n, nlanc), norm2=n2, nsteps=ns, isector=js, ms=ms,
                    hxv=nh)

    def lanczos_poles(self, alpha, beta, norm2, Ei, isign):
        return lanczos_poles(alpha, beta, norm2, Ei, isign)

    def build_gf_electrons(self, Egs, sites, offdiag=False, spins=(1,), nlanc=200, prec="fp64", group=0):
        """GF pole/weight lists for the installed ground state (one state, T=0).

        Returns {(ispin, io, jo): [(poles, weights) particle, (poles, weights) hole]}
        for io == jo in `sites` and, with offdiag, every io != jo pair
        (the mixed seeds (c_io + c_jo)).  Channel 1 = c^dg (isign +1),
        channel 2 = c (isign -1), as GF_ELECTRONS.f90:251-252, 277-278."""
        keys, seeds = [], []
        for s in spins:
            for io in sites:
                keys.append((s, io, io))
                seeds += [("cdg", s, io, 0), ("c", s, io, 0)]
            if offdiag:
                for io in sites:
                    for jo in sites:
                        if io == jo:
                            continue
                        keys.append((s, io, jo))
                        seeds += [("cdg", s, io, jo), ("c", s, io, jo)]
        r = self.tridiag_batched(seeds, nlanc, prec, group)
        out = {}
        for k, key in enumerate(keys):
            chans = []
            for ch, isign in ((0, 1), (1, -1)):
                s = 2 * k + ch
                if r["isector"][s] == 0 or r["nsteps"][s] == 0:
                    chans.append((np.zeros(0), np.zeros(0)))
                    continue
                nl = self._nlanc_of(r, s, nlanc)
                p, w = lanczos_poles(r["alpha"][s, :nl], r["beta"][s, :nl], r["norm2"][s], Egs, isign)
                chans.append((p, w))
            out[key] = chans
        return out, r

    def _nlanc_of(self, r, s, nlanc):
        d, _, _ = self.build_sector(int(r["isector"][s]))
        return int(min(d, nlanc))

    def build_chi_spin(self, Egs, sites, offdiag=False, nlanc=200, prec="fp64", group=0):
        """Spin-chi pole/weight lists: {(io, jo): (poles, weights)}, poles = E_j - Ei
        (CHI_SPIN_ELECTRONS.f90:305-366)."""
        keys, seeds = [], []
        for io in sites:
            keys.append((io, io))
            seeds.append(("sz", 1, io, 0))
        if offdiag:
            for io in sites:
                for jo in sites:
                    if io != jo:
                        keys.append((io, jo))
                        seeds.append(("sz", 1, io, jo))
        r = self.tridiag_batched(seeds, nlanc, prec, group)
        out = {}
        for k, key in enumerate(keys):
            nl = self._nlanc_of(r, k, nlanc)
            out[key] = lanczos_poles(r["alpha"][k, :nl], r["beta"][k, :nl], r["norm2"][k], Egs, 1)
        return out, r
