"""TEST INFRASTRUCTURE ONLY -- the oracle.

Python side of the CPU restatement of the reference's H x v + Lanczos path
(aamaricci/ED_Lattice 1.6.0).  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may use it, and only as the checker.

Contents
  * ctypes binding of edl_oracle.c (the faithful restatement: maps,
    c/cdg, binary_search, directMatVec, seeds, observables);
  * a plain FP64 Lanczos over the oracle's own matrix (scipy CSR built from
    the restated scatter loops).  The reference delegates Lanczos to
    SciFortran (sp_lanc_eigh / sp_lanc_tridiag, not vendored, no pinned
    version): this is the textbook three-term recurrence those routines
    implement, pinned here by dense diagonalisation and closed forms;
  * pole/weight extraction of add_to_lanczos_gf_normal
    (GF_ELECTRONS.f90:417-457) with scipy's tridiagonal eigensolver;
  * an independent dense Hamiltonian built from Jordan-Wigner matrices
    (numpy Kronecker products), used only to pin the restatement.
"""
import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_long

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_build", "libedl_oracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        dp, ip, i32p, i64p = POINTER(c_double), POINTER(c_int), POINTER(c_int32), POINTER(c_int64)
        model = [c_int, c_int, c_int, ip, dp, dp, dp, c_double, c_double, c_double, c_double, c_double, c_int,
                 c_int]
        L.or_binomial.argtypes = [c_int, c_int]
        L.or_build_map.argtypes = [c_int, c_int, ip]
        L.or_binary_search.argtypes = [ip, c_int, c_int]
        L.or_c.argtypes = [c_int, c_int, ip, dp]
        L.or_cdg.argtypes = [c_int, c_int, ip, dp]
        L.or_direct_matvec.argtypes = model + [c_int, c_int, dp, dp, c_double]
        L.or_direct_matvec_omp.argtypes = model + [c_int, c_int, dp, dp, c_double, c_int]
        L.or_matvec_triplets.argtypes = model + [c_int, c_int, c_long, i32p, i32p, dp, dp]
        L.or_matvec_triplets.restype = c_long
        L.or_hxv_rows.argtypes = model + [c_int, c_int, c_int, i64p, dp, dp]
        L.or_apply_op_c_cdg.argtypes = [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, dp, dp]
        L.or_apply_op_sz.argtypes = [c_int, c_int, c_int, c_int, dp, dp]
        L.or_observables.argtypes = [c_int, c_int, c_int, dp, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(POINTER(c_double))


def _ip(a):
    return a.ctypes.data_as(POINTER(c_int))


class Model:
    """The module globals directMatVec reads: Hij (Nspin,Ns,Ns) and input variables."""

    def __init__(self, Hij, Uloc, Nsites=None, Ust=0.0, Jh=0.0, Jx=0.0, Jp=0.0, xmu=0.0, hfmode=True,
                 ed_filling=0, sparse_h=False):
        """sparse_h: the stored-H path's operator (ED_SPARSE_H=T).  Its
        diagonal applies the -xmu*N and hfmode -U/2*N terms unconditionally
        (stored/H_local.f90:61-90: the ED_FILLING test of the direct path,
        direct/HxV_local.f90:63, is commented out there), and its hops use
        conjg(Hij) (stored/H_up.f90:19, stored/H_dw.f90), equal to Hij for the
        real matrices restated here: the direct restatement with the filling
        test off."""
        H = np.asarray(Hij)
        if H.ndim == 2:
            H = H[None]
        self.H = H
        self.Nspin, self.Ns, _ = H.shape
        self.Uloc = np.atleast_1d(np.asarray(Uloc, dtype=np.float64))
        self.Norb = self.Uloc.size
        self.Nsites = np.asarray(Nsites if Nsites is not None else [self.Ns // self.Norb] * self.Norb,
                                 dtype=np.int32)
        self.Ust, self.Jh, self.Jx, self.Jp, self.xmu = Ust, Jh, Jx, Jp, xmu
        self.hfmode = int(bool(hfmode))
        self.ed_filling = 0 if sparse_h else ed_filling
        self._re = np.asfortranarray(np.real(H)).ravel(order="F").copy()
        self._im = np.asfortranarray(np.imag(H)).ravel(order="F").copy()

    def args(self):
        return (self.Ns, self.Nspin, self.Norb, _ip(self.Nsites), _dp(self._re), _dp(self._im), _dp(self.Uloc),
                self.Ust, self.Jh, self.Jx, self.Jp, self.xmu, self.hfmode, self.ed_filling)


def binomial(n, k):
    return lib().or_binomial(n, k)


def build_map(Ns, N):
    m = np.zeros(max(binomial(Ns, N), 1), np.int32)
    n = lib().or_build_map(Ns, N, _ip(m))
    return m[:n]


def binary_search(a, value):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return lib().or_binary_search(_ip(a), a.size, int(value))


def c_op(pos, s):
    out = c_int()
    sg = c_double()
    rc = lib().or_c(pos, s, ctypes.byref(out), ctypes.byref(sg))
    return (None, None) if rc else (out.value, sg.value)


def cdg_op(pos, s):
    out = c_int()
    sg = c_double()
    rc = lib().or_cdg(pos, s, ctypes.byref(out), ctypes.byref(sg))
    return (None, None) if rc else (out.value, sg.value)


def matvec(model, nup, ndw, v, frac=1.0, nthreads=0):
    """directMatVec on a complex(8) vector (faithful scatter restatement).
    The threaded split is race-free only without the non-local block, whose
    moves change both strings: models with it run serially."""
    vin = np.ascontiguousarray(v, dtype=np.complex128)
    out = np.zeros_like(vin)
    if model.Norb > 1 and (model.Jx != 0.0 or model.Jp != 0.0):
        nthreads = 0
    if nthreads:
        rc = lib().or_direct_matvec_omp(*model.args(), nup, ndw, _dp(vin.view(np.float64)),
                                        _dp(out.view(np.float64)), frac, nthreads)
    else:
        rc = lib().or_direct_matvec(*model.args(), nup, ndw, _dp(vin.view(np.float64)),
                                    _dp(out.view(np.float64)), frac)
    if rc:
        raise RuntimeError(f"or_direct_matvec failed ({rc})")
    return out


def hamiltonian_csr(model, nup, ndw):
    """The matrix of directMatVec as scipy CSR (duplicates summed)."""
    import scipy.sparse as sp

    L = lib()
    n = L.or_matvec_triplets(*model.args(), nup, ndw, 0, None, None, None, None)
    if n < 0:
        raise RuntimeError(f"or_matvec_triplets failed ({n})")
    ti = np.zeros(n, np.int32)
    tj = np.zeros(n, np.int32)
    tre = np.zeros(n)
    tim = np.zeros(n)
    L.or_matvec_triplets(*model.args(), nup, ndw, n, ti.ctypes.data_as(POINTER(c_int32)),
                         tj.ctypes.data_as(POINTER(c_int32)), _dp(tre), _dp(tim))
    D = binomial(model.Ns, nup) * binomial(model.Ns, ndw)
    vals = tre + 1j * tim if np.any(tim) else tre
    return sp.csr_matrix((vals, (ti, tj)), shape=(D, D))


def hxv_rows(model, nup, ndw, rows, v):
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    vin = np.ascontiguousarray(v, dtype=np.complex128)
    out = np.zeros(rows.size, np.complex128)
    rc = lib().or_hxv_rows(*model.args(), nup, ndw, rows.size, rows.ctypes.data_as(POINTER(c_int64)),
                           _dp(vin.view(np.float64)), _dp(out.view(np.float64)))
    if rc:
        raise RuntimeError(f"or_hxv_rows failed ({rc})")
    return out


def hxv_rows_real64(model, nup, ndw, rows, v32):
    """Rows of H v for a real fp32 vector with 64-bit element indices (for
    sectors beyond the reference's 32-bit Dim), accumulated in fp64."""
    L = lib()
    L.or_hxv_rows_real64.argtypes = [c_int, c_int, c_int, POINTER(c_int), POINTER(c_double), POINTER(c_double),
                                     POINTER(c_double), c_double, c_double, c_double, c_double, c_double, c_int,
                                     c_int, c_int, c_int, c_int, POINTER(c_int64), POINTER(ctypes.c_float),
                                     POINTER(c_double)]
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    v = np.ascontiguousarray(v32, dtype=np.float32)
    out = np.zeros(rows.size)
    rc = L.or_hxv_rows_real64(*model.args(), nup, ndw, rows.size, rows.ctypes.data_as(POINTER(c_int64)),
                              v.ctypes.data_as(POINTER(ctypes.c_float)), _dp(out))
    if rc:
        raise RuntimeError(f"or_hxv_rows_real64 failed ({rc})")
    return out


def hxv_rows_real64d(model, nup, ndw, rows, v64):
    """hxv_rows_real64 for a real fp64 vector (full-size eigenvector checks)."""
    L = lib()
    L.or_hxv_rows_real64d.argtypes = [c_int, c_int, c_int, POINTER(c_int), POINTER(c_double), POINTER(c_double),
                                      POINTER(c_double), c_double, c_double, c_double, c_double, c_double, c_int,
                                      c_int, c_int, c_int, c_int, POINTER(c_int64), POINTER(c_double),
                                      POINTER(c_double)]
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    v = np.ascontiguousarray(v64, dtype=np.float64)
    out = np.zeros(rows.size)
    rc = L.or_hxv_rows_real64d(*model.args(), nup, ndw, rows.size, rows.ctypes.data_as(POINTER(c_int64)), _dp(v),
                               _dp(out))
    if rc:
        raise RuntimeError(f"or_hxv_rows_real64d failed ({rc})")
    return out


def seed_c_cdg(Ns, op, ipos, ispin, nI, nJ, state):
    """apply_op_C (op=1) / apply_op_CDG (op=2) loop of GF_ELECTRONS.f90:241-245."""
    st = np.ascontiguousarray(state, dtype=np.complex128)
    DJ = binomial(Ns, nJ[0]) * binomial(Ns, nJ[1])
    out = np.zeros(DJ, np.complex128)
    rc = lib().or_apply_op_c_cdg(Ns, op, ipos, ispin, nI[0], nI[1], nJ[0], nJ[1], _dp(st.view(np.float64)),
                                 _dp(out.view(np.float64)))
    if rc:
        raise RuntimeError(f"or_apply_op_c_cdg failed ({rc})")
    return out


def seed_sz(Ns, ipos, n, state):
    st = np.ascontiguousarray(state, dtype=np.complex128)
    out = np.zeros_like(st)
    lib().or_apply_op_sz(Ns, ipos, n[0], n[1], _dp(st.view(np.float64)), _dp(out.view(np.float64)))
    return out


def observables(Ns, nup, ndw, state):
    st = np.ascontiguousarray(state, dtype=np.complex128)
    u, d, o = np.zeros(Ns), np.zeros(Ns), np.zeros(Ns)
    lib().or_observables(Ns, nup, ndw, _dp(st.view(np.float64)), _dp(u), _dp(d), _dp(o))
    return dict(dens_up=u, dens_dw=d, docc=o)


# ----------------------------------------------------------------- Lanczos
def lanczos_tridiag(H, v0, nlanc, brk=1e-14):
    """Three-term Lanczos from v0 (normalised here): alpha[nlanc], beta[nlanc]
    with beta[0] = 0, zero past a breakdown (sp_lanc_tridiag role,
    HAMILTONIAN.f90:258-271)."""
    a = np.zeros(nlanc)
    b = np.zeros(nlanc)
    v = np.asarray(v0, dtype=np.float64)
    v = v / np.sqrt(v @ v)
    vold = np.zeros_like(v)
    bk = 0.0
    nsteps = 0
    for k in range(nlanc):
        w = H @ v - bk * vold
        ak = w @ v
        w = w - ak * v
        a[k] = ak
        nsteps = k + 1
        bn = np.sqrt(w @ w)
        if k + 1 < nlanc:
            b[k + 1] = bn
        if not (bn > brk * (abs(ak) + bk)) or k + 1 >= nlanc:
            break
        vold, v, bk = v, w / bn, bn
    return a, b, nsteps


def tridiag_eigh(a, b):
    """Dense LAPACK eigh of the tridiagonal (diag a, off-diagonal b[1:])."""
    T = np.diag(np.asarray(a, float)) + np.diag(np.asarray(b, float)[1:], 1) + np.diag(np.asarray(b, float)[1:], -1)
    return np.linalg.eigh(T)


def lanczos_ground_state(H, v0, niter, want_vector=True):
    a, b, m = lanczos_tridiag(H, v0, niter)
    w, Z = tridiag_eigh(a[:m], b[:m])
    e0 = w[0]
    if not want_vector:
        return e0, None, (a[:m], b[:m])
    # two-pass replay with the Ritz coefficients
    c = Z[:, 0]
    v = np.asarray(v0, dtype=np.float64)
    v = v / np.sqrt(v @ v)
    vold = np.zeros_like(v)
    bk = 0.0
    x = np.zeros_like(v)
    for k in range(m):
        x += c[k] * v
        if k + 1 == m:
            break
        w_ = H @ v - bk * vold
        w_ = w_ - a[k] * v
        bn = b[k + 1]
        vold, v, bk = v, w_ / bn, bn
    x /= np.sqrt(x @ x)
    return e0, x, (a[:m], b[:m])


def gf_poles(a, b, norm2, Ei, isign):
    """add_to_lanczos_gf_normal: weight_j = norm2*Z(1,j)^2, pole_j = isign*(E_j - Ei)."""
    w, Z = tridiag_eigh(np.asarray(a, float), np.asarray(b, float))
    return isign * (w - Ei), norm2 * Z[0, :] ** 2


# ------------------------------------------------------- dense, independent
def _jw_annihilators(n):
    """Annihilation operators of n modes as dense 2^n x 2^n matrices; basis
    index = sum_p bit_p 2^p; mode p carries the Jordan-Wigner string over
    modes below p."""
    a = np.array([[0.0, 1.0], [0.0, 0.0]])
    z = np.diag([1.0, -1.0])
    eye = np.eye(2)
    ops = []
    for p in range(n):
        m = np.ones((1, 1))
        for q in reversed(range(n)):  # most significant factor first
            f = a if q == p else (z if q < p else eye)
            m = np.kron(m, f)
        ops.append(m)
    return ops


def dense_hamiltonian(model, nup, ndw):
    """H of one (Nup, Ndw) sector from JW matrices: up and down species carry
    separate strings (the two species commute, as in the reference), basis
    index iup + idw*DimUp with each species' strings in ascending order."""
    Ns = model.Ns
    ca = _jw_annihilators(Ns)
    cu = [np.kron(np.eye(2 ** Ns), c) for c in ca]   # up fastest
    cd = [np.kron(c, np.eye(2 ** Ns)) for c in ca]
    nu = [c.T @ c for c in cu]
    nd = [c.T @ c for c in cd]
    Hs = model.H
    dim = 4 ** Ns
    H = np.zeros((dim, dim))
    Nsp = model.Nspin
    for i in range(Ns):
        for j in range(Ns):
            if i != j:
                H += Hs[0, i, j].real * (cu[i].T @ cu[j]) + Hs[Nsp - 1, i, j].real * (cd[i].T @ cd[j])
        H += Hs[0, i, i].real * nu[i] + Hs[Nsp - 1, i, i].real * nd[i]

    def site(isite, iorb):  # pack_indices, 0-based result
        return isite if iorb == 1 else isite + (iorb - 1) * int(model.Nsites[iorb - 2])

    any_u = np.any(model.Uloc != 0)
    pairs = []
    for a_ in range(1, model.Norb + 1):
        for b_ in range(a_ + 1, model.Norb + 1):
            for s_ in range(1, min(model.Nsites[a_ - 1], model.Nsites[b_ - 1]) + 1):
                pairs.append((site(s_, a_) - 1, site(s_, b_) - 1))
    for iorb in range(1, model.Norb + 1):
        for s_ in range(1, model.Nsites[iorb - 1] + 1):
            io = site(s_, iorb) - 1
            if any_u:
                H += model.Uloc[iorb - 1] * nu[io] @ nd[io]
    for io, jo in pairs:
        H += model.Ust * (nu[io] @ nd[jo] + nu[jo] @ nd[io])
        H += (model.Ust - model.Jh) * (nu[io] @ nu[jo] + nd[io] @ nd[jo])
    if model.ed_filling == 0:
        for io in range(Ns):
            H -= model.xmu * (nu[io] + nd[io])
        if model.hfmode:
            if any_u:
                for iorb in range(1, model.Norb + 1):
                    for s_ in range(1, model.Nsites[iorb - 1] + 1):
                        io = site(s_, iorb) - 1
                        H -= 0.5 * model.Uloc[iorb - 1] * (nu[io] + nd[io])
            if model.Norb > 1:
                for io, jo in pairs:
                    tot = nu[io] + nd[io] + nu[jo] + nd[jo]
                    H -= 0.5 * model.Ust * tot + 0.5 * (model.Ust - model.Jh) * tot
    if model.Norb > 1 and (model.Jx != 0 or model.Jp != 0):
        for iorb in range(1, model.Norb + 1):
            for jorb in range(1, model.Norb + 1):
                for s_ in range(1, min(model.Nsites[iorb - 1], model.Nsites[jorb - 1]) + 1):
                    io, jo = site(s_, iorb) - 1, site(s_, jorb) - 1
                    if io != jo:
                        # spin exchange J c^dg_io,up c_jo,up c^dg_jo,dw c_io,dw
                        H += model.Jx * (cu[io].T @ cu[jo]) @ (cd[jo].T @ cd[io])
                    # pair hopping J c^dg_io,up c_jo,up c^dg_io,dw c_jo,dw
                    if io != jo:
                        H += model.Jp * (cu[io].T @ cu[jo]) @ (cd[io].T @ cd[jo])
    # project on the sector, in the reference order
    ups = [s for s in range(2 ** Ns) if bin(s).count("1") == nup]
    dws = [s for s in range(2 ** Ns) if bin(s).count("1") == ndw]
    idx = np.array([u + (2 ** Ns) * d for d in dws for u in ups])
    return H[np.ix_(idx, idx)]


# ------------------------------------------------------ optical conductivity
def seed_current(model, nup, ndw, iorb, state):
    """lanc_ed_build_oc (OC_ELECTRONS.f90:139-197) in its scatter form on the
    (nup, ndw) sector, before the global factor i of OC_ELECTRONS.f90:205:
    for each basis state, each site isite of orbital iorb and its ring
    successor jsite (the last wraps to the first), both spins: the hop
    jsite -> isite with +Hij(s,io,jo) and isite -> jsite with
    -conjg(Hij(s,jsite,isite)), each where that Hij element is non-zero, with
    the c / cdg signs and binary_search of AUX_FUNX.f90.  Python loops:
    small sectors only."""
    Ns, Nsp = model.Ns, model.Nspin
    mu, md = build_map(Ns, nup), build_map(Ns, ndw)
    du = len(mu)
    st = np.asarray(state, dtype=np.complex128)
    out = np.zeros(du * len(md), np.complex128)
    ns = int(model.Nsites[iorb - 1])

    def pack(isite):  # pack_indices, AUX_FUNX.f90:42-53 (1-based)
        return isite if iorb == 1 else isite + (iorb - 1) * int(model.Nsites[iorb - 2])

    for i in range(len(out)):
        if st[i] == 0:
            continue
        iup, idw = i % du, i // du
        mup, mdw = int(mu[iup]), int(md[idw])
        for isite in range(1, ns + 1):
            jsite = isite + 1 if isite < ns else 1
            for io, jo, sgn in ((pack(isite), pack(jsite), 1.0), (pack(jsite), pack(isite), -1.0)):
                for spin, m, hs in ((1, mup, 0), (2, mdw, Nsp - 1)):
                    h = model.H[hs, io - 1, jo - 1]
                    if h == 0 or not (m >> (jo - 1)) & 1 or (m >> (io - 1)) & 1:
                        continue
                    k1, s1 = c_op(jo, m)
                    k2, s2 = cdg_op(io, k1)
                    coef = h if sgn > 0 else -np.conj(h)
                    if spin == 1:
                        j = (binary_search(mu, k2) - 1) + idw * du
                    else:
                        j = iup + (binary_search(md, k2) - 1) * du
                    out[j] += coef * s1 * s2 * st[i]
    return out


def dense_current(model, nup, ndw, iorb):
    """The same current operator built independently from Jordan-Wigner
    matrices (sum over the orbital's ring bonds and both spins of
    H(s,a,b) c+_a c_b - H(s,b,a) c+_b c_a), projected on the sector in the
    reference order.  Full-Fock dense matrices (4^Ns squared): Ns <= 6."""
    assert model.Ns <= 6, "dense_current builds 4^Ns x 4^Ns matrices"
    Ns, Nsp = model.Ns, model.Nspin
    ca = _jw_annihilators(Ns)
    cu = [np.kron(np.eye(2 ** Ns), c) for c in ca]
    cd = [np.kron(c, np.eye(2 ** Ns)) for c in ca]
    ns = int(model.Nsites[iorb - 1])

    def pack(isite):
        return isite if iorb == 1 else isite + (iorb - 1) * int(model.Nsites[iorb - 2])

    K = np.zeros((4 ** Ns, 4 ** Ns))
    for isite in range(1, ns + 1):
        jsite = isite + 1 if isite < ns else 1
        a, b = pack(isite) - 1, pack(jsite) - 1
        if a == b:
            continue
        for ops, hs in ((cu, 0), (cd, Nsp - 1)):
            K += model.H[hs, a, b].real * (ops[a].T @ ops[b]) - model.H[hs, b, a].real * (ops[b].T @ ops[a])
    ups = [s for s in range(2 ** Ns) if bin(s).count("1") == nup]
    dws = [s for s in range(2 ** Ns) if bin(s).count("1") == ndw]
    idx = np.array([u + (2 ** Ns) * d for d in dws for u in ups])
    return K[np.ix_(idx, idx)]


def oc_lehmann(Hdense, seed, E0):
    """Exact current spectrum of a state: poles E_n - E0 and weights
    |<n|seed>|^2 over the eigenpairs of Hdense, seed = K|gs> (the function
    the continued fraction of add_to_lanczos_oc approximates)."""
    w, V = np.linalg.eigh(Hdense)
    amp = V.T @ np.asarray(seed, float)
    return w - E0, amp ** 2
