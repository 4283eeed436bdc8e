"""Hopping matrices Hij of the benchmark lattices and the reference's Hij file format.

Hij is the real array Hij(Nspin, Ns, Ns) that ``Hij_get`` returns
(GRAPH_MATRIX.f90:334-345): off-diagonal entries are the hopping amplitudes
(-t on both (i,j) and (j,i)), diagonal entries the onsite energies, which
the direct H x v reads as ``Hdiag`` (HAMILTONIAN_DIRECT_HxV.f90:35-39).
Sites are numbered x + Lx*y (0-based here, 1-based in files).
"""
import numpy as np

_GAMMA = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 stream (Steele, Lea, Flood 2014); u() is uniform in [0, 1)."""

    def __init__(self, seed):
        self.state = seed & _M64

    def next(self):
        self.state = (self.state + _GAMMA) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def u(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def splitmix_vector(n, seed):
    """Element i = SplitMix64 output for state seed + (i+1)*gamma, mapped to U[-1,1).

    This is the Lanczos start vector the device library generates
    (``k_start_vector``) for a given seed."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _M64) + i * np.uint64(_GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) - 1.0


def bonds_rect(Lx, Ly, pbc_x, pbc_y):
    """Nearest-neighbour bonds (i, j), i < j not required, each listed once."""
    out = []
    for y in range(Ly):
        for x in range(Lx):
            s = x + Lx * y
            if x + 1 < Lx:
                out.append((s, (x + 1) + Lx * y))
            elif pbc_x and Lx > 1:
                out.append((s, 0 + Lx * y))
            if y + 1 < Ly:
                out.append((s, x + Lx * (y + 1)))
            elif pbc_y and Ly > 1:
                out.append((s, x + 0))
    return out


def hij_from_bonds(Ns, bonds, t=1.0, onsite=None, nspin=1):
    """Hij(Nspin, Ns, Ns) with -t_b on both orientations of every bond b."""
    H = np.zeros((nspin, Ns, Ns))
    tt = np.broadcast_to(np.asarray(t, dtype=float), (len(bonds),))
    for (i, j), tb in zip(bonds, tt):
        for s in range(nspin):
            H[s, i, j] += -tb
            H[s, j, i] += -tb
    if onsite is not None:
        for s in range(nspin):
            H[s, np.arange(Ns), np.arange(Ns)] += np.asarray(onsite, dtype=float)
    return H


def ring(L, t=1.0):
    return hij_from_bonds(L, bonds_rect(L, 1, True, False), t)


def square(Lx, Ly, t=1.0, pbc=True):
    return hij_from_bonds(Lx * Ly, bonds_rect(Lx, Ly, pbc, pbc), t)


def disordered_rect(Lx, Ly, seed=20261202, t_range=(0.8, 1.2), eps_range=(-0.5, 0.5), pbc=False):
    """Open (or periodic) rectangle with bond t_ij ~ U[t_range] drawn in bond
    order, then onsite e_i ~ U[eps_range] drawn in site order, from one
    SplitMix64 stream."""
    bonds = bonds_rect(Lx, Ly, pbc, pbc)
    rng = SplitMix64(seed)
    t = [t_range[0] + (t_range[1] - t_range[0]) * rng.u() for _ in bonds]
    eps = [eps_range[0] + (eps_range[1] - eps_range[0]) * rng.u() for _ in range(Lx * Ly)]
    return hij_from_bonds(Lx * Ly, bonds, t, onsite=eps)


def write_hij(path, H):
    """Hij_read format (GRAPH_MATRIX.f90:280-301): one link per line
    `i j a b spin re im`, 1-based, single orbital (a = b = 1)."""
    nspin, Ns, _ = H.shape
    with open(path, "w") as f:
        for s in range(nspin):
            for i in range(Ns):
                for j in range(Ns):
                    if H[s, i, j] != 0.0:
                        f.write(f"{i + 1} {j + 1} 1 1 {s + 1} {H[s, i, j]:.17g} 0.0\n")


def read_hij(path, Ns, nspin=1, norb=1, nsites=None):
    """Inverse of write_hij; orbital a of site i maps to pack_indices(i, a)
    (AUX_FUNX.f90:42-53)."""
    nsites = nsites or [Ns]
    H = np.zeros((nspin, Ns, Ns))
    with open(path) as f:
        for line in f:
            p = line.split()
            if len(p) < 7:
                continue
            i, j, a, b, s = (int(x) for x in p[:5])
            re, im = float(p[5]), float(p[6])
            if im != 0.0:
                raise ValueError("complex Hij is not supported")
            io = i if a == 1 else i + (a - 1) * nsites[a - 2]
            jo = j if b == 1 else j + (b - 1) * nsites[b - 2]
            H[s - 1, io - 1, jo - 1] = re
    return H


# The benchmark workloads: Hij, U, electrons.
CONFIGS = {
    "ring8": dict(H=lambda: ring(8), U=4.0, nup=4, ndw=4),
    "sq4x4": dict(H=lambda: square(4, 4), U=8.0, nup=8, ndw=8),
    "rect3x4": dict(H=lambda: square(3, 4), U=6.0, nup=6, ndw=6),
    "obc3x6": dict(H=lambda: disordered_rect(3, 6), U=4.0, nup=9, ndw=9),
}
