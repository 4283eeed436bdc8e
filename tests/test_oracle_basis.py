"""Pins of the oracle's basis, ranking and sign routines (CPU only).

Each check compares the restatement with what the mathematics fixes: the
ascending popcount-N list is the set of N-subsets in lexicographic order of
their integer value; binary_search must return the 1-based position; the
c/cdg bit-loop sign must equal (-1)^popcount(bits below pos)."""
import itertools

import numpy as np
import pytest


def _subsets(Ns, N):
    return sorted(sum(1 << p for p in c) for c in itertools.combinations(range(Ns), N))


@pytest.mark.parametrize("Ns", [1, 4, 7, 10, 12])
def test_build_map_is_sorted_subsets(orc, Ns):
    for N in range(Ns + 1):
        m = orc.build_map(Ns, N)
        assert m.tolist() == _subsets(Ns, N)


@pytest.mark.parametrize("Ns,N,dim", [(8, 4, 70), (16, 8, 12870), (12, 6, 924), (12, 7, 792), (12, 5, 792),
                                      (18, 9, 48620)])
def test_sector_sizes_of_the_workloads(orc, Ns, N, dim):
    assert orc.binomial(Ns, N) == dim
    if Ns <= 16:
        assert orc.build_map(Ns, N).size == dim


def test_binary_search_positions_and_misses(orc):
    m = orc.build_map(10, 4)
    for k, s in enumerate(m):
        assert orc.binary_search(m, int(s)) == k + 1
    present = set(m.tolist())
    for s in range(1 << 10):
        if s not in present:
            assert orc.binary_search(m, s) == 0


def test_c_cdg_signs_match_popcount(orc):
    Ns = 9
    for s in range(1 << Ns):
        for pos in range(1, Ns + 1):
            below = bin(s & ((1 << (pos - 1)) - 1)).count("1")
            sg = -1.0 if below % 2 else 1.0
            occ = (s >> (pos - 1)) & 1
            out, g = orc.c_op(pos, s)
            if occ:
                assert out == s & ~(1 << (pos - 1)) and g == sg
            else:
                assert out is None  # the reference stops here
            out, g = orc.cdg_op(pos, s)
            if not occ:
                assert out == s | (1 << (pos - 1)) and g == sg
            else:
                assert out is None


def test_hop_sign_is_popcount_between(orc):
    """c^dg_io c_jo on s: sg1*sg2 = (-1)^popcount(s & bits strictly between)."""
    Ns = 8
    for s in range(1 << Ns):
        for io in range(1, Ns + 1):
            for jo in range(1, Ns + 1):
                if io == jo or not (s >> (jo - 1)) & 1 or (s >> (io - 1)) & 1:
                    continue
                k1, g1 = orc.c_op(jo, s)
                k2, g2 = orc.cdg_op(io, k1)
                lo, hi = sorted((io - 1, jo - 1))
                between = ((1 << hi) - 1) & ~((2 << lo) - 1)
                expect = -1.0 if bin(s & between).count("1") % 2 else 1.0
                assert g1 * g2 == expect
                assert k2 == s ^ (1 << (io - 1)) ^ (1 << (jo - 1))


def test_combinatorial_rank_equals_position(orc):
    """rank(s) = sum_k C(p_k, k) (set bits p_1 < p_2 < ...) equals the
    0-based position in the ascending map -- the identity the device uses."""
    from math import comb

    for Ns, N in [(8, 4), (12, 6), (12, 7)]:
        m = orc.build_map(Ns, N)
        for k, s in enumerate(m):
            bits = [p for p in range(Ns) if (s >> p) & 1]
            r = sum(comb(p, i + 1) for i, p in enumerate(bits))
            assert r == k
    assert np.all(np.diff(orc.build_map(16, 8)) > 0)
