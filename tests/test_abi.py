"""The C-ABI library: loads, exports every entry point include/edlat_gpu.h
declares, and its host-only functions work without a GPU (CPU only)."""
import os
import re

import numpy as np

from ed_lattice_amd import _lib, lanczos_poles, sector_index

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "edlat_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(edl_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
        assert n in _lib.SIGNATURES, n
    assert set(_lib.SIGNATURES) == set(names)


def test_sector_label_maths():
    lib = _lib.load()
    Ns = 16
    for nup in range(Ns + 1):
        for ndw in range(Ns + 1):
            isec = sector_index(Ns, nup, ndw)
            assert isec == 1 + nup * (Ns + 1) + ndw
            a, b = (np.zeros(1, np.int32) for _ in range(2))
            assert lib.edl_sector_qn(Ns, isec, a.ctypes.data_as(_lib._ip), b.ctypes.data_as(_lib._ip)) == 0
            assert (a[0], b[0]) == (nup, ndw)
    assert sector_index(16, 8, 8) == 145 and sector_index(12, 7, 6) == 98 and sector_index(18, 9, 9) == 181


def test_context_without_gpu_fails_cleanly():
    import ctypes

    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.edl_ctx_create(ctypes.byref(h), 0)
    if rc == 0:  # a GPU is visible: nothing to check here
        lib.edl_ctx_destroy(h)
        return
    assert not h.value
    assert lib.edl_last_error(None)


def test_pole_extraction_matches_oracle(orc):
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 150):
        a = rng.standard_normal(n)
        b = np.concatenate([[0.0], np.abs(rng.standard_normal(n - 1)) + 0.1])
        p, w = lanczos_poles(a, b, 0.37, -1.25, -1)
        po, wo = orc.gf_poles(a, b, 0.37, -1.25, -1)
        order = np.argsort(p)
        np.testing.assert_allclose(np.sort(p), np.sort(po), atol=1e-11)
        np.testing.assert_allclose(w[order], wo[np.argsort(po)], atol=1e-11)
        assert abs(w.sum() - 0.37) < 1e-12
