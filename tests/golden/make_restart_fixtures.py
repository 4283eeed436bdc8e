"""Writes the restart-file fixtures of tests/golden (test infrastructure;
needs the GPU engine): the Fortran host (ed_lattice_amd/fortran/_build/
edlat_gpu, module EDLAT) run on the 6-site ring, U = 3, sector (3,3), with
GF_FLAG, OFFDIAG_GF_FLAG and CHISPIN_FLAG on and 20 Lanczos steps per seed;
its gfmatrix.restart and ChiSpinMatrix.restart are copied to
ring6_gfmatrix.restart and ring6_ChiSpinMatrix.restart.

    python3 tests/golden/make_restart_fixtures.py [output directory]
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ed_lattice_amd import lattice  # noqa: E402

INPUT = {"NORB": 1, "NSITES": 6, "NSPIN": 1, "ULOC": 3.0, "XMU": 0.0, "HFMODE": "T", "ED_FILLING": 0,
         "ED_SPARSE_H": "F", "ED_METHOD": "lanczos", "LANC_METHOD": "lanczos", "LANC_NSTATES_SECTOR": 1,
         "LANC_NITER": 400, "LANC_NGFITER": 20, "GF_FLAG": "T", "OFFDIAG_GF_FLAG": "T", "CHISPIN_FLAG": "T",
         "LMATS": 20, "LREAL": 20, "HIJ_FILE": "hij.in", "ED_PRECISION": "fp64", "SECTOR_NUP": 3,
         "SECTOR_NDW": 3}


def main(out):
    exe = os.path.join(ROOT, "ed_lattice_amd", "fortran", "_build", "edlat_gpu")
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "inputED.conf"), "w") as f:
            f.write("".join(f"{k}={v}\n" for k, v in INPUT.items()))
        lattice.write_hij(os.path.join(d, "hij.in"), lattice.ring(6))
        subprocess.run([exe], cwd=d, check=True, capture_output=True, timeout=300)
        os.makedirs(out, exist_ok=True)
        for name in ("gfmatrix", "ChiSpinMatrix"):
            shutil.copy(os.path.join(d, f"{name}.restart"), os.path.join(out, f"ring6_{name}.restart"))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else HERE)
