"""Test helper: the reference's restart-file records (read_formatted_gfmatrix,
VARS_GLOBAL.f90:611-632) read the way Fortran list-directed input reads them."""
import numpy as np


class FortranRecords:
    """List-directed reads as Fortran does them: every read starts a new
    record and takes records until it has its values (blank records are
    skipped; a read of zero values consumes one record); r*c repeats."""

    def __init__(self, path):
        with open(path) as f:
            self.lines = f.read().split("\n")
        self.i = 0

    def read(self, n):
        vals, first = [], True
        while first or len(vals) < n:
            first = False
            for tok in self.lines[self.i].replace(",", " ").split():
                if "*" in tok:
                    r, c = tok.split("*")
                    vals += [c] * int(r)
                else:
                    vals.append(tok)
            self.i += 1
        return vals[:n]

    def logical(self):
        return self.read(1)[0].upper().lstrip(".").startswith("T")

    def int(self):
        return int(self.read(1)[0])

    def reals(self, n):
        return np.array([float(t.replace("D", "E").replace("d", "e")) for t in self.read(n)])

    def exhausted(self):
        return all(not ln.strip() for ln in self.lines[self.i:])


def read_gfmatrix(rd):
    """read_formatted_gfmatrix (VARS_GLOBAL.f90:611-632): None when the
    status record is F, else [state][channel] -> (poles, weights)."""
    if not rd.logical():
        return None
    out = []
    for _ in range(rd.int()):
        chans = []
        for _ in range(rd.int()):
            n = rd.int()
            chans.append((rd.reals(n), rd.reals(n)))
        out.append(chans)
    return out
