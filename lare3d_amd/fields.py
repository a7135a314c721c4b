"""Host field arrays with the reference's Fortran bounds (mpi_routines.F90:295-334).

Arrays are numpy float64 in Fortran order; element (i, j, k) of a field with
lower bounds (lx, ly, lz) is a[i - lx, j - ly, k - lz].
"""
import numpy as np

from .abi import LareFields, dptr

# lower bounds per field kind
LOWER = {
    "cell": (-1, -1, -1),
    "node": (-2, -2, -2),
    "bx": (-2, -1, -1),
    "by": (-1, -2, -1),
    "bz": (-1, -1, -2),
}
KIND = {"rho": "cell", "energy": "cell", "eta": "cell", "visc_dep": "cell", "ohmic_dep": "cell",
        "vx": "node", "vy": "node", "vz": "node", "visc3": "node",
        "bx": "bx", "by": "by", "bz": "bz"}
NAMES = ("rho", "energy", "vx", "vy", "vz", "bx", "by", "bz", "visc3", "eta", "visc_dep", "ohmic_dep")


def shape_of(kind, n):
    lo = LOWER[kind]
    return tuple(n[a] + 2 - lo[a] + 1 for a in range(3))


class Fields:
    """The state arrays the hot path reads and writes."""

    def __init__(self, nx, ny, nz):
        self.n = (nx, ny, nz)
        for name in NAMES:
            setattr(self, name, np.zeros(shape_of(KIND[name], self.n), order="F"))

    def view(self, name, i0, i1, j0, j1, k0, k1):
        """Slice by Fortran index ranges (inclusive)."""
        lo = LOWER[KIND[name]]
        a = getattr(self, name)
        return a[i0 - lo[0]:i1 - lo[0] + 1, j0 - lo[1]:j1 - lo[1] + 1, k0 - lo[2]:k1 - lo[2] + 1]

    def interior(self, name):
        nx, ny, nz = self.n
        return self.view(name, 1, nx, 1, ny, 1, nz)

    def copy(self):
        f = Fields(*self.n)
        for name in NAMES:
            getattr(f, name)[...] = getattr(self, name)
        return f

    def to_struct(self, with_visc3=True):
        s = LareFields()
        for name in NAMES:
            a = getattr(self, name)
            if name == "visc3" and not with_visc3:
                a = None
            setattr(s, name, dptr(a))
        return s
