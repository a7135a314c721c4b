"""ctypes wrapper of the CPU restatement (liblare_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the device implementation.  Only
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.
"""
import ctypes as C
import os
import subprocess

from lare3d_amd.abi import LareFields, LareGrid, LareParams, LarePotential, LareScalars
from lare3d_amd.fields import Fields
from lare3d_amd.grid import grid_struct

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblare_oracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    # LARE_ORACLE_LIB names another build of the restatement (the mutation
    # check in oracle/mutation_check.py plants mistakes in copies)
    path = os.environ.get("LARE_ORACLE_LIB") or LIB
    if path == LIB and not os.path.exists(LIB):
        build()
    lib = C.CDLL(path)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    sig = {
        "orc_create": ([C.POINTER(LareGrid), C.POINTER(LareParams)], vp),
        "orc_destroy": ([vp], None),
        "orc_upload": ([vp, C.POINTER(LareFields)], C.c_int),
        "orc_download": ([vp, C.POINTER(LareFields)], C.c_int),
        "orc_get_scalars": ([vp, C.POINTER(LareScalars)], C.c_int),
        "orc_set_scalars": ([vp, C.POINTER(LareScalars)], C.c_int),
        "orc_lagrangian_step": ([vp], None),
        "orc_eulerian_remap": ([vp, C.c_int], None),
        "orc_eta_calc": ([vp], None),
        "orc_conduct_heat": ([vp], None),
        "orc_boundary_conditions": ([vp], None),
        "orc_energy_account": ([vp, dp], None),
        "orc_pow25": ([C.c_double], C.c_double),
        "orc_pow15": ([C.c_double], C.c_double),
        "orc_divb": ([vp], C.c_double),
        "orc_step": ([vp, C.c_int, dp, dp], None),
        "orc_potential_field": ([vp, C.POINTER(LarePotential), C.POINTER(C.c_int), dp, dp, dp], C.c_int),
        "orc_energy_correction": ([vp], None),
        "orc_get_delta_ke": ([vp, dp], C.c_int),
        "orc_neutral_fraction": ([vp], C.c_int),
        "orc_perpendicular_resistivity": ([vp], C.c_int),
        "orc_get_neutral": ([vp, dp, dp], C.c_int),
        "orc_get_fault": ([vp], C.c_int),
        "orc_exp": ([C.c_double], C.c_double),
        "orc_saha_neutral": ([C.POINTER(LareParams), C.c_double, C.c_double], C.c_double),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


class OracleFault(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def exp(x):
    """The restatement's exp (the one the neutral fraction uses)."""
    return lib().orc_exp(x)


def saha_neutral(params, t, rho):
    """get_neutral of neutral.f90:140-154 for one (T, rho)."""
    return lib().orc_saha_neutral(C.byref(params), t, rho)


class Oracle:
    """Same call surface as lare3d_amd.device.Device, on the CPU."""

    def __init__(self, grid, params, fields):
        self.grid = grid
        self.n = grid.n
        self._g = grid_struct(grid)
        self._p = params
        self.h = lib().orc_create(C.byref(self._g), C.byref(self._p))
        self._f = fields.copy()
        lib().orc_upload(self.h, C.byref(self._f.to_struct()))

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    __del__ = close

    def fields(self):
        f = Fields(*self.n)
        lib().orc_download(self.h, C.byref(f.to_struct(with_visc3=False)))
        return f

    def scalars(self):
        s = LareScalars()
        lib().orc_get_scalars(self.h, C.byref(s))
        return s

    def set_scalars(self, **kw):
        s = self.scalars()
        for k, v in kw.items():
            setattr(s, k, v)
        lib().orc_set_scalars(self.h, C.byref(s))

    def lagrangian_step(self):
        lib().orc_lagrangian_step(self.h)
        self._check_fault("lagrangian_step")

    def eulerian_remap(self, step):
        lib().orc_eulerian_remap(self.h, step)

    def eta_calc(self):
        lib().orc_eta_calc(self.h)

    def conduct_heat(self):
        lib().orc_conduct_heat(self.h)
        self._check_fault("conduct_heat")

    def boundary_conditions(self):
        lib().orc_boundary_conditions(self.h)

    def energy_account(self):
        en = (C.c_double * 3)()
        lib().orc_energy_account(self.h, en)
        return tuple(en)

    def divb(self):
        return lib().orc_divb(self.h)

    def step(self, step):
        dt, t = C.c_double(), C.c_double()
        lib().orc_step(self.h, step, C.byref(dt), C.byref(t))
        self._check_fault("step")
        return dt.value, t.value

    def _check_fault(self, what):
        f = lib().orc_get_fault(self.h)
        if f:
            raise OracleFault(f, f"{what}: neutral_fraction temperature bracket inverted (fault {f})")

    def neutral_fraction(self):
        rc = lib().orc_neutral_fraction(self.h)
        if rc < 0:
            raise RuntimeError(f"neutral_fraction: {rc}")
        self._check_fault("neutral_fraction")

    def perpendicular_resistivity(self):
        rc = lib().orc_perpendicular_resistivity(self.h)
        if rc:
            raise RuntimeError(f"perpendicular_resistivity: {rc}")

    def neutral(self, xi=True, eta_perp=False):
        """(xi_n, eta_perp) in the (-1:n+2)^3 layout; None where not asked."""
        import numpy as np
        shape = tuple(n + 4 for n in self.n)
        a = np.zeros(shape, order="F") if xi else None
        b = np.zeros(shape, order="F") if eta_perp else None
        rc = lib().orc_get_neutral(self.h, None if a is None else a.ctypes.data_as(C.POINTER(C.c_double)),
                                   None if b is None else b.ctypes.data_as(C.POINTER(C.c_double)))
        if rc:
            raise RuntimeError(f"get_neutral: {rc}")
        return a, b

    def energy_correction(self):
        lib().orc_energy_correction(self.h)

    def delta_ke(self):
        """delta_ke(-1:n+2)^3 as it stands (rke on)."""
        import numpy as np
        out = np.zeros(tuple(n + 4 for n in self.n), order="F")
        if lib().orc_get_delta_ke(self.h, out.ctypes.data_as(C.POINTER(C.c_double))) != 0:
            raise RuntimeError("delta_ke: rke is off")
        return out

    def potential_field(self, pf, want_phi=False):
        """initial_conditions.f90:114 potential_field; returns (iterations,
        residual, bz_max), plus phi(-1:n+2)^3 when want_phi."""
        import numpy as np
        it, res, bm = C.c_int(), C.c_double(), C.c_double()
        phi = np.zeros(tuple(n + 4 for n in self.n), order="F") if want_phi else None
        out = phi.ctypes.data_as(C.POINTER(C.c_double)) if want_phi else None
        lib().orc_potential_field(self.h, C.byref(pf), C.byref(it), C.byref(res), C.byref(bm), out)
        return (it.value, res.value, bm.value) + ((phi,) if want_phi else ())
