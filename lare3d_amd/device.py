"""Host-side mirror of the reference's hot-path module interface, running on
the MI355X through the C-ABI library (include/lare3d_dev.h).

    dev = Device(grid, params, fields)     # lare_dev_init + lare_dev_upload
    dev.lagrangian_step()                  # lagran.F90:50
    dev.eulerian_remap(step)               # remap.f90:33
    dev.eta_calc()                         # lagran.F90:1176
    dev.conduct_heat()                     # conduct.f90:39
    dev.boundary_conditions()              # boundary.f90:48
    dev.neutral_fraction()                 # neutral.f90:158 (EOS_ION / EOS_PI)
    dev.perpendicular_resistivity()        # neutral.f90:68 (Cowling)
    dev.energy_account()                   # diagnostics.F90:517
    dev.step(step)                         # lare3d.f90:84-87

Errors from the library raise DeviceError; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from .abi import DeviceError, LareScalars, LareStepInfo, check, load_library
from .fields import Fields
from .grid import grid_struct

MAX_KERNELS = 128
NAME_LEN = 48


class Device:
    # The library holds one device state per process (as the reference's
    # module variables do); a second open Device would silently share it.
    _open = None

    def __init__(self, grid, params, fields):
        if Device._open is not None and Device._open.lib is not None:
            raise DeviceError("the device library holds one state per process: "
                              "close the open Device before creating another")
        self.lib = load_library()
        self.grid = grid
        self.n = grid.n
        self._g = grid_struct(grid)
        self._p = params
        check(self.lib, self.lib.lare_dev_init(C.byref(self._g), C.byref(self._p)), "lare_dev_init")
        # the library copies the host arrays during the call; no host copy
        # is kept (at 512 x 512 x 1024 one field set is 26 GB)
        st = fields.to_struct(with_visc3=bool(params.use_viscous_damping))
        check(self.lib, self.lib.lare_dev_upload(C.byref(st)), "lare_dev_upload")
        Device._open = self

    def upload(self, fields):
        """Replace the device state with new host fields (same grid and
        parameters; the surrogate's parameters and rows are kept)."""
        st = fields.to_struct(with_visc3=bool(self._p.use_viscous_damping))
        check(self.lib, self.lib.lare_dev_upload(C.byref(st)), "lare_dev_upload")

    def close(self):
        if self.lib is not None:
            self.lib.lare_dev_finalize()
            self.lib = None
        if Device._open is self:
            Device._open = None

    def fields(self):
        f = Fields(*self.n)
        st = f.to_struct(with_visc3=False)
        check(self.lib, self.lib.lare_dev_download(C.byref(st)), "lare_dev_download")
        return f

    def scalars(self):
        s = LareScalars()
        check(self.lib, self.lib.lare_dev_get_scalars(C.byref(s)), "lare_dev_get_scalars")
        return s

    def set_scalars(self, **kw):
        s = self.scalars()
        for k, v in kw.items():
            setattr(s, k, v)
        check(self.lib, self.lib.lare_dev_set_scalars(C.byref(s)), "lare_dev_set_scalars")

    def lagrangian_step(self):
        check(self.lib, self.lib.lare_dev_lagrangian_step(), "lagrangian_step")

    def eulerian_remap(self, step):
        check(self.lib, self.lib.lare_dev_eulerian_remap(step), "eulerian_remap")

    def eta_calc(self):
        check(self.lib, self.lib.lare_dev_eta_calc(), "eta_calc")

    def energy_correction(self):
        """diagnostics.F90:613-629 (rke on): the remap's kinetic-energy loss
        added to the internal energy."""
        check(self.lib, self.lib.lare_dev_energy_correction(), "energy_correction")

    def conduct_heat(self):
        check(self.lib, self.lib.lare_dev_conduct_heat(), "conduct_heat")

    def boundary_conditions(self):
        check(self.lib, self.lib.lare_dev_boundary_conditions(), "boundary_conditions")

    def neutral_fraction(self):
        """neutral.f90:158-195 over the cell box (EOS_ION / EOS_PI)."""
        check(self.lib, self.lib.lare_dev_neutral_fraction(), "neutral_fraction")

    def perpendicular_resistivity(self):
        """neutral.f90:68-136 at nodes 0..n (cowling_resistivity)."""
        check(self.lib, self.lib.lare_dev_perpendicular_resistivity(), "perpendicular_resistivity")

    def neutral(self, xi=True, eta_perp=False):
        """(xi_n, eta_perp) in the (-1:n+2)^3 layout; None where not asked."""
        shape = tuple(n + 4 for n in self.n)
        a = np.zeros(shape, order="F") if xi else None
        b = np.zeros(shape, order="F") if eta_perp else None
        dp = C.POINTER(C.c_double)
        check(self.lib, self.lib.lare_dev_get_neutral(None if a is None else a.ctypes.data_as(dp),
                                                      None if b is None else b.ctypes.data_as(dp)), "get_neutral")
        return a, b

    def energy_account(self, with_divb=False):
        en = (C.c_double * 3)()
        dvb = C.c_double()
        check(self.lib, self.lib.lare_dev_energy_account(en, C.byref(dvb)), "energy_account")
        return (tuple(en), dvb.value) if with_divb else tuple(en)

    def divb(self):
        return self.energy_account(with_divb=True)[1]

    def step(self, step):
        dt, t = C.c_double(), C.c_double()
        check(self.lib, self.lib.lare_dev_step(step, C.byref(dt), C.byref(t)), f"step {step}")
        return dt.value, t.value

    def step_info(self, step, with_energy=True):
        """step() plus the per-step scalars (and energies) in one read-back."""
        info = LareStepInfo()
        check(self.lib, self.lib.lare_dev_step_info(step, 1 if with_energy else 0, C.byref(info)), f"step {step}")
        return info

    def potential_field(self, pf):
        """initial_conditions.f90:114 potential_field on the device (pf from
        lare3d_amd.initial.potential_inputs); returns (iterations, residual,
        bz_max)."""
        it, res, bm = C.c_int32(), C.c_double(), C.c_double()
        check(self.lib, self.lib.lare_dev_potential_field(C.byref(pf), C.byref(it), C.byref(res), C.byref(bm)),
              "potential_field")
        return it.value, res.value, bm.value

    def synchronize(self):
        check(self.lib, self.lib.lare_dev_synchronize(), "synchronize")

    def profile(self, enable):
        check(self.lib, self.lib.lare_dev_profile(1 if enable else 0), "profile")

    def kernel_stats(self):
        names = C.create_string_buffer(MAX_KERNELS * NAME_LEN)
        counts = (C.c_int32 * MAX_KERNELS)()
        ms = (C.c_double * MAX_KERNELS)()
        n = self.lib.lare_dev_kernel_stats(names, NAME_LEN, counts, ms, MAX_KERNELS)
        if n < 0:
            check(self.lib, n, "kernel_stats")
        out = {}
        raw = names.raw
        for q in range(n):
            nm = raw[q * NAME_LEN:(q + 1) * NAME_LEN].split(b"\0", 1)[0].decode()
            out[nm] = (counts[q], ms[q])
        return out
