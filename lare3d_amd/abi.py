"""ctypes view of the C-ABI declared in include/lare3d_dev.h.

The structs mirror the header field for field.  The same structs describe
the inputs of the device library and of the CPU oracle used by the tests.
"""
import ctypes as C
import os

BC_PERIODIC = 1
BC_USER = 2
BC_OPEN = 3
PREC_FP64 = 0
PREC_FP32 = 1
EOS_IDEAL = 1
EOS_ION = 2
EOS_PI = 3
F_NONFINITE = 1
F_TBOUNDS = 2

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


class LareGrid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("xbc_min", C.c_int32), ("xbc_max", C.c_int32),
                ("ybc_min", C.c_int32), ("ybc_max", C.c_int32),
                ("zbc_min", C.c_int32), ("zbc_max", C.c_int32),
                ("dxb", _dp), ("dxc", _dp), ("dyb", _dp), ("dyc", _dp),
                ("dzb", _dp), ("dzc", _dp), ("grav", _dp)]


class LareParams(C.Structure):
    _fields_ = [("gamma", C.c_double), ("visc1", C.c_double), ("visc2", C.c_double),
                ("eta_background", C.c_double), ("eta0", C.c_double), ("j_max", C.c_double),
                ("kappa_0", C.c_double), ("flux_limiter", C.c_double), ("xi_n", C.c_double),
                ("dt_from_restart", C.c_double),
                ("resistive_mhd", C.c_int32), ("conduction", C.c_int32),
                ("heat_flux_limiter", C.c_int32), ("use_viscous_damping", C.c_int32),
                ("restart", C.c_int32), ("precision", C.c_int32),
                ("alpha_av", C.c_double), ("va_max", C.c_double),
                ("cooling_term", C.c_int32), ("boris", C.c_int32),
                ("shock_limiter", C.c_int32), ("shock_compression", C.c_int32), ("rke", C.c_int32),
                ("cauchy", C.c_int32), ("fourth_order", C.c_int32),
                ("eos_number", C.c_int32), ("cowling_resistivity", C.c_int32), ("pad2_", C.c_int32),
                ("ionise_pot", C.c_double), ("f_bar", C.c_double), ("t_bar", C.c_double),
                ("r_bar", C.c_double), ("eta_bar", C.c_double), ("t_rad", C.c_double)]


class LareStepInfo(C.Structure):
    _fields_ = [("dt", C.c_double), ("dtr", C.c_double), ("time", C.c_double),
                ("total_visc_heating", C.c_double), ("total_ohmic_heating", C.c_double),
                ("en_b", C.c_double), ("en_ke", C.c_double), ("en_int", C.c_double), ("max_divb", C.c_double),
                ("n_s_stages", C.c_int32), ("substeps", C.c_int32), ("peak_substeps", C.c_int32),
                ("sur_accepted", C.c_int32), ("sur_rejected", C.c_int32), ("with_energy", C.c_int32)]


class LarePotential(C.Structure):
    _fields_ = [("omega", C.c_double), ("fractional_error", C.c_double), ("dx1", C.c_double),
                ("dy1", C.c_double), ("dz1", C.c_double), ("flux0", C.POINTER(C.c_double)),
                ("max_iterations", C.c_int32), ("pad_", C.c_int32)]


class LareFields(C.Structure):
    _fields_ = [(n, _dp) for n in ("rho", "energy", "vx", "vy", "vz", "bx", "by", "bz",
                                   "visc3", "eta", "visc_dep", "ohmic_dep")]


class LareScalars(C.Structure):
    _fields_ = [("dt", C.c_double), ("dtr", C.c_double), ("time", C.c_double),
                ("total_visc_heating", C.c_double), ("total_ohmic_heating", C.c_double),
                ("n_s_stages", C.c_int32), ("peak_substeps", C.c_int32),
                ("last_substeps", C.c_int32), ("first_set_dt", C.c_int32),
                ("step", C.c_int32)]


def dptr(a):
    """Pointer to the first element of a float64 numpy array (or NULL)."""
    if a is None:
        return C.cast(None, _dp)
    assert a.dtype.name == "float64" and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(_dp)


# Exported symbols of the device library, with argtypes.
DEVICE_SYMBOLS = {
    "lare_dev_init": [C.POINTER(LareGrid), C.POINTER(LareParams)],
    "lare_dev_upload": [C.POINTER(LareFields)],
    "lare_dev_download": [C.POINTER(LareFields)],
    "lare_dev_get_scalars": [C.POINTER(LareScalars)],
    "lare_dev_set_scalars": [C.POINTER(LareScalars)],
    "lare_dev_lagrangian_step": [],
    "lare_dev_eulerian_remap": [C.c_int32],
    "lare_dev_eta_calc": [],
    "lare_dev_energy_correction": [],
    "lare_dev_conduct_heat": [],
    "lare_dev_boundary_conditions": [],
    "lare_dev_neutral_fraction": [],
    "lare_dev_perpendicular_resistivity": [],
    "lare_dev_get_neutral": [_dp, _dp],
    "lare_dev_energy_account": [_dp, _dp],
    "lare_dev_step": [C.c_int32, _dp, _dp],
    "lare_dev_step_info": [C.c_int32, C.c_int32, C.POINTER(LareStepInfo)],
    "lare_dev_potential_field": [C.POINTER(LarePotential), C.POINTER(C.c_int32), _dp, _dp],
    "lare_dev_energy_account_noalloc": [C.POINTER(LareGrid), C.POINTER(LareFields), _dp, _dp],
    "lare_dev_profile": [C.c_int32],
    "lare_dev_kernel_stats": [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), _dp, C.c_int32],
    "lare_dev_sur_set_params": [_fp, C.c_int32],
    "lare_dev_sur_get_params": [_fp, C.c_int32],
    "lare_dev_sur_set_norm": [_fp, _fp, C.c_float],
    "lare_dev_sur_get_norm": [_fp, _fp, _fp],
    "lare_dev_sur_collect": [C.c_int32, C.c_uint64, C.POINTER(C.c_int32)],
    "lare_dev_sur_rows": [_fp, _fp, C.c_int32],
    "lare_dev_sur_fit_norm": [],
    "lare_dev_sur_train": [C.c_int32, C.c_int32, C.c_float, C.c_uint64, _fp],
    "lare_dev_sur_forward": [_fp, C.c_int32, _fp],
    "lare_dev_sur_grad": [_fp, _fp, C.c_int32, _fp, _fp],
    "lare_dev_sur_enable": [C.c_int32, C.c_double],
    "lare_dev_sur_conduct": [_dp, C.POINTER(C.c_int32)],
    "lare_dev_sur_plain_residual": [_dp],
    "lare_dev_sur_stats": [C.POINTER(C.c_int32), C.POINTER(C.c_int32), _dp],
    "lare_dev_synchronize": [],
    "lare_dev_finalize": [],
    "lare_dev_last_error": [],
}

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "liblare3d_dev.so")


class DeviceError(RuntimeError):
    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


_lib = None


def load_library(path=None):
    """Load the device library built in-tree.  Raises if it is missing:
    there is no CPU fallback for the product path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or os.environ.get("LARE3D_DEV_LIB") or LIB_PATH
    if not os.path.exists(p):
        raise DeviceError(f"device library not built: {p} (run __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, args in DEVICE_SYMBOLS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name == "lare_dev_last_error" else C.c_int
    if path is None:
        _lib = lib
    return lib


def check(lib, rc, what):
    if rc != 0:
        msg = lib.lare_dev_last_error().decode(errors="replace")
        raise DeviceError(f"{what} failed with status {rc}: {msg}", rc)
