"""MI355X device implementation of the Lare3d per-step hot path.

The product is the C-ABI library liblare3d_dev.so (include/lare3d_dev.h)
built from lare3d_amd/csrc; this package is the host-side mirror of the
reference's module interface (lagrangian_step, eulerian_remap, eta_calc,
conduct_heat, boundary_conditions, energy_account) around that library.
"""
from .abi import BC_OPEN, BC_PERIODIC, BC_USER, PREC_FP32, PREC_FP64  # noqa: F401
from .fields import Fields  # noqa: F401
from .grid import Grid  # noqa: F401
