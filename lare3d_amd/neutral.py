"""Normalised constants of the partially ionised equation of state.

The reference computes them once at start-up: setup_neutral
(neutral.f90:31-64) in SI units, then normalise_transport
(normalise.f90:36-52) divides by the normalising temperature, density and
resistivity.  The device takes the results through lare_params
(ionise_pot, f_bar, t_bar, r_bar, eta_bar, t_rad); this module is the host
side that fills them from the run's normalisation (B_norm, L_norm,
rho_norm and the mean mass factor mf of control.f90).
"""
import math

from .abi import EOS_ION, EOS_IDEAL, EOS_PI

# SI constants (CODATA 2010 values, as shared_data.F90:36-57 states them)
PI = 3.141592653589793
MU0_SI = 4.0e-7 * PI           # N / A^2
KB_SI = 1.3806488e-23          # J / K
MH_SI = 1.672621777e-27        # kg
ME_SI = 9.10938291e-31         # kg
HP_SI = 6.62606957e-34         # J s
IONISE_POT_SI = 2.17870364e-18  # J, hydrogen
SIGMA_IN = 5.0e-19             # ion-neutral cross section, m^2 (neutral.f90:34)
T_RAD_SI = 6420.0              # photospheric radiation temperature, K (neutral.f90:47)


def normalisation(B_norm, L_norm, rho_norm, mf):
    """energy0, v0, temp_norm, time_norm of normalise_transport
    (normalise.f90:36-43)."""
    pressure0 = B_norm ** 2 / MU0_SI
    energy0 = pressure0 / rho_norm
    v0 = math.sqrt(energy0)
    mbar = mf * MH_SI
    temp_norm = (mbar / KB_SI) * energy0
    return {"energy0": energy0, "v0": v0, "mbar": mbar, "temp_norm": temp_norm, "time_norm": L_norm / v0}


def neutral_constants(B_norm, L_norm, rho_norm, mf, eos_number):
    """The lare_params fields of the partially ionised EOS for one
    normalisation: setup_neutral's f_bar, t_bar, r_bar, eta_bar and tr
    (neutral.f90:46-61) with normalise_transport's scalings
    (normalise.f90:40-49).  ionise_pot is zero unless eos_number is
    EOS_ION (normalise.f90:48)."""
    if eos_number not in (EOS_IDEAL, EOS_ION, EOS_PI):
        raise ValueError(f"eos_number {eos_number}")
    nm = normalisation(B_norm, L_norm, rho_norm, mf)
    mbar, temp_norm, v0, energy0 = nm["mbar"], nm["temp_norm"], nm["v0"], nm["energy0"]
    f_bar = (2.0 * PI * (ME_SI / HP_SI) * (KB_SI / HP_SI)) ** 1.5
    t_bar = IONISE_POT_SI / KB_SI
    r_bar = 4.0 / mbar
    eta_bar = 2.0 * mbar / (math.sqrt(16.0 * KB_SI / (PI * mbar)) * SIGMA_IN)
    eta_bar_0 = (MU0_SI * L_norm * v0) * rho_norm ** 2 * math.sqrt(temp_norm) / B_norm ** 2
    return {
        "f_bar": f_bar,
        "t_bar": t_bar / temp_norm,
        "r_bar": r_bar * rho_norm / temp_norm ** 1.5,
        "eta_bar": eta_bar / eta_bar_0,
        "ionise_pot": IONISE_POT_SI / (energy0 * mbar) if eos_number == EOS_ION else 0.0,
        "t_rad": T_RAD_SI / temp_norm,
        "temp_norm": temp_norm,
    }


def set_neutral(params, B_norm, L_norm, rho_norm, mf, eos_number, cowling=0):
    """Fill params' partially ionised fields; returns temp_norm."""
    c = neutral_constants(B_norm, L_norm, rho_norm, mf, eos_number)
    params.eos_number = eos_number
    params.cowling_resistivity = cowling
    for k in ("f_bar", "t_bar", "r_bar", "eta_bar", "ionise_pot", "t_rad"):
        setattr(params, k, c[k])
    return c["temp_norm"]
