!  lare_gpu_bridge -- ISO_C_BINDING interface to liblare3d_dev.so
!  (include/lare3d_dev.h).  A Fortran driver USEs this module and calls the
!  device library with its own allocatable arrays (C_LOC of the first
!  element, i.e. with the ghost bounds of mpi_routines.F90:295-334).

MODULE lare_gpu_bridge

  USE, INTRINSIC :: ISO_C_BINDING

  IMPLICIT NONE

  INTEGER(C_INT32_T), PARAMETER :: LARE_BC_PERIODIC = 1, LARE_BC_USER = 2
  INTEGER(C_INT32_T), PARAMETER :: LARE_PREC_FP64 = 0, LARE_PREC_FP32 = 1

  TYPE, BIND(C) :: lare_grid
    INTEGER(C_INT32_T) :: nx, ny, nz
    INTEGER(C_INT32_T) :: xbc_min, xbc_max, ybc_min, ybc_max, zbc_min, zbc_max
    TYPE(C_PTR) :: dxb, dxc, dyb, dyc, dzb, dzc, grav
  END TYPE lare_grid

  TYPE, BIND(C) :: lare_params
    REAL(C_DOUBLE) :: gamma, visc1, visc2
    REAL(C_DOUBLE) :: eta_background, eta0, j_max
    REAL(C_DOUBLE) :: kappa_0, flux_limiter, xi_n, dt_from_restart
    INTEGER(C_INT32_T) :: resistive_mhd, conduction, heat_flux_limiter
    INTEGER(C_INT32_T) :: use_viscous_damping, restart, precision
    REAL(C_DOUBLE) :: alpha_av, va_max
    INTEGER(C_INT32_T) :: cooling_term, boris, shock_limiter, shock_compression, rke, cauchy, fourth_order
    INTEGER(C_INT32_T) :: eos_number, cowling_resistivity, pad2_
    REAL(C_DOUBLE) :: ionise_pot, f_bar, t_bar, r_bar, eta_bar, t_rad
  END TYPE lare_params

  TYPE, BIND(C) :: lare_fields
    TYPE(C_PTR) :: rho, energy, vx, vy, vz, bx, by, bz
    TYPE(C_PTR) :: visc3, eta, visc_dep, ohmic_dep
  END TYPE lare_fields

  TYPE, BIND(C) :: lare_scalars
    REAL(C_DOUBLE) :: dt, dtr, time
    REAL(C_DOUBLE) :: total_visc_heating, total_ohmic_heating
    INTEGER(C_INT32_T) :: n_s_stages, peak_substeps, last_substeps, first_set_dt, step
  END TYPE lare_scalars

  INTERFACE
    INTEGER(C_INT) FUNCTION lare_dev_init(grid, params) BIND(C, NAME='lare_dev_init')
      IMPORT :: C_INT, lare_grid, lare_params
      TYPE(lare_grid), INTENT(IN) :: grid
      TYPE(lare_params), INTENT(IN) :: params
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_upload(fields) BIND(C, NAME='lare_dev_upload')
      IMPORT :: C_INT, lare_fields
      TYPE(lare_fields), INTENT(IN) :: fields
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_download(fields) BIND(C, NAME='lare_dev_download')
      IMPORT :: C_INT, lare_fields
      TYPE(lare_fields), INTENT(INOUT) :: fields
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_get_scalars(s) BIND(C, NAME='lare_dev_get_scalars')
      IMPORT :: C_INT, lare_scalars
      TYPE(lare_scalars), INTENT(OUT) :: s
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_set_scalars(s) BIND(C, NAME='lare_dev_set_scalars')
      IMPORT :: C_INT, lare_scalars
      TYPE(lare_scalars), INTENT(IN) :: s
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_lagrangian_step() BIND(C, NAME='lare_dev_lagrangian_step')
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_eulerian_remap(step) BIND(C, NAME='lare_dev_eulerian_remap')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: step
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_eta_calc() BIND(C, NAME='lare_dev_eta_calc')
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_energy_correction() BIND(C, NAME='lare_dev_energy_correction')
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_conduct_heat() BIND(C, NAME='lare_dev_conduct_heat')
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_boundary_conditions() BIND(C, NAME='lare_dev_boundary_conditions')
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_neutral_fraction() BIND(C, NAME='lare_dev_neutral_fraction')
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_perpendicular_resistivity() &
        BIND(C, NAME='lare_dev_perpendicular_resistivity')
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_get_neutral(xi_n, eta_perp) BIND(C, NAME='lare_dev_get_neutral')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: xi_n, eta_perp
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_energy_account(en, divb) BIND(C, NAME='lare_dev_energy_account')
      IMPORT :: C_INT, C_DOUBLE
      REAL(C_DOUBLE), INTENT(OUT) :: en(3), divb
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_step(step, dt, time) BIND(C, NAME='lare_dev_step')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: step
      REAL(C_DOUBLE), INTENT(OUT) :: dt, time
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_synchronize() BIND(C, NAME='lare_dev_synchronize')
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_finalize() BIND(C, NAME='lare_dev_finalize')
      IMPORT :: C_INT
    END FUNCTION
    TYPE(C_PTR) FUNCTION lare_dev_last_error() BIND(C, NAME='lare_dev_last_error')
      IMPORT :: C_PTR
    END FUNCTION
    ! learned conduction surrogate (no counterpart in the reference)
    INTEGER(C_INT) FUNCTION lare_dev_sur_set_params(params, n) BIND(C, NAME='lare_dev_sur_set_params')
      IMPORT :: C_INT, C_INT32_T, C_FLOAT
      REAL(C_FLOAT), INTENT(IN) :: params(*)
      INTEGER(C_INT32_T), VALUE :: n
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_sur_get_params(params, n) BIND(C, NAME='lare_dev_sur_get_params')
      IMPORT :: C_INT, C_INT32_T, C_FLOAT
      REAL(C_FLOAT), INTENT(OUT) :: params(*)
      INTEGER(C_INT32_T), VALUE :: n
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_sur_set_norm(mu, sd, y_scale) BIND(C, NAME='lare_dev_sur_set_norm')
      IMPORT :: C_INT, C_FLOAT
      REAL(C_FLOAT), INTENT(IN) :: mu(*), sd(*)
      REAL(C_FLOAT), VALUE :: y_scale
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_sur_collect(count, seed, total) BIND(C, NAME='lare_dev_sur_collect')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T
      INTEGER(C_INT32_T), VALUE :: count
      INTEGER(C_INT64_T), VALUE :: seed
      INTEGER(C_INT32_T), INTENT(OUT) :: total
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_sur_fit_norm() BIND(C, NAME='lare_dev_sur_fit_norm')
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_sur_train(steps, batch, lr, seed, loss) BIND(C, NAME='lare_dev_sur_train')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_FLOAT
      INTEGER(C_INT32_T), VALUE :: steps, batch
      REAL(C_FLOAT), VALUE :: lr
      INTEGER(C_INT64_T), VALUE :: seed
      REAL(C_FLOAT), INTENT(OUT) :: loss(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_sur_enable(on, tol) BIND(C, NAME='lare_dev_sur_enable')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), VALUE :: on
      REAL(C_DOUBLE), VALUE :: tol
    END FUNCTION
    INTEGER(C_INT) FUNCTION lare_dev_sur_stats(accepted, rejected, last_residual) &
        BIND(C, NAME='lare_dev_sur_stats')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE
      INTEGER(C_INT32_T), INTENT(OUT) :: accepted, rejected
      REAL(C_DOUBLE), INTENT(OUT) :: last_residual
    END FUNCTION
  END INTERFACE

CONTAINS

  ! Text of the last library error as a Fortran string.
  FUNCTION lare_dev_error_text() RESULT(msg)
    CHARACTER(LEN=512) :: msg
    CHARACTER(KIND=C_CHAR), DIMENSION(:), POINTER :: cs
    INTEGER :: i
    CALL C_F_POINTER(lare_dev_last_error(), cs, [512])
    msg = ' '
    DO i = 1, 512
      IF (cs(i) == C_NULL_CHAR) EXIT
      msg(i:i) = cs(i)
    END DO
  END FUNCTION lare_dev_error_text

END MODULE lare_gpu_bridge
