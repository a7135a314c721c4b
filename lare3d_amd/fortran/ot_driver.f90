!  Orszag-Tang driver through lare_gpu_bridge: the loop of lare3d.f90:80-89
!  with lagrangian_step / eulerian_remap / eta_calc replaced by the device
!  library.  Arrays are allocated with the reference's bounds
!  (mpi_routines.F90:295-334) and the uniform periodic grid of
!  setup.F90:92-247.  Usage: ot_driver n nsteps
PROGRAM ot_driver

  USE, INTRINSIC :: ISO_C_BINDING
  USE lare_gpu_bridge

  IMPLICIT NONE

  INTEGER :: n, nsteps, step, ix, iy, iz, rc, i
  CHARACTER(LEN=32) :: arg
  REAL(C_DOUBLE), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: rho, energy, vx, vy, vz, bx, by, bz
  REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: xb, xc, dxb, dxc, grav
  REAL(C_DOUBLE) :: dt, time, en(3), divb, d, pi2
  TYPE(lare_grid) :: grid
  TYPE(lare_params) :: params
  TYPE(lare_fields) :: fields

  CALL GET_COMMAND_ARGUMENT(1, arg); READ(arg, *) n
  CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg, *) nsteps

  ALLOCATE(rho(-1:n+2,-1:n+2,-1:n+2), energy(-1:n+2,-1:n+2,-1:n+2))
  ALLOCATE(vx(-2:n+2,-2:n+2,-2:n+2), vy(-2:n+2,-2:n+2,-2:n+2), vz(-2:n+2,-2:n+2,-2:n+2))
  ALLOCATE(bx(-2:n+2,-1:n+2,-1:n+2), by(-1:n+2,-2:n+2,-1:n+2), bz(-1:n+2,-1:n+2,-2:n+2))
  ALLOCATE(xb(-2:n+2), xc(-1:n+2), dxb(-1:n+2), dxc(-1:n+1), grav(-1:n+2))

  ! periodic uniform grid on [0, 2 pi] (setup.F90:106-149)
  pi2 = 8.0_C_DOUBLE * ATAN(1.0_C_DOUBLE)
  d = pi2 / REAL(n, C_DOUBLE)
  DO i = -2, n + 2
    xb(i) = REAL(i, C_DOUBLE) * d
  END DO
  xb(n+1) = xb(n) + (xb(1) - xb(0))
  xb(n+2) = xb(n) + (xb(2) - xb(0))
  xb(-1) = xb(0) - (xb(n) - xb(n-1))
  xb(-2) = xb(0) - (xb(n) - xb(n-2))
  DO i = -1, n + 2
    xc(i) = 0.5_C_DOUBLE * (xb(i-1) + xb(i))
    dxb(i) = xb(i) - xb(i-1)
  END DO
  DO i = -1, n + 1
    dxc(i) = xc(i+1) - xc(i)
  END DO
  grav = 0.0_C_DOUBLE

  ! Orszag-Tang: rho = 25/9, p = 5/3, v = (-sin y, sin x, 0.1 sin z), B = (-sin y, sin 2x, 0)
  rho = 25.0_C_DOUBLE / 9.0_C_DOUBLE
  energy = (5.0_C_DOUBLE / 3.0_C_DOUBLE) / ((5.0_C_DOUBLE / 3.0_C_DOUBLE - 1.0_C_DOUBLE) * rho)
  DO iz = -2, n + 2
    DO iy = -2, n + 2
      DO ix = -2, n + 2
        vx(ix,iy,iz) = -SIN(xb(iy))
        vy(ix,iy,iz) = SIN(xb(ix))
        vz(ix,iy,iz) = 0.1_C_DOUBLE * SIN(xb(iz))
      END DO
    END DO
  END DO
  DO iz = -1, n + 2
    DO iy = -1, n + 2
      DO ix = -2, n + 2
        bx(ix,iy,iz) = -SIN(xc(iy))
      END DO
    END DO
  END DO
  DO iz = -1, n + 2
    DO iy = -2, n + 2
      DO ix = -1, n + 2
        by(ix,iy,iz) = SIN(2.0_C_DOUBLE * xc(ix))
      END DO
    END DO
  END DO
  bz = 0.0_C_DOUBLE

  grid%nx = n; grid%ny = n; grid%nz = n
  grid%xbc_min = LARE_BC_PERIODIC; grid%xbc_max = LARE_BC_PERIODIC
  grid%ybc_min = LARE_BC_PERIODIC; grid%ybc_max = LARE_BC_PERIODIC
  grid%zbc_min = LARE_BC_PERIODIC; grid%zbc_max = LARE_BC_PERIODIC
  grid%dxb = C_LOC(dxb); grid%dxc = C_LOC(dxc)
  grid%dyb = C_LOC(dxb); grid%dyc = C_LOC(dxc)
  grid%dzb = C_LOC(dxb); grid%dzc = C_LOC(dxc)
  grid%grav = C_LOC(grav)

  params%gamma = 5.0_C_DOUBLE / 3.0_C_DOUBLE
  params%visc1 = 0.1_C_DOUBLE; params%visc2 = 0.5_C_DOUBLE
  params%eta_background = 0.0_C_DOUBLE; params%eta0 = 0.0_C_DOUBLE; params%j_max = 1.0e10_C_DOUBLE
  params%kappa_0 = 0.0_C_DOUBLE; params%flux_limiter = 0.1_C_DOUBLE; params%xi_n = 0.0_C_DOUBLE
  params%dt_from_restart = 0.0_C_DOUBLE
  params%resistive_mhd = 0; params%conduction = 0; params%heat_flux_limiter = 1
  params%use_viscous_damping = 0; params%restart = 0; params%precision = LARE_PREC_FP64
  params%alpha_av = 0.05_C_DOUBLE; params%va_max = 10.0_C_DOUBLE
  params%cooling_term = 0; params%boris = 0; params%shock_limiter = 0; params%shock_compression = 0
  params%rke = 0; params%cauchy = 0; params%fourth_order = 0
  params%eos_number = 1; params%cowling_resistivity = 0; params%pad2_ = 0
  params%ionise_pot = 0.0_C_DOUBLE; params%f_bar = 0.0_C_DOUBLE; params%t_bar = 0.0_C_DOUBLE
  params%r_bar = 0.0_C_DOUBLE; params%eta_bar = 0.0_C_DOUBLE; params%t_rad = 0.0_C_DOUBLE

  fields%rho = C_LOC(rho); fields%energy = C_LOC(energy)
  fields%vx = C_LOC(vx); fields%vy = C_LOC(vy); fields%vz = C_LOC(vz)
  fields%bx = C_LOC(bx); fields%by = C_LOC(by); fields%bz = C_LOC(bz)
  fields%visc3 = C_NULL_PTR; fields%eta = C_NULL_PTR
  fields%visc_dep = C_NULL_PTR; fields%ohmic_dep = C_NULL_PTR

  rc = lare_dev_init(grid, params);  CALL check(rc, 'init')
  rc = lare_dev_upload(fields);      CALL check(rc, 'upload')
  rc = lare_dev_boundary_conditions(); CALL check(rc, 'boundary_conditions')
  rc = lare_dev_eta_calc();          CALL check(rc, 'eta_calc')

  DO step = 1, nsteps
    rc = lare_dev_step(step, dt, time); CALL check(rc, 'step')
    rc = lare_dev_energy_account(en, divb); CALL check(rc, 'energy_account')
    WRITE(*, '(A,I5,5ES24.16)') 'step', step, dt, time, en(1), en(2), en(3)
  END DO
  rc = lare_dev_download(fields); CALL check(rc, 'download')
  WRITE(*, '(A,ES24.16)') 'rho_sum', SUM(rho(1:n,1:n,1:n))
  rc = lare_dev_finalize()

CONTAINS

  SUBROUTINE check(code, what)
    INTEGER(C_INT), INTENT(IN) :: code
    CHARACTER(LEN=*), INTENT(IN) :: what
    IF (code /= 0) THEN
      WRITE(*, '(A,A,A,I0,A,A)') 'lare_dev ', what, ' failed: ', code, ' ', TRIM(lare_dev_error_text())
      STOP 1
    END IF
  END SUBROUTINE check

END PROGRAM ot_driver
