/*
 * lare3d_dev.h -- C-ABI of the MI355X device implementation of the Lare3d
 * per-step hot path (Lagrangian predictor-corrector with edge shock
 * viscosity, Van Leer x/y/z remap with constrained-transport face-B remap,
 * RKL2 Braginskii conduction, boundary/halo fill, CFL timestep reduction,
 * per-step energy diagnostics).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an int status: 0 = ok, < 0 = -(hipError_t) or a library error
 * (see LARE_E_*), > 0 = a physics fault detected on the device.  The text
 * of the last error is available from lare_dev_last_error().
 *
 * Array convention (the data contract of the reference): host arrays are
 * Fortran column-major (x fastest) and are passed as a pointer to their
 * first element, i.e. C_LOC(a) of the allocatable with its ghost bounds as
 * allocated in mpi_routines.F90:295-334:
 *   cell  (rho, energy, eta, visc_dep, ohmic_dep): (-1:nx+2, -1:ny+2, -1:nz+2)
 *   node  (vx, vy, vz, vx1.., visc3)            : (-2:nx+2, -2:ny+2, -2:nz+2)
 *   bx (-2:nx+2, -1:ny+2, -1:nz+2), by (-1:nx+2, -2:ny+2, -1:nz+2),
 *   bz (-1:nx+2, -1:ny+2, -2:nz+2)
 *   1-D metrics dxb(-1:nx+2), dxc(-1:nx+1) (same for y, z), grav(-1:nz+2).
 *
 * Threading: the caller is single threaded (one process per GPU).  The
 * library is not re-entrant; all work is queued on one HIP stream.
 */
#ifndef LARE3D_DEV_H
#define LARE3D_DEV_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary codes, as constants.f90 of the reference (shared_data.F90:65-66). */
#define LARE_BC_PERIODIC 1
#define LARE_BC_USER 2
#define LARE_BC_OPEN 3   /* far-field characteristic boundary (openboundary.f90) */

/* Precision modes of the bulk fields. */
#define LARE_PREC_FP64 0
#define LARE_PREC_FP32 1

/* Library error codes (negative, disjoint from -hipError_t which is > -1000). */
#define LARE_E_STATE   -1001 /* call order: not initialised / no fields     */
#define LARE_E_ARG     -1002 /* invalid argument / unsupported option       */
#define LARE_E_NOMEM   -1003 /* device allocation failed                    */
/* Physics faults (positive). */
#define LARE_F_NONFINITE 1   /* non-finite dt, or a state with a non-finite
                                value or rho <= 0 entering a step            */
#define LARE_F_TBOUNDS 2     /* neutral_fraction's temperature bracket is
                                inverted in some cell (energy <= 0): the
                                reference STOPs (neutral.f90:172-175)        */

/* Equations of state (shared_data.F90:73, control.f90:187-199). */
#define LARE_EOS_IDEAL 1     /* xi_n uniform (lare_params.xi_n); 0 also
                                means EOS_IDEAL                              */
#define LARE_EOS_ION 2       /* partially ionised, with ionisation energy    */
#define LARE_EOS_PI 3        /* partially ionised, ionise_pot = 0            */

/* Grid: sizes, boundary codes and the 1-D metrics of setup.F90:92-247.
 * Replaces the module state nx/ny/nz, x/y/zbc_*, dxb/dxc/dyb/dyc/dzb/dzc
 * and grav of shared_data.F90:96-116,135-137. */
typedef struct lare_grid {
  int32_t nx, ny, nz;
  int32_t xbc_min, xbc_max, ybc_min, ybc_max, zbc_min, zbc_max;
  const double *dxb, *dxc; /* dxb(-1:nx+2), dxc(-1:nx+1) */
  const double *dyb, *dyc;
  const double *dzb, *dzc;
  const double *grav;      /* grav(-1:nz+2); NULL means zero gravity */
} lare_grid;

/* Run parameters of control.f90 (control_variables) and normalise.f90.
 * With EOS_IDEAL xi_n is the uniform value of lare3d.f90:72 and the
 * ionisation potential is 0 (normalise.f90:48); EOS_ION and EOS_PI keep a
 * neutral fraction per cell (neutral.f90:158-195) from the normalised
 * constants of setup_neutral and normalise_transport (neutral.f90:31-64,
 * normalise.f90:36-52), which the caller computes once. */
typedef struct lare_params {
  double gamma;            /* setup.F90:46 default 5/3                    */
  double visc1, visc2;     /* shock viscosity coefficients               */
  double eta_background, eta0, j_max; /* resistivity (lagran.F90:1176)  */
  double kappa_0;          /* parallel conductivity (normalise.f90:54-55) */
  double flux_limiter;     /* Larsen flux limiter (conduct.f90:90)        */
  double xi_n;             /* neutral fraction, uniform (0 or 1)          */
  double dt_from_restart;  /* used by the first set_dt when restart != 0  */
  int32_t resistive_mhd;
  int32_t conduction;
  int32_t heat_flux_limiter;
  int32_t use_viscous_damping;
  int32_t restart;
  int32_t precision;       /* LARE_PREC_FP64 or LARE_PREC_FP32            */
  double alpha_av;         /* cooling_term averaging weight (control.f90:206) */
  double va_max;           /* Boris Alfven speed limit (control.f90:184)  */
  int32_t cooling_term;    /* lagran.F90:464-476, :1290-1302              */
  int32_t boris;           /* lagran.F90:1092-1108, :375-377, :556        */
  int32_t shock_limiter;   /* the SHOCKLIMITER variant of edge_viscosity
                              (lagran.F90:801-803, :831-845)              */
  int32_t shock_compression; /* the SHOCKCOMPRESSION variant (:807-808)   */
  int32_t rke;             /* remap kinetic-energy correction (control.f90:148;
                              xremap.f90:664-695, diagnostics.F90:613-629) */
  int32_t cauchy;          /* the CAUCHY predictor: cell-centred B advanced by
                              the Cauchy solution instead of the predictor
                              remap (lagran.F90:285-348, :945-1048) */
  int32_t fourth_order;    /* FOURTHORDER: resistive substeps by classical
                              fourth-order Runge-Kutta (lagran.F90:1329-1415) */
  int32_t eos_number;      /* LARE_EOS_* (control.f90:191)                 */
  int32_t cowling_resistivity; /* Cowling eta_perp (control.f90:162,
                              neutral.f90:68-136); needs eos_number
                              EOS_ION or EOS_PI                           */
  int32_t pad2_;
  double ionise_pot;       /* normalised ionisation potential (used with
                              EOS_ION only, normalise.f90:47-48)           */
  double f_bar;            /* (2 pi m_e k_B / h^2)^1.5, SI (neutral.f90:50-51) */
  double t_bar;            /* ionisation temperature / temp_norm (:54, normalise.f90:40) */
  double r_bar;            /* 4 / mbar, normalised (:57-58, normalise.f90:41) */
  double eta_bar;          /* Cowling coefficient, normalised (:61, normalise.f90:42-43) */
  double t_rad;            /* radiation temperature tr / temp_norm (:47, normalise.f90:49) */
} lare_params;

/* Field pointers (Fortran arrays, see the array convention above).
 * visc3 may be NULL when use_viscous_damping is 0; visc_dep/ohmic_dep/eta
 * may be NULL on upload (then they start at zero). */
typedef struct lare_fields {
  double *rho, *energy;
  double *vx, *vy, *vz;
  double *bx, *by, *bz;
  double *visc3;
  double *eta, *visc_dep, *ohmic_dep;
} lare_fields;

/* Scalar state of shared_data.F90:126-139 that the step reads and writes. */
typedef struct lare_scalars {
  double dt, dtr, time;
  double total_visc_heating, total_ohmic_heating;
  int32_t n_s_stages;      /* last RKL2 stage count                       */
  int32_t peak_substeps;
  int32_t last_substeps;   /* resistive substeps of the last step         */
  int32_t first_set_dt;    /* 1 until set_dt has run once (restart logic) */
  int32_t step;            /* step counter of lare3d.f90:82 (Strang order) */
} lare_scalars;

/* Scalars of one step, as lare_dev_step_info returns them. */
typedef struct lare_step_info {
  double dt, dtr, time;    /* shared_data.F90:126-139                    */
  double total_visc_heating, total_ohmic_heating;
  double en_b, en_ke, en_int; /* energy_account of the new state         */
  double max_divb;         /* max |div B| dx of the new state             */
  int32_t n_s_stages;      /* RKL2 stages of the step (conduction)        */
  int32_t substeps;        /* resistive substeps of the step              */
  int32_t peak_substeps;
  int32_t sur_accepted, sur_rejected; /* surrogate super-steps so far     */
  int32_t with_energy;     /* 1 if the energy fields are set              */
} lare_step_info;

/* Inputs of potential_field (initial_conditions.f90:114-296): red-black
 * SOR for the scalar potential phi of a current-free field, B = -grad phi,
 * scaled so that the largest bz on the lower z boundary is 1. */
typedef struct lare_potential {
  double omega;            /* over-relaxation w (:131, 1.6)                */
  double fractional_error; /* stop when max |residual| < this (:132, 5e-5) */
  double dx1, dy1, dz1;    /* (n_global / length)^2 per axis (:126-128)    */
  const double *flux0;     /* nx*ny host values, x fastest: the lower
                              boundary sets phi(ix,iy,0) = phi(ix,iy,1) +
                              flux0[(ix-1) + nx*(iy-1)] (:285-293, where
                              the reference's example is amp*dzc(1)*
                              (1 - tanh((r - radius)/0.2)))             */
  int32_t max_iterations;  /* iteration limit (:135, 1000000)              */
  int32_t pad_;
} lare_potential;

/* Initialise the device state for one grid (allocates all device memory).
 * Replaces the allocation part of mpi_initialise (mpi_routines.F90:295-334)
 * and grid() for the device copy.  Re-initialising frees the old state. */
int lare_dev_init(const lare_grid *grid, const lare_params *params);

/* Copy host fields to the device (after set_initial_conditions). */
int lare_dev_upload(const lare_fields *fields);

/* Copy device fields to host arrays (dump steps only). */
int lare_dev_download(lare_fields *fields);

/* Scalar state access (dt, time, heating totals ...). */
int lare_dev_get_scalars(lare_scalars *out);
int lare_dev_set_scalars(const lare_scalars *in);

/* lagran.F90:50 lagrangian_step (incl. set_dt :1070, conduct_heat, resistive
 * substeps and the predictor/corrector). */
int lare_dev_lagrangian_step(void);

/* remap.f90:33 eulerian_remap(i) -- Strang-ordered x/y/z Van Leer remap. */
int lare_dev_eulerian_remap(int32_t step);

/* lagran.F90:1176 eta_calc. */
int lare_dev_eta_calc(void);

/* diagnostics.F90:613-629 energy_correction (called by lare3d.f90:86 after
 * eulerian_remap when rke): the kinetic energy the last remap lost added to
 * the internal energy, then energy_bcs.  LARE_E_STATE when rke is off. */
int lare_dev_energy_correction(void);

/* conduct.f90:39 conduct_heat -- one RKL2 super-step of length dt. */
int lare_dev_conduct_heat(void);

/* boundary.f90:48 boundary_conditions (B, energy, density, velocity). */
int lare_dev_boundary_conditions(void);

/* neutral.f90:158-195 neutral_fraction: the neutral fraction xi_n of every
 * cell (-1..n+2) from energy and rho, by the reference's bisection for the
 * temperature.  lare_dev_step calls it before each lagrangian_step
 * (lare3d.f90:83), lagrangian_step before each resistive substep
 * (lagran.F90:147) and conduct_heat between RKL2 stages (conduct.f90:404);
 * the caller calls it once after the initial boundary_conditions
 * (lare3d.f90:70).  LARE_E_STATE with EOS_IDEAL, LARE_F_TBOUNDS for an
 * inverted bracket. */
int lare_dev_neutral_fraction(void);

/* neutral.f90:68-136 perpendicular_resistivity: the Cowling resistivity
 * eta_perp at nodes 0..n from rho, energy, xi_n and B.  lagrangian_step
 * calls it before each resistive substep (lagran.F90:148); the caller calls
 * it once at start-up (lare3d.f90:74).  LARE_E_STATE without
 * cowling_resistivity. */
int lare_dev_perpendicular_resistivity(void);

/* Copy xi_n (cell layout (-1:nx+2, -1:ny+2, -1:nz+2)) and eta_perp (the
 * same layout; set at 0..n, zero elsewhere) to host arrays; either may be
 * NULL.  The arrays exist with EOS_ION/EOS_PI and cowling_resistivity
 * (the dump variables of diagnostics.F90:377-394); LARE_E_STATE when a
 * requested one does not. */
int lare_dev_get_neutral(double *xi_n, double *eta_perp);

/* diagnostics.F90:517 energy_account: en[0] = magnetic, en[1] = kinetic,
 * en[2] = internal (FP64 sums).  divb (may be NULL) receives max |div B|
 * over interior cells, scaled by the cell width (a device diagnostic with no
 * reference counterpart). */
int lare_dev_energy_account(double en[3], double *divb);

/* energy_account on the caller's own device arrays, allocating nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same results as lare_dev_energy_account
 * for the same field values, and needs no lare_dev_init.  dev_fields holds
 * DEVICE pointers to FP64 arrays in the Fortran layout above (rho, energy,
 * vx, vy, vz, bx, by, bz; the others are ignored); grid supplies the sizes
 * and the host metrics dxb, dyb, dzb.  Partial sums live in a fixed device
 * array of about 8.1 MiB that each GPU holds until the process ends: at most
 * 4096 cells per axis and 262144 reduction blocks ((nx+1)(ny+1)/256 rounded
 * up, times nz+1 -- about 400^3 cells); larger grids return LARE_E_ARG before
 * any device call.  Calls that overlap on one GPU share that array and race.
 * Runs on the null stream and returns after the sums reach the host. */
int lare_dev_energy_account_noalloc(const lare_grid *grid, const lare_fields *dev_fields, double en[3],
                                    double *divb);

/* One full iteration of the lare3d.f90:82-88 loop body:
 * lagrangian_step, eulerian_remap(step), energy_correction when rke is on,
 * eta_calc.  Returns dt and time. */
int lare_dev_step(int32_t step, double *dt, double *time);

/* initial_conditions.f90:114 potential_field on the device: replaces the
 * face B of the loaded state by the scaled potential field.  Needs a
 * BC_USER z axis (the flux patch sits on the lower z boundary) and
 * nz <= ny + 2 (the reference's search for the largest bz runs over
 * bz(1:nx, 1:nz, 0), :221).  *iterations receives the SOR iterations done,
 * *residual the max |residual| of the last one, *bz_max the scale. */
int lare_dev_potential_field(const lare_potential *pf, int32_t *iterations, double *residual, double *bz_max);

/* lare_dev_step, then the scalars the driver reads after a step (dt, dtr,
 * time, heating totals, stage and substep counts, surrogate counts) and,
 * when with_energy != 0, energy_account (diagnostics.F90:517) and max
 * |div B| of the new state, read back together in one device-to-host copy.
 * LARE_F_NONFINITE reports a state that holds a non-finite value or
 * rho <= 0 (the field and cell are in lare_dev_last_error) or a non-finite
 * dt. */
int lare_dev_step_info(int32_t step, int32_t with_energy, lare_step_info *info);

/* Per-kernel timing with HIP events on the library stream (for roofline
 * measurement).  enable != 0 starts recording; lare_dev_kernel_stats fills
 * up to cap entries of (name, count, total_ms) and returns the number. */
int lare_dev_profile(int32_t enable);
int lare_dev_kernel_stats(char *names, int32_t name_len, int32_t *counts,
                          double *total_ms, int32_t cap);

/* ------------------------------------------------------------------------
 * Learned conduction surrogate.  No counterpart in the reference: it stands
 * in for whole conduct_heat super-steps (conduct.f90:39) and falls back to
 * them.  A 3x3x3-stencil MLP (27 cells x 10 channels -> 64 -> 64 -> 1, GELU)
 * predicts the temperature increment of one RKL2 super-step; BF16 MFMA with
 * FP32 accumulation for inference and training.  Each use is checked by the
 * trapezoidal residual |E1 - E0 - dt/2 (L(E0) + L(E1))| / |E0| (L2 over the
 * interior, L = heat-flux divergence / rho); above the tolerance the step is
 * redone by the plain RKL2 solver.
 *
 * Parameter vector (float, LARE_SUR_NPARAM): W1[336][64] (rows 324..335 and
 * the rows of channels 10, 11 of each cell meet zero inputs), b1[64],
 * W2[64][64], b2[64], W3[64], b3.  Inputs per row: 27 neighbours (dz, dy, dx
 * in -1..1, x fastest) x 12 channels (T, unit B x3, rho, log dt, |B|,
 * dT/dx, dT/dy, dT/dz, 0, 0), standardised, then 12 zeros (336 floats). */
#define LARE_SUR_K 336
#define LARE_SUR_NCH 10
#define LARE_SUR_NPARAM (336 * 64 + 64 + 64 * 64 + 64 + 64 + 1)

/* Load / read the FP32 master parameters (n == LARE_SUR_NPARAM).  Loading
 * resets the Adam moments. */
int lare_dev_sur_set_params(const float *params, int32_t n);
int lare_dev_sur_get_params(float *params, int32_t n);
/* Channel standardisation (mean and std of the 10 channels) and the scale of
 * the network output (dT = y * y_scale). */
int lare_dev_sur_set_norm(const float *mu, const float *sd, float y_scale);
int lare_dev_sur_get_norm(float *mu, float *sd, float *y_scale);
/* Training data from the current device state and dt (set_scalars): count
 * random interior cells (hash of seed), their raw input rows, and the
 * temperature increment one plain RKL2 super-step gives them.  The state is
 * left unchanged.  *total receives the number of rows held. */
int lare_dev_sur_collect(int32_t count, uint64_t seed, int32_t *total);
/* Copy the first `rows` held rows (LARE_SUR_K floats each) and targets to
 * the host (raw before lare_dev_sur_fit_norm, standardised after). */
int lare_dev_sur_rows(float *X, float *y, int32_t rows);
/* Standardisation from the collected rows (centre-cell channel statistics,
 * with the std of the log dt channel floored at 1; RMS target); standardises
 * the held rows. */
int lare_dev_sur_fit_norm(void);
/* Adam (lr, beta 0.9 / 0.999) for `steps` mini-batches of `batch` rows
 * (a multiple of 16) drawn by hash of seed; loss_out[steps] receives the
 * mean squared error of each mini-batch (may be NULL). */
int lare_dev_sur_train(int32_t steps, int32_t batch, float lr, uint64_t seed, float *loss_out);
/* Network output for `rows` standardised input rows (LARE_SUR_K floats each). */
int lare_dev_sur_forward(const float *X, int32_t rows, float *y);
/* Gradient of the mean squared error against targets t over the rows. */
int lare_dev_sur_grad(const float *X, const float *t, int32_t rows, float *grad, float *loss);
/* Use the surrogate for the conduction of lagrangian_step (on != 0), with the
 * relative residual tolerance tol. */
int lare_dev_sur_enable(int32_t on, double tol);
/* One checked surrogate super-step of length dt on the current state (the
 * surrogate form of lare_dev_conduct_heat).  *accepted = 0 means the
 * residual exceeded the tolerance and RKL2 did the step. */
int lare_dev_sur_conduct(double *residual, int32_t *accepted);
/* The same residual check applied to the plain RKL2 super-step from the
 * current state (the state is left unchanged): the floor of the check. */
int lare_dev_sur_plain_residual(double *residual);
/* Counts of accepted and rejected surrogate steps, and the last residual. */
int lare_dev_sur_stats(int32_t *accepted, int32_t *rejected, double *last_residual);

/* Block until all queued device work is done. */
int lare_dev_synchronize(void);

/* Free all device memory. */
int lare_dev_finalize(void);

/* Text of the last error (static storage, never NULL). */
const char *lare_dev_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LARE3D_DEV_H */
