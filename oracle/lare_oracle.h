/*
 * lare_oracle.h -- CPU restatement of the Lare3d per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the device
 * implementation: only tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py may load it.  It is never part of the product path.
 *
 * Every routine restates the reference Fortran (file:line cited at each
 * function in lare_oracle.c) with the reference's loop ranges, loop order
 * and operation order, in FP64.  Arrays use the reference's index bounds;
 * storage is a common (-2:n+2)^3 box per array, and every element outside
 * the reference's allocated bounds of that array holds NaN, so that a read
 * outside the reference's bounds shows up as a non-finite result.
 *
 * The structs of include/lare3d_dev.h describe grid, parameters, fields and
 * scalars, so that a test hands the same inputs to the oracle and the device.
 */
#ifndef LARE_ORACLE_H
#define LARE_ORACLE_H

#include "../include/lare3d_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

orc_state *orc_create(const lare_grid *grid, const lare_params *params);
void orc_destroy(orc_state *o);
int orc_upload(orc_state *o, const lare_fields *f);
int orc_download(orc_state *o, lare_fields *f);
int orc_get_scalars(orc_state *o, lare_scalars *s);
int orc_set_scalars(orc_state *o, const lare_scalars *s);

void orc_lagrangian_step(orc_state *o);
void orc_eulerian_remap(orc_state *o, int step);
void orc_eta_calc(orc_state *o);
void orc_conduct_heat(orc_state *o);
void orc_boundary_conditions(orc_state *o);
void orc_energy_account(orc_state *o, double en[3]);
/* max over interior cells of |div B| * min(dxb,dyb,dzb) (device diagnostic) */
double orc_divb(orc_state *o);
/* correctly rounded t^2.5 and t^1.5 (the conduction temperature powers) */
double orc_pow25(double t);
double orc_pow15(double t);
void orc_step(orc_state *o, int step, double *dt, double *time);
/* diagnostics.F90:613-629 energy_correction (rke) */
void orc_energy_correction(orc_state *o);
/* delta_ke(-1:n+2)^3 (Fortran order) as it stands; -1 when rke is off */
int orc_get_delta_ke(orc_state *o, double *out);
/* neutral.f90:158-195 neutral_fraction and :68-136 perpendicular_resistivity;
 * LARE_E_STATE when the option is off; neutral_fraction returns
 * LARE_F_TBOUNDS once any cell's bracket was inverted */
int orc_neutral_fraction(orc_state *o);
int orc_perpendicular_resistivity(orc_state *o);
/* xi_n and eta_perp (0..n, zero elsewhere) in the (-1:n+2)^3 layout */
int orc_get_neutral(orc_state *o, double *xi_n, double *eta_perp);
/* LARE_F_TBOUNDS after an inverted bracket in any neutral_fraction, else 0 */
int orc_get_fault(orc_state *o);
/* the exp and get_neutral of the restatement (neutral.f90:140-154) */
double orc_exp(double x);
double orc_saha_neutral(const lare_params *p, double t_v, double rho_v);
/* initial_conditions.f90:114-296 potential_field (replaces bx, by, bz);
 * phi_out (may be NULL) receives phi(-1:nx+2, -1:ny+2, -1:nz+2) */
int orc_potential_field(orc_state *o, const lare_potential *pf, int *iterations, double *residual, double *bz_max,
                        double *phi_out);

#ifdef __cplusplus
}
#endif

#endif
