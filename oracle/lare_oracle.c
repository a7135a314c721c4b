/*
 * lare_oracle.c -- CPU restatement (FP64) of the Lare3d per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device implementation.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
 * it.  Each function cites the reference routine it restates (paths are
 * relative to the reference's src/ directory).  Loop ranges, nesting and the
 * order of floating-point operations follow the reference; compile with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Scope (as the device implementation): the three equations of state of
 * control.f90:187-199 (EOS_IDEAL with a uniform xi_n; EOS_ION and EOS_PI
 * with the per-cell neutral fraction and the Cowling resistivity of
 * neutral.f90), the run-time options of lare_params (boris, cooling_term,
 * rke, the CAUCHY, SHOCKLIMITER, SHOCKCOMPRESSION and FOURTHORDER
 * variants); each axis is either periodic on both sides or has BC_USER
 * or BC_OPEN on each side (openboundary.f90); one MPI rank (so every MPI_SENDRECV is
 * the periodic self-exchange of mpiboundary.f90).
 */
#include "lare_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define NONE_ZERO DBL_MIN   /* TINY(1.0_num), shared_data.F90:62 */
#define LARGEST DBL_MAX     /* HUGE(1.0_num), shared_data.F90:63 */
#define DT_MULT 0.9         /* dt_multiplier, shared_data.F90:60 */
#define SIXTH (1.0 / 6.0)   /* shared_data.F90:64 */
#define MIN_B 1.0e-10       /* conduct.f90:28 */
#define POW 2.5             /* conduct.f90:27 */

/* temperature**pow (conduct.f90:27, pow = 2.5) and temperature**1.5
 * (conduct.f90:83) evaluated correctly rounded: t^2 and t*sqrt(t) are
 * carried in double-double (FMA-exact products, sqrt(t) = r + (t - r^2) /
 * (2r)) and rounded once, so the result is the double nearest the exact
 * value (the relative error before the final rounding is near 2^-105).
 * A libm pow is not always that: glibc's is one ULP off in about 0.08% of
 * arguments.  The device computes the same operations. */
double orc_pow25(double t) {
  if (!(t > 0.0) || t == INFINITY) return t == 0.0 ? 0.0 : (t > 0.0 ? t : NAN);
  const double r = sqrt(t);
  const double e = fma(-r, r, t);
  const double rl = e / (2.0 * r);
  const double h2 = t * t;
  const double l2 = fma(t, t, -h2);
  const double p = h2 * r;
  if (!(p < INFINITY)) return p;
  const double pl = fma(h2, r, -p);
  return p + (pl + (h2 * rl + l2 * r));
}
double orc_pow15(double t) {
  if (!(t > 0.0) || t == INFINITY) return t == 0.0 ? 0.0 : (t > 0.0 ? t : NAN);
  const double r = sqrt(t);
  const double e = fma(-r, r, t);
  const double rl = e / (2.0 * r);
  const double p = t * r;
  if (!(p < INFINITY)) return p;
  const double pl = fma(t, r, -p);
  return p + (pl + t * rl);
}

struct orc_state {
  int nx, ny, nz;
  long ex, ey, ez, ntot;
  int bc[6];               /* x_min, x_max, y_min, y_max, z_min, z_max */
  int per[3];              /* axis periodic */
  lare_params p;
  double *m[7];            /* dxb, dxc, dyb, dyc, dzb, dzc, grav ; index i+2 */
  /* shared_data fields */
  double *rho, *energy, *vx, *vy, *vz, *vx1, *vy1, *vz1, *bx, *by, *bz;
  double *eta, *cv, *cv1, *visc3, *visc_dep, *ohmic_dep, *p_visc, *temperature;
  double *larsen;          /* larsen_factor(0:nx,0:ny,0:nz), conduct.f90:41 */
  double *gamma_boris;     /* gamma_boris(-1:n+2)^3, mpi_routines.F90:311 */
  double *cool_term_v, *cool_term_b; /* (1:n)^3, mpi_routines.F90:315-319 */
  double *delta_ke;        /* delta_ke(-1:n+2)^3 when rke, mpi_routines.F90:313 */
  double *xi_n;            /* xi_n(-1:n+2)^3 with EOS_ION / EOS_PI (neutral.f90:37) */
  double *eta_perp;        /* eta_perp, set at 0..n, with cowling (neutral.f90:41) */
  double ip;               /* ionise_pot: 0 unless EOS_ION (normalise.f90:48) */
  int pi_eos;              /* EOS_ION or EOS_PI */
  int any_open;            /* some side is BC_OPEN (boundary.f90:35-38) */
  int fault;               /* LARE_F_TBOUNDS once neutral_fraction's bracket failed */
  /* lagran module arrays live only within one lagrangian_step */
  double *bx1, *by1, *bz1, *pressure, *rho_vl, *cv_vl, *visc_heat;
  double *alpha1, *alpha2, *alpha3, *fx_visc, *fy_visc, *fz_visc;
  /* remap module arrays live only within one sweep */
  double *flux, *dxb1, *dxc1, *cv2, *rho1, *dm, *rho_v, *rho_v1;
  /* scalars */
  double dt, dt2, dtr, time, visc2_norm, tvh, toh;
  int n_s_stages, peak_substeps, last_substeps, first_set_dt, first_shock;
  int predictor_step, xpass, ypass, zpass, step;
};

#define IDX(i, j, k) ((((long)(k) + 2) * o->ey + ((long)(j) + 2)) * o->ex + ((long)(i) + 2))
#define A(a, i, j, k) (a)[IDX(i, j, k)]
#define DXB(i) o->m[0][(i) + 2]
#define DXC(i) o->m[1][(i) + 2]
#define DYB(i) o->m[2][(i) + 2]
#define DYC(i) o->m[3][(i) + 2]
#define DZB(i) o->m[4][(i) + 2]
#define DZC(i) o->m[5][(i) + 2]
#define GRAV(i) o->m[6][(i) + 2]
#define SQ(x) ((x) * (x))

static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double min3(double a, double b, double c) { return dmin(dmin(a, b), c); }

/* Allocate a (-2:n+2)^3 box; NaN outside [lo,hi] per axis, zero inside. */
static double *box(orc_state *o, int lx, int hx, int ly, int hy, int lz, int hz) {
  double *a = (double *)malloc(sizeof(double) * o->ntot);
  for (long q = 0; q < o->ntot; ++q) a[q] = NAN;
  for (int k = lz; k <= hz; ++k)
    for (int j = ly; j <= hy; ++j)
      for (int i = lx; i <= hx; ++i) A(a, i, j, k) = 0.0;
  return a;
}
static double *cellbox(orc_state *o) { return box(o, -1, o->nx + 2, -1, o->ny + 2, -1, o->nz + 2); }
static double *nodebox(orc_state *o) { return box(o, -2, o->nx + 2, -2, o->ny + 2, -2, o->nz + 2); }
static void drop(double **a) { free(*a); *a = NULL; }
static void copy_all(orc_state *o, double *dst, const double *src) {
  memcpy(dst, src, sizeof(double) * o->ntot);
}

/* ------------------------------------------------------------------------ */
/* Halo exchange on one rank (mpiboundary.f90).  Slabs span the full extent  */
/* of the other two axes; exchanges run x, then y, then z.                   */
/* kind 0: cell-like (n+1:n+2)<-(1:2), (-1:0)<-(n-1:n)  (cell_xface)        */
/* kind 1: normal face (n+1:n+2)<-(1:2), (-2:0)<-(n-2:n) (bx_xface/_xface1) */
/* kind 2: node (n:n+2)<-(0:2), (-2:-1)<-(n-2:n-1)  (node_xface1/node_xface) */
/* ------------------------------------------------------------------------ */
static void copy_plane(orc_state *o, double *a, int axis, int dst, int src) {
  int n[3] = {o->nx, o->ny, o->nz};
  int u = (axis + 1) % 3, w = (axis + 2) % 3;
  for (int jw = -2; jw <= n[w] + 2; ++jw)
    for (int ju = -2; ju <= n[u] + 2; ++ju) {
      int d[3], q[3];
      d[axis] = dst; q[axis] = src;
      d[u] = q[u] = ju; d[w] = q[w] = jw;
      A(a, d[0], d[1], d[2]) = A(a, q[0], q[1], q[2]);
    }
}
static int nax(orc_state *o, int axis) { return axis == 0 ? o->nx : (axis == 1 ? o->ny : o->nz); }
static void exchange(orc_state *o, double *a, int axis, int kind) {
  if (!o->per[axis]) return;  /* MPI_PROC_NULL neighbours: no-op */
  int n = nax(o, axis);
  if (kind == 0) {
    for (int l = 0; l < 2; ++l) copy_plane(o, a, axis, n + 1 + l, 1 + l);
    for (int l = 0; l < 2; ++l) copy_plane(o, a, axis, -1 + l, n - 1 + l);
  } else if (kind == 1) {
    for (int l = 0; l < 2; ++l) copy_plane(o, a, axis, n + 1 + l, 1 + l);
    for (int l = 0; l < 3; ++l) copy_plane(o, a, axis, -2 + l, n - 2 + l);
  } else {
    for (int l = 0; l < 3; ++l) copy_plane(o, a, axis, n + l, l);
    for (int l = 0; l < 2; ++l) copy_plane(o, a, axis, -2 + l, n - 2 + l);
  }
}
static int user(orc_state *o, int side) { return o->bc[side] == LARE_BC_USER; }

/* cell-centred mirror of energy_bcs / density_bcs / temperature_bcs
 * (boundary.f90:129-246) for one field. */
static void mirror_cell(orc_state *o, double *a) {
  for (int ax = 0; ax < 3; ++ax) {
    int n = nax(o, ax);
    if (user(o, 2 * ax)) { copy_plane(o, a, ax, 0, 1); copy_plane(o, a, ax, -1, 2); }
    if (user(o, 2 * ax + 1)) { copy_plane(o, a, ax, n + 1, n); copy_plane(o, a, ax, n + 2, n - 1); }
  }
}
/* energy_bcs (boundary.f90:129) = energy_mpi + mirror; density_bcs alike */
static void cell_bcs(orc_state *o, double *a) {
  for (int ax = 0; ax < 3; ++ax) exchange(o, a, ax, 0);
  mirror_cell(o, a);
}
/* temperature_bcs (boundary.f90:212-246): exchanges rho (density_mpi,
 * :214), then mirrors temperature. */
static void temperature_bcs(orc_state *o) {
  for (int ax = 0; ax < 3; ++ax) exchange(o, o->rho, ax, 0);
  mirror_cell(o, o->temperature);
}
/* bfield_bcs (boundary.f90:63-121) after bfield_mpi (mpiboundary.f90:31-114) */
static void bfield_bcs(orc_state *o) {
  double *b[3] = {o->bx, o->by, o->bz};
  for (int ax = 0; ax < 3; ++ax)
    for (int c = 0; c < 3; ++c) exchange(o, b[c], ax, c == ax ? 1 : 0);
  for (int ax = 0; ax < 3; ++ax) {
    int n = nax(o, ax);
    if (user(o, 2 * ax))
      for (int c = 0; c < 3; ++c) {
        if (c == ax) { copy_plane(o, b[c], ax, -1, 1); copy_plane(o, b[c], ax, -2, 2); }
        else { copy_plane(o, b[c], ax, 0, 1); copy_plane(o, b[c], ax, -1, 2); }
      }
    if (user(o, 2 * ax + 1))
      for (int c = 0; c < 3; ++c) {
        if (c == ax) { copy_plane(o, b[c], ax, n + 1, n - 1); copy_plane(o, b[c], ax, n + 2, n - 2); }
        else { copy_plane(o, b[c], ax, n + 1, n); copy_plane(o, b[c], ax, n + 2, n - 1); }
      }
  }
}
/* zero-velocity walls of velocity_bcs / remap_v_bcs (boundary.f90:255-343) */
static void zero_planes(orc_state *o, double *a, int ax, int from, int to) {
  int n[3] = {o->nx, o->ny, o->nz};
  int u = (ax + 1) % 3, w = (ax + 2) % 3;
  for (int l = from; l <= to; ++l)
    for (int jw = -2; jw <= n[w] + 2; ++jw)
      for (int ju = -2; ju <= n[u] + 2; ++ju) {
        int d[3];
        d[ax] = l; d[u] = ju; d[w] = jw;
        A(a, d[0], d[1], d[2]) = 0.0;
      }
}
static void node_bcs(orc_state *o, double *u, double *v, double *w) {
  double *f[3] = {u, v, w};
  for (int ax = 0; ax < 3; ++ax)
    for (int c = 0; c < 3; ++c) exchange(o, f[c], ax, 2);
  for (int ax = 0; ax < 3; ++ax) {
    int n = nax(o, ax);
    if (user(o, 2 * ax)) for (int c = 0; c < 3; ++c) zero_planes(o, f[c], ax, -2, 0);
    if (user(o, 2 * ax + 1)) for (int c = 0; c < 3; ++c) zero_planes(o, f[c], ax, n, n + 2);
  }
}
static void velocity_bcs(orc_state *o) { node_bcs(o, o->vx, o->vy, o->vz); }
static void remap_v_bcs(orc_state *o) { node_bcs(o, o->vx1, o->vy1, o->vz1); }
/* boundary.f90:48-55 */
void orc_boundary_conditions(orc_state *o) {
  bfield_bcs(o);
  cell_bcs(o, o->energy);
  cell_bcs(o, o->rho);
  velocity_bcs(o);
}

/* dm_x_bcs / dm_y_bcs / dm_z_bcs (xremap.f90:899-917 and y/z analogues) */
static void dm_bcs(orc_state *o, int ax) {
  int n = nax(o, ax);
  double *dm = o->dm;
  if (o->per[ax]) {
    for (int l = 0; l < 2; ++l) copy_plane(o, dm, ax, n + 1 + l, 1 + l);
    for (int l = 0; l < 2; ++l) copy_plane(o, dm, ax, -1 + l, n - 1 + l);
    return;
  }
  /* MPI_PROC_NULL on both sides: clamp over the (0:n+1) transverse box */
  for (int k = 0; k <= o->nz + 1; ++k)
    for (int j = 0; j <= o->ny + 1; ++j)
      for (int i = 0; i <= o->nx + 1; ++i) {
        int d[3] = {i, j, k};
        if (d[ax] != 0) continue;
        int s[3] = {i, j, k};
        d[ax] = n + 1; s[ax] = n;
        A(dm, d[0], d[1], d[2]) = A(dm, s[0], s[1], s[2]);
      }
  for (int k = 0; k <= o->nz + 1; ++k)
    for (int j = 0; j <= o->ny + 1; ++j)
      for (int i = 0; i <= o->nx + 1; ++i) {
        int d[3] = {i, j, k};
        if (d[ax] != 0) continue;
        int s[3] = {i, j, k};
        d[ax] = -1; s[ax] = 0;
        A(dm, d[0], d[1], d[2]) = A(dm, s[0], s[1], s[2]);
      }
}

/* ------------------------------------------------------------------------ */
/* Van Leer limited flux core shared by every *_flux routine                */
/* (e.g. xremap.f90:369-398, :500-528, :565-596, :628-659).                 */
/* ------------------------------------------------------------------------ */
typedef struct { double sign_v, vad_p, vad_m; } upw_t;
/* branch-free upwind switch (xremap.f90:377-379); SIGN(1,-0.0) = -1 */
static inline upw_t upwind(double v_advect) {
  upw_t r;
  r.sign_v = copysign(1.0, v_advect);
  r.vad_p = (r.sign_v + 1.0) * 0.5;
  r.vad_m = 1.0 - r.vad_p;
  return r;
}
/* Full limiter given phi (xremap.f90:390-396). */
static inline double vl_Di(double fm, double fi, double fp, double fp2, double vad_p, double vad_m,
                           double sign_v, double dci, double dcu, double dbu, double phi) {
  double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
  double dfu = dfm * vad_p + dfp * vad_m;
  double Da = (2.0 - phi) * fabs(dfi) / dci + (1.0 + phi) * fabs(dfu) / dcu;
  Da = Da * SIXTH;
  double ss = 0.5 * (copysign(1.0, dfi) + copysign(1.0, dfu));
  return sign_v * ss * min3(fabs(Da) * dbu, fabs(dfi), fabs(dfu));
}

/* ------------------------------------------------------------------------ */
/* Remap sweeps.  The three directions are written out separately, as in   */
/* the reference, because loop ranges and update orders differ per         */
/* direction (xremap.f90:121-157, yremap.f90:122-158, zremap.f90:121-157).  */
/* ------------------------------------------------------------------------ */

/* geometry loops of remap_x/y/z (xremap.f90:57-115, yremap.f90:58-116,
 * zremap.f90:57-115): cv1, cv2, d?b1, d?c1 over the whole cell box. */
static void remap_geometry(orc_state *o, int ax) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double dt = o->dt;
  for (int iz = -1; iz <= nz + 2; ++iz) {
    int izm = iz - 1;
    for (int iy = -1; iy <= ny + 2; ++iy) {
      int iym = iy - 1;
      for (int ix = -1; ix <= nx + 2; ++ix) {
        int ixm = ix - 1;
        double vxb = (A(o->vx1, ix, iy, iz) + A(o->vx1, ix, iym, iz) + A(o->vx1, ix, iy, izm) + A(o->vx1, ix, iym, izm)) * 0.25;
        double vxbm = (A(o->vx1, ixm, iy, iz) + A(o->vx1, ixm, iym, iz) + A(o->vx1, ixm, iy, izm) + A(o->vx1, ixm, iym, izm)) * 0.25;
        double vyb = (A(o->vy1, ix, iy, iz) + A(o->vy1, ixm, iy, iz) + A(o->vy1, ix, iy, izm) + A(o->vy1, ixm, iy, izm)) * 0.25;
        double vybm = (A(o->vy1, ix, iym, iz) + A(o->vy1, ixm, iym, iz) + A(o->vy1, ix, iym, izm) + A(o->vy1, ixm, iym, izm)) * 0.25;
        double vzb = (A(o->vz1, ix, iy, iz) + A(o->vz1, ixm, iy, iz) + A(o->vz1, ix, iym, iz) + A(o->vz1, ixm, iym, iz)) * 0.25;
        double vzbm = (A(o->vz1, ix, iy, izm) + A(o->vz1, ixm, iy, izm) + A(o->vz1, ix, iym, izm) + A(o->vz1, ixm, iym, izm)) * 0.25;
        double dv, cvv = A(o->cv, ix, iy, iz);
        if (ax == 0) {
          dv = ((double)o->ypass * (vyb - vybm) / DYB(iy) + (double)o->zpass * (vzb - vzbm) / DZB(iz) + (vxb - vxbm) / DXB(ix)) * dt;
          A(o->cv1, ix, iy, iz) = cvv * (1.0 + dv);
          dv = ((double)o->ypass * (vyb - vybm) / DYB(iy) + (double)o->zpass * (vzb - vzbm) / DZB(iz)) * dt;
          A(o->cv2, ix, iy, iz) = cvv * (1.0 + dv);
          A(o->dxb1, ix, iy, iz) = DXB(ix) + (vxb - vxbm) * dt;
        } else if (ax == 1) {
          dv = ((double)o->xpass * (vxb - vxbm) / DXB(ix) + (double)o->zpass * (vzb - vzbm) / DZB(iz) + (vyb - vybm) / DYB(iy)) * dt;
          A(o->cv1, ix, iy, iz) = cvv * (1.0 + dv);
          dv = ((double)o->xpass * (vxb - vxbm) / DXB(ix) + (double)o->zpass * (vzb - vzbm) / DZB(iz)) * dt;
          A(o->cv2, ix, iy, iz) = cvv * (1.0 + dv);
          A(o->dxb1, ix, iy, iz) = DYB(iy) + (vyb - vybm) * dt;
        } else {
          dv = ((double)o->xpass * (vxb - vxbm) / DXB(ix) + (double)o->ypass * (vyb - vybm) / DYB(iy) + (vzb - vzbm) / DZB(iz)) * dt;
          A(o->cv1, ix, iy, iz) = cvv * (1.0 + dv);
          dv = ((double)o->xpass * (vxb - vxbm) / DXB(ix) + (double)o->ypass * (vyb - vybm) / DYB(iy)) * dt;
          A(o->cv2, ix, iy, iz) = cvv * (1.0 + dv);
          A(o->dxb1, ix, iy, iz) = DZB(iz) + (vzb - vzbm) * dt;
        }
      }
    }
  }
  /* d?c1 = 0.5*(d?b1 + d?b1(+1)) with the +1 index clamped at n+2 */
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) {
        int p[3] = {ix, iy, iz};
        int n = nax(o, ax);
        p[ax] = p[ax] + 1 < n + 2 ? p[ax] + 1 : n + 2;
        A(o->dxc1, ix, iy, iz) = 0.5 * (A(o->dxb1, ix, iy, iz) + A(o->dxb1, p[0], p[1], p[2]));
      }
}

/* Generic CT flux of a transverse field f along sweep axis ax
 * (vx_by_flux xremap.f90:340-403, vx_bz_flux :407-470, vy_bx_flux
 * yremap.f90:341-404, vy_bz_flux :408-471, vz_bx_flux zremap.f90:340-403,
 * vz_by_flux :407-470).  vax: axis along which the sweep velocity is averaged
 * (v(i) + v(i - e_vax)); gax: axis along which d?b1 is averaged
 * (d(i) + d(i + e_gax)).  Loop box [-1, n+1]^3; the sweep index is clamped
 * to [-1, n+2] for the m and p2 points. */
static void ct_flux(orc_state *o, int ax, const double *f, int vax, int gax) {
  const double *vel = ax == 0 ? o->vx1 : (ax == 1 ? o->vy1 : o->vz1);
  int n = nax(o, ax);
  double dt = o->dt;
  for (int iz = -1; iz <= o->nz + 1; ++iz)
    for (int iy = -1; iy <= o->ny + 1; ++iy)
      for (int ix = -1; ix <= o->nx + 1; ++ix) {
        int c[3] = {ix, iy, iz};
        int s = c[ax];
        int sm = s - 1 > -1 ? s - 1 : -1, sp = s + 1, sp2 = s + 2 < n + 2 ? s + 2 : n + 2;
        int v2[3] = {ix, iy, iz};
        v2[vax] -= 1;
        double v_advect = (A(vel, ix, iy, iz) + A(vel, v2[0], v2[1], v2[2])) * 0.5;
#define PT(arr, sv, dg) ({ int q_[3] = {ix, iy, iz}; q_[ax] = (sv); q_[gax] += (dg); A(arr, q_[0], q_[1], q_[2]); })
        double db = (PT(o->dxb1, s, 0) + PT(o->dxb1, s, 1)) * 0.5;
        double dbp = (PT(o->dxb1, sp, 0) + PT(o->dxb1, sp, 1)) * 0.5;
        double dbp2 = (PT(o->dxb1, sp2, 0) + PT(o->dxb1, sp2, 1)) * 0.5;
        double dbm = (PT(o->dxb1, sm, 0) + PT(o->dxb1, sm, 1)) * 0.5;
        double fm = PT(f, sm, 0) / dbm;
        double fi = PT(f, s, 0) / db;
        double fp = PT(f, sp, 0) / dbp;
        double fp2 = PT(f, sp2, 0) / dbp2;
        upw_t r = upwind(v_advect);
        double fu = fi * r.vad_p + fp * r.vad_m;
        double dbu = PT(o->dxb1, s, 0) * r.vad_p + PT(o->dxb1, sp, 0) * r.vad_m;
        double dci = PT(o->dxc1, s, 0);
        double dcu = PT(o->dxc1, sm, 0) * r.vad_p + PT(o->dxc1, sp, 0) * r.vad_m;
        double du = db * r.vad_p + dbp * r.vad_m;
        double phi = fabs(v_advect) * dt / du;
        double Di = vl_Di(fm, fi, fp, fp2, r.vad_p, r.vad_m, r.sign_v, dci, dcu, dbu, phi);
#undef PT
        A(o->flux, ix, iy, iz) = (fu + Di * (1.0 - phi)) * v_advect * dt;
      }
}

/* CT update of field g over a box: g = g - F + F(-e_ax) ("minus") or
 * g = g + F - F(-e_ax) ("plus"). */
static void ct_update(orc_state *o, double *g, int ax, int plus, int lx, int hx, int ly, int hy, int lz, int hz) {
  for (int iz = lz; iz <= hz; ++iz)
    for (int iy = ly; iy <= hy; ++iy)
      for (int ix = lx; ix <= hx; ++ix) {
        int q[3] = {ix, iy, iz};
        q[ax] -= 1;
        double F = A(o->flux, ix, iy, iz), Fm = A(o->flux, q[0], q[1], q[2]);
        if (plus) A(g, ix, iy, iz) = A(g, ix, iy, iz) + F - Fm;
        else A(g, ix, iy, iz) = A(g, ix, iy, iz) - F + Fm;
      }
}

/* mass flux (x_mass_flux xremap.f90:474-533 and y/z analogues): sweep axis
 * 0..n, transverse 0..n+1. */
static void mass_flux(orc_state *o, int ax) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  const double *vel = ax == 0 ? o->vx1 : (ax == 1 ? o->vy1 : o->vz1);
  double dt = o->dt;
  int hi[3] = {nx + 1, ny + 1, nz + 1};
  hi[ax] = nax(o, ax);
  for (int iz = 0; iz <= hi[2]; ++iz)
    for (int iy = 0; iy <= hi[1]; ++iy)
      for (int ix = 0; ix <= hi[0]; ++ix) {
        int ixm = ix - 1, iym = iy - 1, izm = iz - 1;
        double v_advect, area;
        if (ax == 0) {
          area = DYB(iy) * DZB(iz);
          v_advect = (A(vel, ix, iy, iz) + A(vel, ix, iym, iz) + A(vel, ix, iy, izm) + A(vel, ix, iym, izm)) * 0.25;
        } else if (ax == 1) {
          area = DXB(ix) * DZB(iz);
          v_advect = (A(vel, ix, iy, iz) + A(vel, ixm, iy, iz) + A(vel, ix, iy, izm) + A(vel, ixm, iy, izm)) * 0.25;
        } else {
          area = DXB(ix) * DYB(iy);
          v_advect = (A(vel, ix, iy, iz) + A(vel, ixm, iy, iz) + A(vel, ix, iym, iz) + A(vel, ixm, iym, iz)) * 0.25;
        }
        int c[3] = {ix, iy, iz}, m[3] = {ix, iy, iz}, p[3] = {ix, iy, iz}, p2[3] = {ix, iy, iz};
        m[ax] -= 1; p[ax] += 1; p2[ax] += 2;
#define AT(arr, q) A(arr, q[0], q[1], q[2])
        double fm = AT(o->rho, m), fi = AT(o->rho, c), fp = AT(o->rho, p), fp2 = AT(o->rho, p2);
        upw_t r = upwind(v_advect);
        double fu = fi * r.vad_p + fp * r.vad_m;
        double dbu = AT(o->dxb1, c) * r.vad_p + AT(o->dxb1, p) * r.vad_m;
        double dci = AT(o->dxc1, c);
        double dcu = AT(o->dxc1, m) * r.vad_p + AT(o->dxc1, p) * r.vad_m;
        double phi = fabs(v_advect) * dt / dbu;
        double Di = vl_Di(fm, fi, fp, fp2, r.vad_p, r.vad_m, r.sign_v, dci, dcu, dbu, phi);
        AT(o->dm, c) = (fu + Di * (1.0 - phi)) * v_advect * dt * area;
      }
}

/* energy flux in mass coordinates (x_energy_flux xremap.f90:539-601 and
 * analogues): box 0..n in every axis. */
static void energy_flux(orc_state *o, int ax) {
  const double *vel = ax == 0 ? o->vx1 : (ax == 1 ? o->vy1 : o->vz1);
  for (int iz = 0; iz <= o->nz; ++iz)
    for (int iy = 0; iy <= o->ny; ++iy)
      for (int ix = 0; ix <= o->nx; ++ix) {
        int ixm = ix - 1, iym = iy - 1, izm = iz - 1;
        double v_advect, area;
        if (ax == 0) {
          area = DYB(iy) * DZB(iz);
          v_advect = (A(vel, ix, iy, iz) + A(vel, ix, iym, iz) + A(vel, ix, iy, izm) + A(vel, ix, iym, izm)) * 0.25;
        } else if (ax == 1) {
          area = DXB(ix) * DZB(iz);
          v_advect = (A(vel, ix, iy, iz) + A(vel, ixm, iy, iz) + A(vel, ix, iy, izm) + A(vel, ixm, iy, izm)) * 0.25;
        } else {
          area = DXB(ix) * DYB(iy);
          v_advect = (A(vel, ix, iy, iz) + A(vel, ixm, iy, iz) + A(vel, ix, iym, iz) + A(vel, ixm, iym, iz)) * 0.25;
        }
        int c[3] = {ix, iy, iz}, m[3] = {ix, iy, iz}, p[3] = {ix, iy, iz}, p2[3] = {ix, iy, iz};
        m[ax] -= 1; p[ax] += 1; p2[ax] += 2;
        double fm = AT(o->energy, m), fi = AT(o->energy, c), fp = AT(o->energy, p), fp2 = AT(o->energy, p2);
        upw_t r = upwind(v_advect);
        double fu = fi * r.vad_p + fp * r.vad_m;
        double dbu = AT(o->dxb1, c) * r.vad_p + AT(o->dxb1, p) * r.vad_m;
        double dci = AT(o->dxc1, c);
        double dcu = AT(o->dxc1, m) * r.vad_p + AT(o->dxc1, p) * r.vad_m;
        double phi = fabs(v_advect) * o->dt / dbu;
        double Di = vl_Di(fm, fi, fp, fp2, r.vad_p, r.vad_m, r.sign_v, dci, dcu, dbu, phi);
        double rhou = AT(o->rho1, c) * r.vad_p + AT(o->rho1, p) * r.vad_m;
        double dmu = fabs(AT(o->dm, c)) / area / dbu / rhou;
        AT(o->flux, c) = (fu + Di * (1.0 - dmu)) * AT(o->dm, c);
      }
}

/* momentum flux of one velocity component (x_momx_flux xremap.f90:605-699
 * and the eight analogues): sweep axis -1..n, transverse 0..n; with rke the
 * kinetic energy the remap loses at each edge (xremap.f90:664-695,
 * yremap.f90:665-696, zremap.f90:664-695); then flux *= dm over the same
 * box. */
static void mom_flux(orc_state *o, int ax, const double *v) {
  const double *vel = ax == 0 ? o->vx1 : (ax == 1 ? o->vy1 : o->vz1);
  int lo[3] = {0, 0, 0};
  lo[ax] = -1;
  for (int iz = lo[2]; iz <= o->nz; ++iz)
    for (int iy = lo[1]; iy <= o->ny; ++iy)
      for (int ix = lo[0]; ix <= o->nx; ++ix) {
        double area = ax == 0 ? DYC(iy) * DZC(iz) : (ax == 1 ? DXC(ix) * DZC(iz) : DXC(ix) * DYC(iy));
        double v_advect = A(vel, ix, iy, iz);
        int c[3] = {ix, iy, iz}, m[3] = {ix, iy, iz}, p[3] = {ix, iy, iz}, p2[3] = {ix, iy, iz};
        m[ax] -= 1; p[ax] += 1; p2[ax] += 2;
        double fm = AT(v, m), fi = AT(v, c), fp = AT(v, p), fp2 = AT(v, p2);
        upw_t r = upwind(v_advect);
        double fu = fi * r.vad_p + fp * r.vad_m;
        double dbu = AT(o->dxc1, c) * r.vad_p + AT(o->dxc1, p) * r.vad_m;
        double dci = AT(o->dxb1, p);
        double dcu = AT(o->dxb1, c) * r.vad_p + AT(o->dxb1, p2) * r.vad_m;
        double phi = fabs(v_advect) * o->dt / dbu;
        double Di = vl_Di(fm, fi, fp, fp2, r.vad_p, r.vad_m, r.sign_v, dci, dcu, dbu, phi);
        double rhou = AT(o->rho_v, c) * r.vad_p + AT(o->rho_v, p) * r.vad_m;
        double dmu = fabs(AT(o->dm, c)) / area / dbu / rhou;
        AT(o->flux, c) = fu + Di * (1.0 - dmu);
      }
  if (o->p.rke) {
    /* edges between nodes s and s+1 for s = 0..n-1 along the sweep,
     * transverse 0..n; each edge adds dk to the four cells around it (sweep
     * coordinate s+1), the lower transverse axis varying fastest */
    int hi[3] = {o->nx, o->ny, o->nz};
    hi[ax] -= 1;
    const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
    for (int iz = 0; iz <= hi[2]; ++iz)
      for (int iy = 0; iy <= hi[1]; ++iy)
        for (int ix = 0; ix <= hi[0]; ++ix) {
          int c[3] = {ix, iy, iz}, m[3] = {ix, iy, iz}, p[3] = {ix, iy, iz};
          m[ax] -= 1;
          p[ax] += 1;
          double mc = AT(o->rho_v1, c) * AT(o->cv2, c);
          double mp = AT(o->rho_v1, p) * AT(o->cv2, p);
          double ai = (AT(v, c) - AT(o->flux, m)) * AT(o->dm, m) / mc - (AT(v, c) - AT(o->flux, c)) * AT(o->dm, c) / mc;
          double aip = (AT(v, p) - AT(o->flux, c)) * AT(o->dm, c) / mp - (AT(v, p) - AT(o->flux, p)) * AT(o->dm, p) / mp;
          double dk = (AT(v, p) - AT(v, c)) * (AT(o->flux, c) - 0.5 * (AT(v, p) + AT(v, c))) -
                      0.5 * ai * (AT(v, c) - AT(o->flux, c)) + 0.5 * aip * (AT(v, p) - AT(o->flux, c));
          dk = dk * AT(o->dm, c) * 0.5;
          for (int q = 0; q < 4; ++q) {
            int d[3] = {ix, iy, iz};
            d[ax] += 1;
            d[t1] += q & 1;
            d[t2] += q >> 1;
            AT(o->delta_ke, d) = AT(o->delta_ke, d) + dk;
          }
        }
  }
  for (int iz = lo[2]; iz <= o->nz; ++iz)
    for (int iy = lo[1]; iy <= o->ny; ++iy)
      for (int ix = lo[0]; ix <= o->nx; ++ix) A(o->flux, ix, iy, iz) = A(o->flux, ix, iy, iz) * A(o->dm, ix, iy, iz);
}

/* 8-point vertex sum in the reference's order (xremap.f90:233-236). */
static inline double sum8(orc_state *o, const double *a, int ix, int iy, int iz) {
  int ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
  return A(a, ix, iy, iz) + A(a, ixp, iy, iz) + A(a, ix, iyp, iz) + A(a, ixp, iyp, iz) +
         A(a, ix, iy, izp) + A(a, ixp, iy, izp) + A(a, ix, iyp, izp) + A(a, ixp, iyp, izp);
}

/* remap_x / remap_y / remap_z (xremap.f90:37-336, yremap.f90:37-337,
 * zremap.f90:37-336) */
static void remap_dir(orc_state *o, int ax) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  o->cv2 = cellbox(o);
  o->dxb1 = cellbox(o);
  o->dxc1 = cellbox(o);
  {
    int l[3] = {-1, -1, -1};
    l[ax] = -2;
    o->flux = box(o, l[0], nx + 2, l[1], ny + 2, l[2], nz + 2);
  }
  remap_geometry(o, ax);

  /* constrained transport of the face fluxes */
  if (ax == 0) {
    ct_flux(o, 0, o->by, 2, 1);                                     /* vx_by_flux */
    ct_update(o, o->by, 0, 0, 0, nx + 1, 0, ny + 1, 0, nz + 1);     /* :121-128 */
    ct_update(o, o->bx, 1, 1, 0, nx + 1, 0, ny + 1, 0, nz + 1);     /* :130-137 */
    ct_flux(o, 0, o->bz, 1, 2);                                     /* vx_bz_flux */
    ct_update(o, o->bz, 0, 0, 1, nx, 1, ny, 0, nz);                 /* :141-148 */
    ct_update(o, o->bx, 2, 1, 0, nx, 0, ny + 1, 0, nz + 1);         /* :150-157 */
  } else if (ax == 1) {
    ct_flux(o, 1, o->bx, 2, 0);                                     /* vy_bx_flux */
    ct_update(o, o->bx, 1, 0, 0, nx + 1, 0, ny + 1, 0, nz + 1);     /* :122-129 */
    ct_update(o, o->by, 0, 1, 0, nx + 1, 0, ny + 1, 0, nz + 1);     /* :131-138 */
    ct_flux(o, 1, o->bz, 0, 2);                                     /* vy_bz_flux */
    ct_update(o, o->bz, 1, 0, 0, nx + 1, 0, ny + 1, 0, nz + 1);     /* :142-149 */
    ct_update(o, o->by, 2, 1, 1, nx, 0, ny, 1, nz);                 /* :151-158 */
  } else {
    ct_flux(o, 2, o->bx, 1, 0);                                     /* vz_bx_flux */
    ct_update(o, o->bx, 2, 0, 0, nx + 1, 0, ny + 1, 0, nz + 1);     /* :121-128 */
    ct_update(o, o->bz, 0, 1, 0, nx + 1, 0, ny + 1, 0, nz + 1);     /* :130-137 */
    ct_flux(o, 2, o->by, 0, 1);                                     /* vz_by_flux */
    ct_update(o, o->by, 2, 0, 0, nx + 1, 0, ny + 1, 0, nz + 1);     /* :141-148 */
    ct_update(o, o->bz, 1, 1, 0, nx + 1, 0, ny + 1, 0, nz + 1);     /* :150-157 */
  }

  if (o->predictor_step) {  /* xremap.f90:159-162: no flag reset, no BCs */
    drop(&o->cv2); drop(&o->flux); drop(&o->dxb1); drop(&o->dxc1);
    return;
  }

  o->rho1 = cellbox(o);
  o->dm = cellbox(o);
  o->rho_v = cellbox(o);
  o->rho_v1 = cellbox(o);
  copy_all(o, o->rho1, o->rho);
  mass_flux(o, ax);
  dm_bcs(o, ax);

  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        int m[3] = {ix, iy, iz};
        m[ax] -= 1;
        A(o->rho, ix, iy, iz) = (A(o->rho1, ix, iy, iz) * A(o->cv1, ix, iy, iz) + AT(o->dm, m) - A(o->dm, ix, iy, iz)) / A(o->cv2, ix, iy, iz);
      }

  energy_flux(o, ax);
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        int m[3] = {ix, iy, iz};
        m[ax] -= 1;
        A(o->energy, ix, iy, iz) = (A(o->energy, ix, iy, iz) * A(o->cv1, ix, iy, iz) * A(o->rho1, ix, iy, iz) + AT(o->flux, m) - A(o->flux, ix, iy, iz)) /
                                   (A(o->cv2, ix, iy, iz) * A(o->rho, ix, iy, iz));
      }

  /* vertex quantities (xremap.f90:201-290): sweep axis -1..n+1 for rho_v */
  {
    int lo[3] = {0, 0, 0}, hi[3] = {nx, ny, nz};
    lo[ax] = -1;
    hi[ax] = nax(o, ax) + 1;
    for (int iz = lo[2]; iz <= hi[2]; ++iz)
      for (int iy = lo[1]; iy <= hi[1]; ++iy)
        for (int ix = lo[0]; ix <= hi[0]; ++ix) {
          int ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
          double s = A(o->rho1, ix, iy, iz) * A(o->cv1, ix, iy, iz) + A(o->rho1, ixp, iy, iz) * A(o->cv1, ixp, iy, iz) +
                     A(o->rho1, ix, iyp, iz) * A(o->cv1, ix, iyp, iz) + A(o->rho1, ixp, iyp, iz) * A(o->cv1, ixp, iyp, iz) +
                     A(o->rho1, ix, iy, izp) * A(o->cv1, ix, iy, izp) + A(o->rho1, ixp, iy, izp) * A(o->cv1, ixp, iy, izp) +
                     A(o->rho1, ix, iyp, izp) * A(o->cv1, ix, iyp, izp) + A(o->rho1, ixp, iyp, izp) * A(o->cv1, ixp, iyp, izp);
          A(o->rho_v, ix, iy, iz) = s / sum8(o, o->cv1, ix, iy, iz);
        }
  }
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) A(o->flux, ix, iy, iz) = sum8(o, o->cv1, ix, iy, iz);
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) A(o->cv1, ix, iy, iz) = A(o->flux, ix, iy, iz) * 0.125;
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) A(o->flux, ix, iy, iz) = sum8(o, o->cv2, ix, iy, iz);
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) A(o->cv2, ix, iy, iz) = A(o->flux, ix, iy, iz) * 0.125;
  {
    /* in-place average of the sweep velocity along the sweep axis (Q8) */
    double *vel = ax == 0 ? o->vx1 : (ax == 1 ? o->vy1 : o->vz1);
    int lo[3] = {0, 0, 0}, hi[3] = {nx, ny, nz};
    lo[ax] = -2;
    hi[ax] = nax(o, ax) + 1;
    for (int iz = lo[2]; iz <= hi[2]; ++iz)
      for (int iy = lo[1]; iy <= hi[1]; ++iy)
        for (int ix = lo[0]; ix <= hi[0]; ++ix) {
          int p[3] = {ix, iy, iz};
          p[ax] += 1;
          A(o->flux, ix, iy, iz) = (A(vel, ix, iy, iz) + AT(vel, p)) * 0.5;
        }
    for (int iz = lo[2]; iz <= hi[2]; ++iz)
      for (int iy = lo[1]; iy <= hi[1]; ++iy)
        for (int ix = lo[0]; ix <= hi[0]; ++ix) A(vel, ix, iy, iz) = A(o->flux, ix, iy, iz);
    /* dm vertex average, sweep axis -1..n */
    int l2[3] = {0, 0, 0};
    l2[ax] = -1;
    for (int iz = l2[2]; iz <= nz; ++iz)
      for (int iy = l2[1]; iy <= ny; ++iy)
        for (int ix = l2[0]; ix <= nx; ++ix) A(o->flux, ix, iy, iz) = sum8(o, o->dm, ix, iy, iz);
    for (int iz = l2[2]; iz <= nz; ++iz)
      for (int iy = l2[1]; iy <= ny; ++iy)
        for (int ix = l2[0]; ix <= nx; ++ix) A(o->dm, ix, iy, iz) = A(o->flux, ix, iy, iz) * 0.125;
  }
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) {
        int m[3] = {ix, iy, iz};
        m[ax] -= 1;
        A(o->rho_v1, ix, iy, iz) = (A(o->rho_v, ix, iy, iz) * A(o->cv1, ix, iy, iz) + AT(o->dm, m) - A(o->dm, ix, iy, iz)) / A(o->cv2, ix, iy, iz);
      }

  double *vv[3] = {o->vx, o->vy, o->vz};
  for (int c = 0; c < 3; ++c) {
    mom_flux(o, ax, vv[c]);
    for (int iz = 0; iz <= nz; ++iz)
      for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) {
          int m[3] = {ix, iy, iz};
          m[ax] -= 1;
          A(vv[c], ix, iy, iz) = (A(o->rho_v, ix, iy, iz) * A(vv[c], ix, iy, iz) * A(o->cv1, ix, iy, iz) + AT(o->flux, m) - A(o->flux, ix, iy, iz)) /
                                 (A(o->cv2, ix, iy, iz) * A(o->rho_v1, ix, iy, iz));
        }
  }
#undef AT

  orc_boundary_conditions(o);
  drop(&o->rho1); drop(&o->dm); drop(&o->cv2); drop(&o->flux);
  drop(&o->dxb1); drop(&o->dxc1); drop(&o->rho_v); drop(&o->rho_v1);
  if (ax == 0) o->xpass = 0;
  else if (ax == 1) o->ypass = 0;
  else o->zpass = 0;
}

/* remap.f90:33-131 eulerian_remap */
void orc_eulerian_remap(orc_state *o, int step) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  if (o->p.rke)
    for (int iz = -1; iz <= nz + 2; ++iz)
      for (int iy = -1; iy <= ny + 2; ++iy)
        for (int ix = -1; ix <= nx + 2; ++ix) A(o->delta_ke, ix, iy, iz) = 0.0;
  o->xpass = o->ypass = o->zpass = 1;
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) {
        A(o->bx, ix, iy, iz) = A(o->bx, ix, iy, iz) * DYB(iy) * DZB(iz);
        A(o->by, ix, iy, iz) = A(o->by, ix, iy, iz) * DXB(ix) * DZB(iz);
        A(o->bz, ix, iy, iz) = A(o->bz, ix, iy, iz) * DXB(ix) * DYB(iy);
      }
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy) A(o->bx, -2, iy, iz) = A(o->bx, -2, iy, iz) * DYB(iy) * DZB(iz);
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int ix = -1; ix <= nx + 2; ++ix) A(o->by, ix, -2, iz) = A(o->by, ix, -2, iz) * DXB(ix) * DZB(iz);
  for (int iy = -1; iy <= ny + 2; ++iy)
    for (int ix = -1; ix <= nx + 2; ++ix) A(o->bz, ix, iy, -2) = A(o->bz, ix, iy, -2) * DXB(ix) * DYB(iy);

  static const int order[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}};
  int c = ((step % 6) + 6) % 6;  /* MODULO(i, 6) */
  for (int s = 0; s < 3; ++s) remap_dir(o, order[c][s]);

  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) {
        A(o->bx, ix, iy, iz) = A(o->bx, ix, iy, iz) / DYB(iy) / DZB(iz);
        A(o->by, ix, iy, iz) = A(o->by, ix, iy, iz) / DXB(ix) / DZB(iz);
        A(o->bz, ix, iy, iz) = A(o->bz, ix, iy, iz) / DXB(ix) / DYB(iy);
      }
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy) A(o->bx, -2, iy, iz) = A(o->bx, -2, iy, iz) / DYB(iy) / DZB(iz);
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int ix = -1; ix <= nx + 2; ++ix) A(o->by, ix, -2, iz) = A(o->by, ix, -2, iz) / DXB(ix) / DZB(iz);
  for (int iy = -1; iy <= ny + 2; ++iy)
    for (int ix = -1; ix <= nx + 2; ++ix) A(o->bz, ix, iy, -2) = A(o->bz, ix, iy, -2) / DXB(ix) / DYB(iy);
  bfield_bcs(o);
}

/* ------------------------------------------------------------------------ */
/* Partially ionised plasma (neutral.f90)                                    */
/* ------------------------------------------------------------------------ */

/* exp(x): Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2 (ln2 split in a
 * 32-bit head, so k*LN2_HI is exact, and a tail), the Taylor polynomial of
 * degree 13 in r by Horner's rule with fused multiply-adds (truncation below
 * 0.05 ulp), and the exact scaling by 2^k.  Within an ulp of the libm exp
 * that the reference's EXP calls (neutral.f90:151-152); written out so that
 * the device computes the same operations and hence the same bits. */
#define LN2_HI 0x1.62e42feep-1
#define LN2_LO 0x1.a39ef35793c76p-33
#define INV_LN2 0x1.71547652b82fep+0
double orc_exp(double x) {
  if (x != x) return x;
  if (x > 710.0) return INFINITY;
  if (x < -746.0) return 0.0;
  const double k = rint(x * INV_LN2);
  double r = fma(-k, LN2_HI, x);
  r = fma(-k, LN2_LO, r);
  double p = 0x1.6124613a86d09p-33;       /* 1/13! */
  p = fma(p, r, 0x1.1eed8eff8d898p-29);   /* 1/12! */
  p = fma(p, r, 0x1.ae64567f544e4p-26);   /* 1/11! */
  p = fma(p, r, 0x1.27e4fb7789f5cp-22);   /* 1/10! */
  p = fma(p, r, 0x1.71de3a556c734p-19);   /* 1/9! */
  p = fma(p, r, 0x1.a01a01a01a01ap-16);   /* 1/8! */
  p = fma(p, r, 0x1.a01a01a01a01ap-13);   /* 1/7! */
  p = fma(p, r, 0x1.6c16c16c16c17p-10);   /* 1/6! */
  p = fma(p, r, 0x1.1111111111111p-7);    /* 1/5! */
  p = fma(p, r, 0x1.5555555555555p-5);    /* 1/4! */
  p = fma(p, r, 0x1.5555555555555p-3);    /* 1/3! */
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

/* neutral.f90:140-154 get_neutral: the Saha-type neutral fraction at
 * temperature t_v and density rho_v, with the dilute photospheric radiation
 * field (t_rad = tr, dilution 1/2) */
double orc_saha_neutral(const lare_params *p, double t_v, double rho_v) {
  const double t_rad = p->t_rad, dilution = 0.5;
  const double bof = 1.0 / (dilution * p->f_bar * t_rad * sqrt(t_v)) *
                     orc_exp((0.25 * (t_v / t_rad - 1.0) + 1.0) * p->t_bar / t_v);
  const double r = 0.5 * (-1.0 + sqrt(1.0 + p->r_bar * rho_v * bof));
  return r / (1.0 + r);
}

/* neutral.f90:158-195 neutral_fraction over the cell box -1..n+2: the
 * temperature by bisection of T = (gamma-1)(e - (1-xi(T)) ionise_pot) /
 * (2 - xi(T)) between (gamma-1) max((e - ionise_pot)/2, tiny) and
 * (gamma-1) e, at most 100 halvings, stopping when the step or the residual
 * falls below 1e-8; xi_n from the last probe.  An inverted bracket (the
 * reference's STOP) sets the fault and leaves that cell's bisection to run
 * its course. */
static void neutral_fraction(orc_state *o) {
  const double gm1 = o->p.gamma - 1.0, ip = o->ip;
  for (int iz = -1; iz <= o->nz + 2; ++iz)
    for (int iy = -1; iy <= o->ny + 2; ++iy)
      for (int ix = -1; ix <= o->nx + 2; ++ix) {
        const double rho0 = A(o->rho, ix, iy, iz), e0 = A(o->energy, ix, iy, iz);
        const double ta1 = gm1 * dmax((e0 - ip) / 2.0, NONE_ZERO), ta2 = gm1 * e0;
        if (ta1 > ta2) o->fault = LARE_F_TBOUNDS;
        double dx = ta2 - ta1, t = ta1, x = t;
        for (int loop = 1; loop <= 100; ++loop) {
          dx = dx / 2.0;
          x = t + dx;
          const double xi_a = orc_saha_neutral(&o->p, x, rho0);
          const double fa = x - gm1 * (e0 - (1.0 - xi_a) * ip) / (2.0 - xi_a);
          if (fa <= 0.0) t = x;
          if (fabs(dx) < 1.0e-8 || fabs(fa) < 1.0e-8) break;
        }
        A(o->xi_n, ix, iy, iz) = orc_saha_neutral(&o->p, x, rho0);
      }
}

/* neutral.f90:68-136 perpendicular_resistivity at nodes 0..n: the volume
 * averaged density, the temperature averaged over the eight cells, the
 * neutral fraction there, and eta_perp = eta_bar xi/(1-xi) |B|^2 / rho^2 /
 * sqrt(T), capped at 100.  The vertex field is the reference's: bxv and
 * byv take bz at the upper faces (:97-98) where bx and by would be the
 * symmetric choice. */
static void perpendicular_resistivity(orc_state *o) {
  const double gm1 = o->p.gamma - 1.0, ip = o->ip;
  for (int iz = 0; iz <= o->nz; ++iz)
    for (int iy = 0; iy <= o->ny; ++iy)
      for (int ix = 0; ix <= o->nx; ++ix) {
        const int ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
        const double *r = o->rho, *c = o->cv;
        double rho_v = A(r, ix, iy, iz) * A(c, ix, iy, iz) + A(r, ixp, iy, iz) * A(c, ixp, iy, iz) +
                       A(r, ix, iyp, iz) * A(c, ix, iyp, iz) + A(r, ixp, iyp, iz) * A(c, ixp, iyp, iz) +
                       A(r, ix, iy, izp) * A(c, ix, iy, izp) + A(r, ixp, iy, izp) * A(c, ixp, iy, izp) +
                       A(r, ix, iyp, izp) * A(c, ix, iyp, izp) + A(r, ixp, iyp, izp) * A(c, ixp, iyp, izp);
        rho_v = rho_v / sum8(o, c, ix, iy, iz);
        double bxv = A(o->bx, ix, iy, iz) + A(o->bx, ix, iyp, iz) + A(o->bz, ix, iy, izp) + A(o->bz, ix, iyp, izp);
        double byv = A(o->by, ix, iy, iz) + A(o->by, ixp, iy, iz) + A(o->bz, ix, iy, izp) + A(o->bz, ixp, iy, izp);
        double bzv = A(o->bz, ix, iy, iz) + A(o->bz, ixp, iy, iz) + A(o->bz, ix, iyp, iz) + A(o->bz, ixp, iyp, iz);
        bxv = 0.25 * bxv;
        byv = 0.25 * byv;
        bzv = 0.25 * bzv;
        const double bfieldsq = bxv * bxv + byv * byv + bzv * bzv;
        double t_v = 0.0;
        for (int iz2 = iz; iz2 <= iz + 1; ++iz2)
          for (int iy2 = iy; iy2 <= iy + 1; ++iy2)
            for (int ix2 = ix; ix2 <= ix + 1; ++ix2) {
              const double xi = A(o->xi_n, ix2, iy2, iz2);
              t_v = t_v + (A(o->energy, ix2, iy2, iz2) - (1.0 - xi) * ip) / (2.0 - xi);
            }
        t_v = t_v * gm1 / 8.0;
        const double xi_v = orc_saha_neutral(&o->p, t_v, rho_v);
        const double f = dmax(1.0 - xi_v, NONE_ZERO);
        double ep = 0.0;
        if (f > 0) ep = o->p.eta_bar * xi_v / f * bfieldsq / (rho_v * rho_v) / sqrt(t_v);
        A(o->eta_perp, ix, iy, iz) = dmin(ep, 100.0);
      }
}

int orc_neutral_fraction(orc_state *o) {
  if (!o->xi_n) return LARE_E_STATE;
  neutral_fraction(o);
  return o->fault;
}

int orc_perpendicular_resistivity(orc_state *o) {
  if (!o->eta_perp) return LARE_E_STATE;
  perpendicular_resistivity(o);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Conduction (conduct.f90)                                                  */
/* ------------------------------------------------------------------------ */
/* temperature of conduct.f90:72-73, (gamma-1)/(2-xi)*(e - (1-xi)*ionise_pot),
 * with the neutral fraction of cell q (uniform xi_n under EOS_IDEAL) */
static inline double xi_at(orc_state *o, long q) { return o->xi_n ? o->xi_n[q] : o->p.xi_n; }
static inline double temp_at(orc_state *o, double e, long q) {
  const double xi = xi_at(o, q);
  return (o->p.gamma - 1.0) / (2.0 - xi) * (e - (1.0 - xi) * o->ip);
}

/* conduct.f90:56-143 calc_s_stages */
static void calc_s_stages(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double k0 = o->p.kappa_0;
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) A(o->temperature, ix, iy, iz) = temp_at(o, A(o->energy, ix, iy, iz), IDX(ix, iy, iz));
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) {
        double *T = o->temperature;
        double temp = orc_pow25(A(T, ix, iy, iz));
        double q_fs = o->p.flux_limiter * 42.85 * A(o->rho, ix, iy, iz) * orc_pow15(A(T, ix, iy, iz));
        double q_fs2 = q_fs * q_fs;
        double q_spx = -k0 * temp * (A(T, ix + 1, iy, iz) - A(T, ix - 1, iy, iz)) * 0.5 / DXB(ix);
        double q_spy = -k0 * temp * (A(T, ix, iy + 1, iz) - A(T, ix, iy - 1, iz)) * 0.5 / DYB(iy);
        double q_spz = -k0 * temp * (A(T, ix, iy, iz + 1) - A(T, ix, iy, iz - 1)) * 0.5 / DZB(iz);
        double q_sp2 = q_spx * q_spx + q_spy * q_spy + q_spz * q_spz;
        A(o->larsen, ix, iy, iz) = q_fs / sqrt(q_fs2 + q_sp2);
      }
  if (!o->p.heat_flux_limiter)
    for (int iz = 0; iz <= nz; ++iz)
      for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix) A(o->larsen, ix, iy, iz) = 1.0;
  double dt_parab = 1.0e10;
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        double kappa1 = k0 * A(o->larsen, ix, iy, iz);
        double temp = temp_at(o, A(o->energy, ix, iy, iz), IDX(ix, iy, iz));
        double r = A(o->rho, ix, iy, iz);
        double tp = orc_pow25(temp);
        double dt1 = r * SQ(DXB(ix)) / (2.0 * kappa1 * tp);
        double dt2 = r * SQ(DYB(iy)) / (2.0 * kappa1 * tp);
        double dt3 = r * SQ(DZB(iz)) / (2.0 * kappa1 * tp);
        dt_parab = dmin(dt_parab, dmin(dt1, dmin(dt2, dt3)));
      }
  dt_parab = DT_MULT * dt_parab / sqrt(3.0);
  double stages = 0.5 * (sqrt(9.0 + 16.0 * (o->dt / dt_parab)) - 1.0);
  int ns = (int)ceil(stages);
  if (ns % 2 == 0) ns = ns + 1;
  o->n_s_stages = ns;
}

/* conduct.f90:152-291 heat_flux (scatter form, reference order) */
static void heat_flux(orc_state *o, const double *T, double *flux) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double k0 = o->p.kappa_0;
  for (long q = 0; q < o->ntot; ++q)
    if (!isnan(flux[q])) flux[q] = 0.0;
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) {
        int ixp = ix + 1, ixm = ix - 1, iyp = iy + 1, iym = iy - 1, izp = iz + 1, izm = iz - 1;
        double L = A(o->larsen, ix, iy, iz);
        /* X flux */
        double byf = 0.25 * (A(o->by, ix, iy, iz) + A(o->by, ixp, iy, iz) + A(o->by, ix, iym, iz) + A(o->by, ixp, iym, iz));
        double bzf = 0.25 * (A(o->bz, ix, iy, iz) + A(o->bz, ixp, iy, iz) + A(o->bz, ix, iy, izm) + A(o->bz, ixp, iy, izm));
        double bxc = A(o->bx, ix, iy, iz);
        double modb = bxc * bxc + byf * byf + bzf * bzf + MIN_B;
        double tb = (A(T, ixp, iy, iz) + A(T, ix, iy, iz)) / 2.0;
        double tb_p = (A(T, ixp, iyp, iz) + A(T, ix, iyp, iz)) / 2.0;
        double tb_m = (A(T, ixp, iym, iz) + A(T, ix, iym, iz)) / 2.0;
        double tg_a1 = (tb_p - tb_m) / (DYC(iy) + DYC(iym));
        tb_p = (A(T, ixp, iy, izp) + A(T, ix, iy, izp)) / 2.0;
        tb_m = (A(T, ixp, iy, izm) + A(T, ix, iy, izm)) / 2.0;
        double tg_a2 = (tb_p - tb_m) / (DZC(iz) + DZC(izm));
        double tg = (A(T, ixp, iy, iz) - A(T, ix, iy, iz)) / DXC(ix);
        tb = dmax(tb, 0.0);
        double fc_sp = -L * k0 * orc_pow25(tb) * (bxc * (tg * bxc + tg_a1 * byf + tg_a2 * bzf) + tg * MIN_B) / modb;
        A(flux, ix, iy, iz) = A(flux, ix, iy, iz) - fc_sp / DXB(ix);
        A(flux, ixp, iy, iz) = A(flux, ixp, iy, iz) + fc_sp / DXB(ix);
        /* Y flux */
        double bxf = 0.25 * (A(o->bx, ix, iy, iz) + A(o->bx, ix, iyp, iz) + A(o->bx, ixm, iy, iz) + A(o->bx, ixm, iyp, iz));
        bzf = 0.25 * (A(o->bz, ix, iy, iz) + A(o->bz, ix, iyp, iz) + A(o->bz, ix, iy, izm) + A(o->bz, ix, iyp, izm));
        double byc = A(o->by, ix, iy, iz);
        modb = bxf * bxf + byc * byc + bzf * bzf + MIN_B;
        tb = (A(T, ix, iyp, iz) + A(T, ix, iy, iz)) / 2.0;
        tb_p = (A(T, ixp, iy, iz) + A(T, ixp, iyp, iz)) / 2.0;
        tb_m = (A(T, ixm, iyp, iz) + A(T, ixm, iy, iz)) / 2.0;
        tg_a1 = (tb_p - tb_m) / (DXC(ix) + DXC(ixm));
        tb_p = (A(T, ix, iyp, izp) + A(T, ix, iy, izp)) / 2.0;
        tb_m = (A(T, ix, iyp, izm) + A(T, ix, iy, izm)) / 2.0;
        tg_a2 = (tb_p - tb_m) / (DZC(iz) + DZC(izm));
        tg = (A(T, ix, iyp, iz) - A(T, ix, iy, iz)) / DYC(iy);
        tb = dmax(tb, 0.0);
        fc_sp = -L * k0 * orc_pow25(tb) * (byc * (tg * byc + tg_a1 * bxf + tg_a2 * bzf) + tg * MIN_B) / modb;
        A(flux, ix, iy, iz) = A(flux, ix, iy, iz) - fc_sp / DYB(iy);
        A(flux, ix, iyp, iz) = A(flux, ix, iyp, iz) + fc_sp / DYB(iy);
        /* Z flux */
        bxf = 0.25 * (A(o->bx, ix, iy, iz) + A(o->bx, ix, iy, izp) + A(o->bx, ixm, iy, iz) + A(o->bx, ixm, iy, izp));
        byf = 0.25 * (A(o->by, ix, iy, iz) + A(o->by, ix, iy, izp) + A(o->by, ix, iym, iz) + A(o->by, ix, iym, izp));
        double bzc = A(o->bz, ix, iy, iz);
        modb = bxf * bxf + byf * byf + bzc * bzc + MIN_B;
        tb = (A(T, ix, iy, izp) + A(T, ix, iy, iz)) / 2.0;
        tb_p = (A(T, ixp, iy, izp) + A(T, ixp, iy, iz)) / 2.0;
        tb_m = (A(T, ixm, iy, izp) + A(T, ixm, iy, iz)) / 2.0;
        tg_a1 = (tb_p - tb_m) / (DXC(ix) + DXC(ixm));
        tb_p = (A(T, ix, iyp, izp) + A(T, ix, iyp, iz)) / 2.0;
        tb_m = (A(T, ix, iym, izp) + A(T, ix, iym, iz)) / 2.0;
        tg_a2 = (tb_p - tb_m) / (DYC(iy) + DYC(iym));
        tg = (A(T, ix, iy, izp) - A(T, ix, iy, iz)) / DZC(iz);
        tb = dmax(tb, 0.0);
        fc_sp = -L * k0 * orc_pow25(tb) * (bzc * (tg * bzc + tg_a1 * bxf + tg_a2 * byf) + tg * MIN_B) / modb;
        A(flux, ix, iy, iz) = A(flux, ix, iy, iz) - fc_sp / DZB(iz);
        A(flux, ix, iy, izp) = A(flux, ix, iy, izp) + fc_sp / DZB(iz);
      }
}

/* conduct.f90:299-416 heat_conduct_sts2 (RKL2, Meyer et al. 2012) */
static void heat_conduct_sts2(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz, s = o->n_s_stages;
  double dt = o->dt;
  double *flux = cellbox(o);
  double *Y[4];
  for (int j = 0; j < 4; ++j) Y[j] = cellbox(o);
  double *Lc = box(o, 1, nx, 1, ny, 1, nz);
  double *a = (double *)calloc(s + 1, sizeof(double)), *b = (double *)calloc(s + 1, sizeof(double));
  double *mu_t = (double *)calloc(s + 1, sizeof(double)), *mu = (double *)calloc(s + 1, sizeof(double));
  double *nu = (double *)calloc(s + 1, sizeof(double)), *gt = (double *)calloc(s + 1, sizeof(double));
  double ds = (double)s;
  double omega_1 = 4.0 / (ds * ds + ds - 2.0);
  for (int j = 0; j <= 2 && j <= s; ++j) b[j] = 1.0 / 3.0;
  for (int j = 3; j <= s; ++j) {
    double dj = (double)j;
    b[j] = (dj * dj + dj - 2.0) / (2.0 * dj * (dj + 1.0));
  }
  for (int j = 0; j <= s; ++j) a[j] = 1.0 - b[j];
  mu_t[1] = omega_1 / 3.0;
  for (int j = 2; j <= s; ++j) {
    double dj = (double)j;
    mu_t[j] = ((2.0 * dj - 1.0) / dj) * omega_1 * (b[j] / b[j - 1]);
    mu[j] = ((2.0 * dj - 1.0) / dj) * (b[j] / b[j - 1]);
    nu[j] = -1.0 * ((dj - 1.0) / dj) * (b[j] / b[j - 2]);
    gt[j] = -1.0 * a[j - 1] * mu_t[j];
  }
  for (int j = 0; j < 4; ++j) copy_all(o, Y[j], o->energy);
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) A(o->temperature, ix, iy, iz) = temp_at(o, A(Y[0], ix, iy, iz), IDX(ix, iy, iz));
  heat_flux(o, o->temperature, flux);
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        A(Lc, ix, iy, iz) = (1.0 / A(o->rho, ix, iy, iz)) * A(flux, ix, iy, iz);
        double c0 = mu_t[1] * dt * A(Lc, ix, iy, iz);
        A(Y[1], ix, iy, iz) = A(Y[0], ix, iy, iz) + c0;
      }
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        A(Y[2], ix, iy, iz) = A(Y[1], ix, iy, iz);
        A(Y[1], ix, iy, iz) = A(Y[0], ix, iy, iz);
      }
  for (int j = 2; j <= s; ++j) {
    for (int iz = -1; iz <= nz + 2; ++iz)
      for (int iy = -1; iy <= ny + 2; ++iy)
        for (int ix = -1; ix <= nx + 2; ++ix) A(o->temperature, ix, iy, iz) = temp_at(o, A(Y[2], ix, iy, iz), IDX(ix, iy, iz));
    temperature_bcs(o);
    heat_flux(o, o->temperature, flux);
    for (int iz = 1; iz <= nz; ++iz)
      for (int iy = 1; iy <= ny; ++iy)
        for (int ix = 1; ix <= nx; ++ix) {
          double c0 = gt[j] * dt * A(Lc, ix, iy, iz);
          double Lj = (1.0 / A(o->rho, ix, iy, iz)) * A(flux, ix, iy, iz);
          double c1 = mu_t[j] * dt * Lj;
          double y3 = mu[j] * A(Y[2], ix, iy, iz) + nu[j] * A(Y[1], ix, iy, iz) + (1.0 - mu[j] - nu[j]) * A(Y[0], ix, iy, iz);
          A(Y[3], ix, iy, iz) = y3 + c1 + c0;
        }
    if (j < s) {
      copy_all(o, Y[1], Y[2]);
      copy_all(o, o->energy, Y[3]);
      cell_bcs(o, o->energy);
      if (o->pi_eos) neutral_fraction(o);
      copy_all(o, Y[2], o->energy);
      for (int iz = 1; iz <= nz; ++iz)
        for (int iy = 1; iy <= ny; ++iy)
          for (int ix = 1; ix <= nx; ++ix) A(Y[3], ix, iy, iz) = 0.0;
    }
  }
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) A(o->energy, ix, iy, iz) = A(Y[3], ix, iy, iz);
  cell_bcs(o, o->energy);
  free(a); free(b); free(mu_t); free(mu); free(nu); free(gt);
  free(Lc); free(flux);
  for (int j = 0; j < 4; ++j) free(Y[j]);
}

/* conduct.f90:39-48 conduct_heat */
void orc_conduct_heat(orc_state *o) {
  calc_s_stages(o);
  heat_conduct_sts2(o);
}

/* ------------------------------------------------------------------------ */
/* Lagrangian step (lagran.F90)                                              */
/* ------------------------------------------------------------------------ */

/* lagran.F90:1176-1230 eta_calc */
void orc_eta_calc(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  if (!o->p.resistive_mhd) {
    for (int iz = -1; iz <= nz + 2; ++iz)
      for (int iy = -1; iy <= ny + 2; ++iy)
        for (int ix = -1; ix <= nx + 2; ++ix) A(o->eta, ix, iy, iz) = 0.0;
    return;
  }
  for (int iz = -1; iz <= nz + 1; ++iz) {
    int izp = iz + 1;
    for (int iy = -1; iy <= ny + 1; ++iy) {
      int iyp = iy + 1;
      for (int ix = -1; ix <= nx + 1; ++ix) {
        int ixp = ix + 1;
        double jx = (A(o->bz, ix, iyp, iz) - A(o->bz, ix, iy, iz)) / DYC(iy) - (A(o->by, ix, iy, izp) - A(o->by, ix, iy, iz)) / DZC(iz);
        double jxp = (A(o->bz, ixp, iyp, iz) - A(o->bz, ixp, iy, iz)) / DYC(iy) - (A(o->by, ixp, iy, izp) - A(o->by, ixp, iy, iz)) / DZC(iz);
        double jy = (A(o->bx, ix, iy, izp) - A(o->bx, ix, iy, iz)) / DZC(iz) - (A(o->bz, ixp, iy, iz) - A(o->bz, ix, iy, iz)) / DXC(ix);
        double jyp = (A(o->bx, ix, iyp, izp) - A(o->bx, ix, iyp, iz)) / DZC(iz) - (A(o->bz, ixp, iyp, iz) - A(o->bz, ix, iyp, iz)) / DXC(ix);
        double jz = (A(o->by, ixp, iy, iz) - A(o->by, ix, iy, iz)) / DXC(ix) - (A(o->bx, ix, iyp, iz) - A(o->bx, ix, iy, iz)) / DYC(iy);
        double jzp = (A(o->by, ixp, iy, izp) - A(o->by, ix, iy, izp)) / DXC(ix) - (A(o->bx, ix, iyp, izp) - A(o->bx, ix, iy, izp)) / DYC(iy);
        jx = (jx + jxp) * 0.5;
        jy = (jy + jyp) * 0.5;
        jz = (jz + jzp) * 0.5;
        if (sqrt(jx * jx + jy * jy + jz * jz) > o->p.j_max)
          A(o->eta, ix, iy, iz) = o->p.eta_background + o->p.eta0;
        else
          A(o->eta, ix, iy, iz) = o->p.eta_background;
      }
    }
  }
}

/* vertex current of rkstep / resistive_effects (lagran.F90:1479-1494) */
static inline void vertex_j(orc_state *o, int ix, int iy, int iz, double *jx, double *jy, double *jz) {
  int ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
  double jx1 = (A(o->bz, ix, iyp, iz) - A(o->bz, ix, iy, iz)) / DYC(iy) - (A(o->by, ix, iy, izp) - A(o->by, ix, iy, iz)) / DZC(iz);
  double jx2 = (A(o->bz, ixp, iyp, iz) - A(o->bz, ixp, iy, iz)) / DYC(iy) - (A(o->by, ixp, iy, izp) - A(o->by, ixp, iy, iz)) / DZC(iz);
  double jy1 = (A(o->bx, ix, iy, izp) - A(o->bx, ix, iy, iz)) / DZC(iz) - (A(o->bz, ixp, iy, iz) - A(o->bz, ix, iy, iz)) / DXC(ix);
  double jy2 = (A(o->bx, ix, iyp, izp) - A(o->bx, ix, iyp, iz)) / DZC(iz) - (A(o->bz, ixp, iyp, iz) - A(o->bz, ix, iyp, iz)) / DXC(ix);
  double jz1 = (A(o->by, ixp, iy, iz) - A(o->by, ix, iy, iz)) / DXC(ix) - (A(o->bx, ix, iyp, iz) - A(o->bx, ix, iy, iz)) / DYC(iy);
  double jz2 = (A(o->by, ixp, iy, izp) - A(o->by, ix, iy, izp)) / DXC(ix) - (A(o->bx, ix, iyp, izp) - A(o->bx, ix, iy, izp)) / DYC(iy);
  *jx = (jx1 + jx2) * 0.5;
  *jy = (jy1 + jy2) * 0.5;
  *jz = (jz1 + jz2) * 0.5;
}

/* lagran.F90:1239-1452 resistive_effects with rkstep :1460-1580 and bstep
 * :1582-1635. */
/* rkstep: EMFs and curlb at nodes 0..n from the current B and eta.  With
 * Cowling resistivity (:1504-1577) the vertex current splits into the part
 * along the vertex field and the rest; eta_perp adds to the resistivity of
 * the perpendicular part. */
static void rkstep(orc_state *o, double *fxr, double *fyr, double *fzr, double *curlb) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) {
        double jx, jy, jz;
        vertex_j(o, ix, iy, iz, &jx, &jy, &jz);
        double e = A(o->eta, ix, iy, iz);
        if (!o->p.cowling_resistivity) {
          A(fxr, ix, iy, iz) = -jx * e * DXC(ix) * 0.5;
          A(fyr, ix, iy, iz) = -jy * e * DYC(iy) * 0.5;
          A(fzr, ix, iy, iz) = -jz * e * DZC(iz) * 0.5;
          A(curlb, ix, iy, iz) = e * (jx * jx + jy * jy + jz * jz);
          continue;
        }
        const int ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
        const double bxv = (A(o->bx, ix, iy, iz) + A(o->bx, ix, iyp, iz) + A(o->bx, ix, iy, izp) + A(o->bx, ix, iyp, izp)) * 0.25;
        const double byv = (A(o->by, ix, iy, iz) + A(o->by, ixp, iy, iz) + A(o->by, ix, iy, izp) + A(o->by, ixp, iy, izp)) * 0.25;
        const double bzv = (A(o->bz, ix, iy, iz) + A(o->bz, ixp, iy, iz) + A(o->bz, ix, iyp, iz) + A(o->bz, ixp, iyp, iz)) * 0.25;
        const double magn_b = bxv * bxv + byv * byv + bzv * bzv;
        double jpx = 0.0, jpy = 0.0, jpz = 0.0;
        if (magn_b > NONE_ZERO) {
          jpx = (jx * bxv + jy * byv + jz * bzv) * bxv / magn_b;
          jpy = (jx * bxv + jy * byv + jz * bzv) * byv / magn_b;
          jpz = (jx * bxv + jy * byv + jz * bzv) * bzv / magn_b;
        }
        const double jqx = jx - jpx, jqy = jy - jpy, jqz = jz - jpz;
        const double mpar = sqrt(jpx * jpx + jpy * jpy + jpz * jpz);
        const double mperp = sqrt(jqx * jqx + jqy * jqy + jqz * jqz);
        const double ep = A(o->eta_perp, ix, iy, iz);
        A(curlb, ix, iy, iz) = e * (mpar * mpar) + (ep + e) * (mperp * mperp);
        A(fxr, ix, iy, iz) = -((jpx + jqx) * e + jqx * ep) * DXC(ix) * 0.5;
        A(fyr, ix, iy, iz) = -((jpy + jqy) * e + jqy * ep) * DYC(iy) * 0.5;
        A(fzr, ix, iy, iz) = -((jpz + jqz) * e + jqz * ep) * DZC(iz) * 0.5;
      }
}

/* bstep(kx, ky, kz, dt) (:1582-1635): B = B1 + curl of the EMFs, then
 * bfield_bcs */
static void bstep(orc_state *o, const double *fxr, const double *fyr, const double *fzr, double dt) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  for (int iz = 1; iz <= nz; ++iz) {
    int izm = iz - 1;
    for (int iy = 1; iy <= ny; ++iy) {
      int iym = iy - 1;
      double area = DYB(iy) * DZB(iz);
      for (int ix = 0; ix <= nx; ++ix)
        A(o->bx, ix, iy, iz) = A(o->bx1, ix, iy, iz) +
            (A(fzr, ix, iy, iz) - A(fzr, ix, iym, iz) + A(fzr, ix, iy, izm) - A(fzr, ix, iym, izm) - A(fyr, ix, iy, iz) + A(fyr, ix, iy, izm) -
             A(fyr, ix, iym, iz) + A(fyr, ix, iym, izm)) * dt / area;
    }
  }
  for (int iz = 1; iz <= nz; ++iz) {
    int izm = iz - 1;
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        int ixm = ix - 1;
        double area = DXB(ix) * DZB(iz);
        A(o->by, ix, iy, iz) = A(o->by1, ix, iy, iz) +
            (A(fxr, ix, iy, iz) - A(fxr, ix, iy, izm) + A(fxr, ixm, iy, iz) - A(fxr, ixm, iy, izm) - A(fzr, ix, iy, iz) + A(fzr, ixm, iy, iz) -
             A(fzr, ix, iy, izm) + A(fzr, ixm, iy, izm)) * dt / area;
      }
  }
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy) {
      int iym = iy - 1;
      for (int ix = 1; ix <= nx; ++ix) {
        int ixm = ix - 1;
        double area = DXB(ix) * DYB(iy);
        A(o->bz, ix, iy, iz) = A(o->bz1, ix, iy, iz) +
            (A(fyr, ix, iy, iz) - A(fyr, ixm, iy, iz) + A(fyr, ix, iym, iz) - A(fyr, ixm, iym, iz) - A(fxr, ix, iy, iz) + A(fxr, ix, iym, iz) -
             A(fxr, ixm, iy, iz) + A(fxr, ixm, iym, iz)) * dt / area;
      }
    }
  bfield_bcs(o);
}

/* total_ohmic_heating increments of curlb c over nodes 0..n (:1309-1328) */
static void ohmic_total(orc_state *o, const double *c, double dt) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) {
        double w1 = dt * DXC(ix) * DYC(iy) * DZC(iz) * A(c, ix, iy, iz);
        if (ix == 0 || ix == nx) w1 = w1 * 0.5;
        if (iy == 0 || iy == ny) w1 = w1 * 0.5;
        if (iz == 0 || iz == nz) w1 = w1 * 0.5;
        o->toh = o->toh + w1;
      }
}

/* the fourth-order Runge-Kutta substep of FOURTHORDER (:1329-1415): four
 * rksteps at B1, B1 + dt/2 k1, B1 + dt/2 k2 and B1 + dt k3; B = B1 + dt/6
 * (k1 + 2 k2 + 2 k3 + k4), and the heating of the weighted curlb.  No
 * ohmic_dep, cooling or energy floor in this variant. */
static void resistive_rk4(orc_state *o, double *fxr, double *fyr, double *fzr, double *curlb) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  const double dt = o->dt, half_dt = dt * 0.5, dt6 = dt * (1.0 / 6.0);
  double *k[4][4];
  for (int q = 0; q < 4; ++q)
    for (int c = 0; c < 4; ++c) k[q][c] = box(o, 0, nx, 0, ny, 0, nz);
  const double h[3] = {half_dt, half_dt, dt};
  for (int q = 0; q < 4; ++q) {
    if (q > 0) {
      bstep(o, k[q - 1][0], k[q - 1][1], k[q - 1][2], h[q - 1]);
      rkstep(o, fxr, fyr, fzr, curlb);
    }
    copy_all(o, k[q][0], fxr);
    copy_all(o, k[q][1], fyr);
    copy_all(o, k[q][2], fzr);
    copy_all(o, k[q][3], curlb);
  }
  for (int c = 0; c < 4; ++c)
    for (int iz = 0; iz <= nz; ++iz)
      for (int iy = 0; iy <= ny; ++iy)
        for (int ix = 0; ix <= nx; ++ix)
          A(k[0][c], ix, iy, iz) = A(k[0][c], ix, iy, iz) + 2.0 * A(k[1][c], ix, iy, iz) + 2.0 * A(k[2][c], ix, iy, iz) +
                                   A(k[3][c], ix, iy, iz);
  bstep(o, k[0][0], k[0][1], k[0][2], dt6);
  const double *c1 = k[0][3];
  for (int iz = 1; iz <= nz; ++iz) {
    int izm = iz - 1;
    for (int iy = 1; iy <= ny; ++iy) {
      int iym = iy - 1;
      for (int ix = 1; ix <= nx; ++ix) {
        int ixm = ix - 1;
        A(o->energy, ix, iy, iz) = A(o->energy, ix, iy, iz) +
            (A(c1, ix, iy, iz) + A(c1, ixm, iy, iz) + A(c1, ix, iym, iz) + A(c1, ixm, iym, iz) + A(c1, ix, iy, izm) +
             A(c1, ixm, iy, izm) + A(c1, ix, iym, izm) + A(c1, ixm, iym, izm)) * dt6 / (8.0 * A(o->rho, ix, iy, iz));
      }
    }
  }
  cell_bcs(o, o->energy);
  ohmic_total(o, c1, dt6);
  for (int q = 0; q < 4; ++q)
    for (int c = 0; c < 4; ++c) free(k[q][c]);
}

static void resistive_effects(orc_state *o, double *fxr, double *fyr, double *fzr, double *curlb) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double dt = o->dt;
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) {
        A(o->bx1, ix, iy, iz) = A(o->bx, ix, iy, iz);
        A(o->by1, ix, iy, iz) = A(o->by, ix, iy, iz);
        A(o->bz1, ix, iy, iz) = A(o->bz, ix, iy, iz);
      }
  rkstep(o, fxr, fyr, fzr, curlb);
  if (o->p.fourth_order) {
    resistive_rk4(o, fxr, fyr, fzr, curlb);
    return;
  }
  bstep(o, fxr, fyr, fzr, dt);
  /* ohmic heating :1270-1288 */
  for (int iz = 1; iz <= nz; ++iz) {
    int izm = iz - 1;
    for (int iy = 1; iy <= ny; ++iy) {
      int iym = iy - 1;
      for (int ix = 1; ix <= nx; ++ix) {
        int ixm = ix - 1;
        double lh = (A(curlb, ix, iy, iz) + A(curlb, ixm, iy, iz) + A(curlb, ix, iym, iz) + A(curlb, ixm, iym, iz) + A(curlb, ix, iy, izm) +
                     A(curlb, ixm, iy, izm) + A(curlb, ix, iym, izm) + A(curlb, ixm, iym, izm)) * dt / (8.0 * A(o->rho, ix, iy, iz));
        A(o->ohmic_dep, ix, iy, iz) = A(o->ohmic_dep, ix, iy, iz) + lh;
        A(o->energy, ix, iy, iz) = A(o->energy, ix, iy, iz) + lh;
      }
    }
  }
  if (o->p.cooling_term) {
    /* :1290-1302.  The loop sets no ixm or iym: they keep the last values of
     * the heating loop above, nx - 1 and ny - 1 */
    const int ixm = nx - 1, iym = ny - 1;
    const double aav = o->p.alpha_av;
    for (int iz = 1; iz <= nz; ++iz)
      for (int iy = 1; iy <= ny; ++iy)
        for (int ix = 1; ix <= nx; ++ix) {
          double lh = (A(curlb, ix, iy, iz) + A(curlb, ixm, iy, iz) + A(curlb, ix, iym, iz) + A(curlb, ixm, iym, iz)) *
                      dt / (4.0 * A(o->rho, ix, iy, iz));
          A(o->cool_term_b, ix, iy, iz) = aav * lh + (1.0 - aav) * A(o->cool_term_b, ix, iy, iz);
          A(o->energy, ix, iy, iz) = A(o->energy, ix, iy, iz) - A(o->cool_term_b, ix, iy, iz);
        }
  }
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) A(o->energy, ix, iy, iz) = dmax(A(o->energy, ix, iy, iz), 0.0);
  cell_bcs(o, o->energy);
  ohmic_total(o, curlb, dt);
  /* jx_r/jy_r/jz_r and the final rkstep (:1417-1445) only feed output. */
}

/* ------------------------------------------------------------------------ */
/* Open boundaries (openboundary.f90)                                        */
/* ------------------------------------------------------------------------ */

/* The state of open_bcs for one boundary point: index 1 the interior values
 * carried out by the Riemann invariants, index 0 the boundary values the
 * solver sets, and the far field behind the boundary (openboundary.f90:23-27,
 * with the normal direction first: "x" is the normal component). */
typedef struct {
  double vxbc[2], vybc[2], vzbc[2], bxbc[2], bybc[2], bzbc[2], rbc[2], ebc[2];
  double pfar, rhofar, uxfar, uyfar, uzfar, bxfar, byfar, bzfar, pbc, vnorm;
} obc_t;

/* open_bcs_fast (openboundary.f90:327-382): a perpendicular far field */
static void open_bcs_fast(obc_t *b, double gamma) {
  const double c0far = sqrt(gamma * b->pfar / b->rhofar);
  const double ctfar = sqrt((b->byfar * b->byfar + b->bzfar * b->bzfar) / b->rhofar);
  const double cffar = sqrt(c0far * c0far + ctfar * ctfar);
  const double c0 = sqrt(gamma * (gamma - 1.0) * b->ebc[1]);
  const double ct = sqrt((b->bybc[1] * b->bybc[1] + b->bzbc[1] * b->bzbc[1]) / b->rbc[1]);
  const double cf = sqrt(c0 * c0 + ct * ct);
  const double vtest[3] = {b->vnorm + cf, b->vnorm - cf, b->vnorm};
  double pstar[3], pmagstar[3], vstar[3], rhostar[3], bystar[3], bzstar[3];
  for (int i = 0; i < 3; ++i) {
    if (vtest[i] >= 0.0) {
      pstar[i] = b->pbc + 0.5 * (b->bybc[1] * b->bybc[1] + b->bzbc[1] * b->bzbc[1]);
      pmagstar[i] = 0.5 * (b->bybc[1] * b->bybc[1] + b->bzbc[1] * b->bzbc[1]);
      vstar[i] = b->vxbc[1];
      rhostar[i] = b->rbc[1];
      bystar[i] = b->bybc[1];
      bzstar[i] = b->bzbc[1];
    } else {
      pstar[i] = b->pfar + 0.5 * (b->byfar * b->byfar + b->bzfar * b->bzfar);
      pmagstar[i] = 0.5 * (b->byfar * b->byfar + b->bzfar * b->bzfar);
      vstar[i] = b->uxfar;
      rhostar[i] = b->rhofar;
      bystar[i] = b->byfar;
      bzstar[i] = b->bzfar;
    }
  }
  b->bxbc[0] = 0.5 * (b->bxbc[1] + b->bxfar);
  b->bybc[0] = 0.5 * (bystar[0] + bystar[1]);
  b->bzbc[0] = 0.5 * (bzstar[0] + bzstar[1]);
  double pg = 0.5 * (pstar[0] + pstar[1] + b->rhofar * cffar * (vstar[0] - vstar[1])) - 0.5 * (pmagstar[0] + pmagstar[1]);
  pg = dmax(pg, 0.0);
  const double rhog = rhostar[2] + (pg - (pstar[2] - pmagstar[2])) / (c0far * c0far);
  b->rbc[0] = dmax(rhog, NONE_ZERO);
  b->ebc[0] = pg / (gamma - 1.0) / b->rbc[0];
}

/* open_bcs_alfven (openboundary.f90:386-455): a far field along the normal */
static void open_bcs_alfven(obc_t *b, double gamma) {
  const double lambdayfar = -b->bxfar * b->byfar, lambdazfar = -b->bxfar * b->bzfar;
  const double c0far = sqrt(gamma * b->pfar / b->rhofar);
  const double cxfar = sqrt(b->bxfar * b->bxfar / b->rhofar);
  const double c0 = sqrt(gamma * (gamma - 1.0) * b->ebc[1]);
  const double cx = sqrt(b->bxbc[1] * b->bxbc[1] / b->rbc[1]);
  const double vtest[5] = {b->vnorm + c0, b->vnorm - c0, b->vnorm + cx, b->vnorm - cx, b->vnorm};
  double pstar[5], pmagstar[5], lys[5], lzs[5], uxs[5], uys[5], uzs[5], rhostar[5];
  for (int i = 0; i < 5; ++i) {
    if (vtest[i] >= 0.0) {
      pstar[i] = b->pbc + 0.5 * (b->bybc[1] * b->bybc[1] + b->bzbc[1] * b->bzbc[1] - b->bxbc[1] * b->bxbc[1]);
      pmagstar[i] = 0.5 * (b->bybc[1] * b->bybc[1] + b->bzbc[1] * b->bzbc[1] - b->bxbc[1] * b->bxbc[1]);
      lys[i] = -b->bybc[1] * b->bxbc[1];
      lzs[i] = -b->bzbc[1] * b->bxbc[1];
      uxs[i] = b->vxbc[1];
      uys[i] = b->vybc[1];
      uzs[i] = b->vzbc[1];
      rhostar[i] = b->rbc[1];
    } else {
      pstar[i] = b->pfar + 0.5 * (b->byfar * b->byfar + b->bzfar * b->bzfar - b->bxfar * b->bxfar);
      pmagstar[i] = 0.5 * (b->byfar * b->byfar + b->bzfar * b->bzfar - b->bxfar * b->bxfar);
      lys[i] = lambdayfar;
      lzs[i] = lambdazfar;
      uxs[i] = b->uxfar;
      uys[i] = b->uyfar;
      uzs[i] = b->uzfar;
      rhostar[i] = b->rhofar;
    }
  }
  b->bxbc[0] = 0.5 * (b->bxbc[1] + b->bxfar);
  if (fabs(b->bxbc[0]) <= NONE_ZERO) b->bxbc[0] = NONE_ZERO;
  double lambdag = 0.5 * (lys[2] + lys[3] + b->rhofar * cxfar * (uys[2] - uys[3]));
  b->bybc[0] = -lambdag / b->bxbc[0];
  lambdag = 0.5 * (lzs[2] + lzs[3] + b->rhofar * cxfar * (uzs[2] - uzs[3]));
  b->bzbc[0] = -lambdag / b->bxbc[0];
  double pg = 0.5 * (pstar[0] + pstar[1] + b->rhofar * c0far * (uxs[0] - uxs[1])) -
              0.5 * (b->bybc[0] * b->bybc[0] + b->bzbc[0] * b->bzbc[0] - b->bxbc[0] * b->bxbc[0]);
  pg = dmax(pg, NONE_ZERO);
  const double rhog = rhostar[4] + (pg - (pstar[4] - pmagstar[4])) / (c0far * c0far);
  b->rbc[0] = dmax(rhog, NONE_ZERO);
  b->ebc[0] = pg / (gamma - 1.0) / b->rbc[0];
}

/* element of arr at the transverse position q with layer l along axis a */
static inline double *obc_at(orc_state *o, double *arr, int *q, int a, int l) {
  q[a] = l;
  return &A(arr, q[0], q[1], q[2]);
}

/* open_bcs (openboundary.f90:30-323): the six faces in the reference's
 * order, each over the transverse points 0..n+1.  Face (a, side): the
 * normal component is axis a, signed so that it points out of the domain
 * (the low side flips it); the two transverse slots are the reference's:
 * for x (y, z), for y (x, z), for z (y, x).  Layers along a: the boundary
 * node and normal face (0 or n), the first interior cell (1 or n), the far
 * field in the second ghost layer (-1 / -2 or n+2), and the values set in
 * the first ghost layer (normal face -1 or n+1, cells 0 or n+1). */
static void open_bcs(orc_state *o) {
  const double gm = o->p.gamma;
  double *v[3] = {o->vx, o->vy, o->vz}, *bf[3] = {o->bx, o->by, o->bz};
  for (int a = 0; a < 3; ++a)
    for (int hi = 0; hi <= 1; ++hi) {
      if (o->bc[2 * a + hi] != LARE_BC_OPEN) continue;
      const int n = nax(o, a);
      const double sg = hi ? 1.0 : -1.0;
      /* transverse slots (y, z) of the face */
      const int ty = a == 1 ? 0 : 1, tz = a == 2 ? 0 : 2;
      const int lb = hi ? n : 0, lc = hi ? n : 1, lfc = hi ? n + 2 : -1, lfn = hi ? n + 2 : -2;
      const int lset_n = hi ? n + 1 : -1, lset_c = hi ? n + 1 : 0;
      const int u = a == 0 ? 1 : 0, w = a == 2 ? 1 : 2;
      const int nu = nax(o, u), nw = nax(o, w);
      for (int jw = 0; jw <= nw + 1; ++jw)
        for (int ju = 0; ju <= nu + 1; ++ju) {
          int q[3];
          q[u] = ju; q[w] = jw;
#define AT(arr, l) (*obc_at(o, (arr), q, a, (l)))
          obc_t b;
          b.vxbc[1] = sg * AT(v[a], lb);
          b.vybc[1] = AT(v[ty], lb);
          b.vzbc[1] = AT(v[tz], lb);
          b.bxbc[1] = sg * AT(bf[a], lb);
          b.bybc[1] = AT(bf[ty], lc);
          b.bzbc[1] = AT(bf[tz], lc);
          b.rbc[1] = AT(o->rho, lc);
          b.ebc[1] = AT(o->energy, lc);
          b.pbc = (gm - 1.0) * AT(o->energy, lc) * AT(o->rho, lc);
          b.pfar = (gm - 1.0) * AT(o->energy, lfc) * AT(o->rho, lfc);
          b.uxfar = sg * AT(v[a], lfn);
          b.uyfar = AT(v[ty], lfn);
          b.uzfar = AT(v[tz], lfn);
          b.bxfar = sg * AT(bf[a], lfn);
          b.byfar = AT(bf[ty], lfc);
          b.bzfar = AT(bf[tz], lfc);
          b.rhofar = AT(o->rho, lfc);
          b.vnorm = sg * AT(v[a], lb);
          const double bperp = sqrt(b.byfar * b.byfar + b.bzfar * b.bzfar);
          if (fabs(bperp) <= NONE_ZERO) open_bcs_alfven(&b, gm);
          else open_bcs_fast(&b, gm);
          AT(bf[a], lset_n) = sg * b.bxbc[0];
          AT(bf[ty], lset_c) = b.bybc[0];
          AT(bf[tz], lset_c) = b.bzbc[0];
          AT(o->rho, lset_c) = b.rbc[0];
          AT(o->energy, lset_c) = b.ebc[0];
#undef AT
        }
    }
}

/* lagran.F90:483-524 viscous_damping */
static void viscous_damping(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double *q[3];
  const double *v[3] = {o->vx, o->vy, o->vz};
  for (int c = 0; c < 3; ++c) q[c] = box(o, 0, nx + 1, 0, ny + 1, 0, nz + 1);
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) {
        int ixp = ix + 1, ixm = ix - 1, iyp = iy + 1, iym = iy - 1, izp = iz + 1, izm = iz - 1;
        for (int c = 0; c < 3; ++c) {
          const double *u = v[c];
          double u0 = A(u, ix, iy, iz);
          A(q[c], ix, iy, iz) = A(o->visc3, ix, iy, iz) *
              (((A(u, ixp, iy, iz) - u0) / DXB(ixp) - (u0 - A(u, ixm, iy, iz)) / DXB(ix)) / DXC(ix) +
               ((A(u, ix, iyp, iz) - u0) / DYB(iyp) - (u0 - A(u, ix, iym, iz)) / DYB(iy)) / DYC(iy) +
               ((A(u, ix, iy, izp) - u0) / DZB(izp) - (u0 - A(u, ix, iy, izm)) / DZB(iz)) / DZC(iz));
        }
      }
  double *w[3] = {o->vx, o->vy, o->vz};
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix)
        for (int c = 0; c < 3; ++c) A(w[c], ix, iy, iz) = A(w[c], ix, iy, iz) + o->dt * A(q[c], ix, iy, iz);
  for (int c = 0; c < 3; ++c) free(q[c]);
}

/* edge_viscosity (lagran.F90:795-855), SHOCKLIMITER and SHOCKCOMPRESSION off */
static double edge_viscosity(orc_state *o, const double *cs_v, double dvdots, double dx, double dxm, double dxp,
                             int i1, int j1, int k1, int i2, int j2, int k2, int i0, int j0, int k0, int i3,
                             int j3, int k3) {
  if (o->p.shock_compression) dvdots = dmin(0.0, dvdots); /* SHOCKCOMPRESSION (:807-808) */
  else dvdots = -fabs(dvdots);
  double r1 = A(o->rho_vl, i1, j1, k1), r2 = A(o->rho_vl, i2, j2, k2);
  double rho_edge = 2.0 * r1 * r2 / (r1 + r2);
  double cs_edge = dmin(A(cs_v, i1, j1, k1), A(cs_v, i2, j2, k2));
  double dvx = A(o->vx, i1, j1, k1) - A(o->vx, i2, j2, k2);
  double dvy = A(o->vy, i1, j1, k1) - A(o->vy, i2, j2, k2);
  double dvz = A(o->vz, i1, j1, k1) - A(o->vz, i2, j2, k2);
  double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
  double dv = sqrt(dv2);
  double psi = 0.0;
  if (dv * o->dt / dx < 1.0e-14) dvdots = 0.0;
  else dvdots = dvdots / dv;
  if (o->p.shock_limiter) {
    /* SHOCKLIMITER (:831-845) */
    double dvxm = A(o->vx, i0, j0, k0) - A(o->vx, i1, j1, k1);
    double dvxp = A(o->vx, i2, j2, k2) - A(o->vx, i3, j3, k3);
    double dvym = A(o->vy, i0, j0, k0) - A(o->vy, i1, j1, k1);
    double dvyp = A(o->vy, i2, j2, k2) - A(o->vy, i3, j3, k3);
    double dvzm = A(o->vz, i0, j0, k0) - A(o->vz, i1, j1, k1);
    double dvzp = A(o->vz, i2, j2, k2) - A(o->vz, i3, j3, k3);
    double rl, rr;
    if (dv * o->dt / dx < 1.0e-14) {
      rl = 1.0;
      rr = 1.0;
    } else {
      rl = (dvxp * dvx + dvyp * dvy + dvzp * dvz) * dx / (dxp * dv2);
      rr = (dvxm * dvx + dvym * dvy + dvzm * dvz) * dx / (dxm * dv2);
    }
    psi = dmin(dmin(dmin(0.5 * (rr + rl), 2.0 * rl), 2.0 * rr), 1.0);
    psi = dmax(0.0, psi);
  }
  double vn = o->visc2_norm;
  double q_k_bar = rho_edge * (vn * dv + sqrt(vn * vn * dv2 + SQ(o->p.visc1 * cs_edge)));
  return q_k_bar * (1.0 - psi) * dvdots;
}

static inline double vdiff2(orc_state *o, int a0, int b0, int c0, int a1, int b1, int c1) {
  double dx = A(o->vx, a0, b0, c0) - A(o->vx, a1, b1, c1);
  double dy = A(o->vy, a0, b0, c0) - A(o->vy, a1, b1, c1);
  double dz = A(o->vz, a0, b0, c0) - A(o->vz, a1, b1, c1);
  return dx * dx + dy * dy + dz * dz;
}
static inline double vdot1(orc_state *o, int a0, int b0, int c0, int a1, int b1, int c1) {
  return (A(o->vx, a0, b0, c0) - A(o->vx, a1, b1, c1)) * (A(o->vx1, a0, b0, c0) - A(o->vx1, a1, b1, c1)) +
         (A(o->vy, a0, b0, c0) - A(o->vy, a1, b1, c1)) * (A(o->vy1, a0, b0, c0) - A(o->vy1, a1, b1, c1)) +
         (A(o->vz, a0, b0, c0) - A(o->vz, a1, b1, c1)) * (A(o->vz1, a0, b0, c0) - A(o->vz1, a1, b1, c1));
}
/* 12-edge sum of shock_viscosity :722-736 and shock_heating :916-928 */
static inline double edge_heat(orc_state *o, int ix, int iy, int iz, const double a[12]) {
  int ixm = ix - 1, iym = iy - 1, iyp = iy + 1, izp = iz + 1;
  double h = -0.25 * DYB(iy) * DZB(iz) * A(o->alpha1, ix, iy, iz) * a[0]
             - 0.25 * DXB(ix) * DZB(iz) * A(o->alpha2, ix, iy, iz) * a[1]
             - 0.25 * DYB(iy) * DZB(iz) * A(o->alpha1, ix, iyp, iz) * a[2]
             - 0.25 * DXB(ix) * DZB(iz) * A(o->alpha2, ixm, iy, iz) * a[3]
             - 0.25 * DYB(iy) * DZB(iz) * A(o->alpha1, ix, iy, izp) * a[4]
             - 0.25 * DXB(ix) * DZB(iz) * A(o->alpha2, ix, iy, izp) * a[5]
             - 0.25 * DYB(iy) * DZB(iz) * A(o->alpha1, ix, iyp, izp) * a[6]
             - 0.25 * DXB(ix) * DZB(iz) * A(o->alpha2, ixm, iy, izp) * a[7]
             - 0.25 * DXB(ix) * DYB(iy) * A(o->alpha3, ix, iy, iz) * a[8]
             - 0.25 * DXB(ix) * DYB(iy) * A(o->alpha3, ixm, iy, iz) * a[9]
             - 0.25 * DXB(ix) * DYB(iy) * A(o->alpha3, ixm, iym, iz) * a[10]
             - 0.25 * DXB(ix) * DYB(iy) * A(o->alpha3, ix, iym, iz) * a[11];
  return h / A(o->cv, ix, iy, iz);
}
/* node pairs of the 12 cell edges, in the order a1..a12 */
static void edge_pairs(int ix, int iy, int iz, int e[12][6]) {
  int ixm = ix - 1, iym = iy - 1, izm = iz - 1;
  int t[12][6] = {
      {ixm, iym, izm, ix, iym, izm}, {ix, iym, izm, ix, iy, izm}, {ix, iy, izm, ixm, iy, izm}, {ixm, iy, izm, ixm, iym, izm},
      {ixm, iym, iz, ix, iym, iz},   {ix, iym, iz, ix, iy, iz},   {ix, iy, iz, ixm, iy, iz},   {ixm, iy, iz, ixm, iym, iz},
      {ix, iy, izm, ix, iy, iz},     {ixm, iy, izm, ixm, iy, iz}, {ixm, iym, izm, ixm, iym, iz}, {ix, iym, izm, ix, iym, iz}};
  memcpy(e, t, sizeof(t));
}

/* lagran.F90:532-857 shock_viscosity */
static void shock_viscosity(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double *cs = cellbox(o);
  double *cs_v = box(o, -1, nx + 1, -1, ny + 1, -1, nz + 1);
  if (o->first_shock) {
    o->first_shock = 0;
    o->visc2_norm = 0.25 * (o->p.gamma + 1.0) * o->p.visc2;
  }
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) A(o->p_visc, ix, iy, iz) = 0.0;
  for (int iz = 0; iz <= nz + 1; ++iz)
    for (int iy = 0; iy <= ny + 1; ++iy)
      for (int ix = 0; ix <= nx + 1; ++ix) A(o->visc_heat, ix, iy, iz) = 0.0;
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int ix = -1; ix <= nx + 2; ++ix)
      for (int iy = -1; iy <= ny + 2; ++iy) {
        double rmin = dmax(A(o->rho, ix, iy, iz), NONE_ZERO);
        double b2 = SQ(A(o->bx1, ix, iy, iz)) + SQ(A(o->by1, ix, iy, iz)) + SQ(A(o->bz1, ix, iy, iz));
        /* gamma_boris of the previous set_dt (1 before the first) */
        A(cs, ix, iy, iz) = sqrt((o->p.gamma * A(o->pressure, ix, iy, iz) + A(o->gamma_boris, ix, iy, iz) * b2) / rmin);
      }
  for (int iz = -1; iz <= nz + 1; ++iz)
    for (int iy = -1; iy <= ny + 1; ++iy)
      for (int ix = -1; ix <= nx + 1; ++ix) {
        int ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
        double s = A(cs, ix, iy, iz) * A(o->cv, ix, iy, iz) + A(cs, ixp, iy, iz) * A(o->cv, ixp, iy, iz) +
                   A(cs, ix, iyp, iz) * A(o->cv, ix, iyp, iz) + A(cs, ixp, iyp, iz) * A(o->cv, ixp, iyp, iz) +
                   A(cs, ix, iy, izp) * A(o->cv, ix, iy, izp) + A(cs, ixp, iy, izp) * A(o->cv, ixp, iy, izp) +
                   A(cs, ix, iyp, izp) * A(o->cv, ix, iyp, izp) + A(cs, ixp, iyp, izp) * A(o->cv, ixp, iyp, izp);
        A(cs_v, ix, iy, iz) = 0.125 * s / A(o->cv_vl, ix, iy, iz);
      }
  for (int iz = 0; iz <= nz + 2; ++iz)
    for (int iy = 0; iy <= ny + 2; ++iy)
      for (int ix = 0; ix <= nx + 1; ++ix) {
        int i1 = ix - 1, j1 = iy - 1, k1 = iz - 1, i2 = ix, j2 = iy - 1, k2 = iz - 1;
        double dvdots = -(A(o->vx, i1, j1, k1) - A(o->vx, i2, j2, k2));
        A(o->alpha1, ix, iy, iz) = edge_viscosity(o, cs_v, dvdots, DXB(ix), DXB(ix - 1), DXB(ix + 1), i1, j1, k1, i2,
                                                  j2, k2, i1 - 1, j1, k1, i2 + 1, j2, k2);
      }
  for (int iz = 0; iz <= nz + 2; ++iz)
    for (int iy = 0; iy <= ny + 1; ++iy)
      for (int ix = -1; ix <= nx + 1; ++ix) {
        int i1 = ix, j1 = iy - 1, k1 = iz - 1, i2 = ix, j2 = iy, k2 = iz - 1;
        double dvdots = -(A(o->vy, i1, j1, k1) - A(o->vy, i2, j2, k2));
        A(o->alpha2, ix, iy, iz) = edge_viscosity(o, cs_v, dvdots, DYB(iy), DYB(iy - 1), DYB(iy + 1), i1, j1, k1, i2,
                                                  j2, k2, i1, j1 - 1, k1, i2, j2 + 1, k2);
      }
  for (int iz = 0; iz <= nz + 1; ++iz)
    for (int iy = -1; iy <= ny + 1; ++iy)
      for (int ix = -1; ix <= nx + 1; ++ix) {
        int i1 = ix, j1 = iy, k1 = iz - 1, i2 = ix, j2 = iy, k2 = iz;
        double dvdots = -(A(o->vz, i1, j1, k1) - A(o->vz, i2, j2, k2));
        A(o->alpha3, ix, iy, iz) = edge_viscosity(o, cs_v, dvdots, DZB(iz), DZB(iz - 1), DZB(iz + 1), i1, j1, k1, i2,
                                                  j2, k2, i1, j1, k1 - 1, i2, j2, k2 + 1);
      }
  for (int iz = 0; iz <= nz + 1; ++iz)
    for (int iy = 0; iy <= ny + 1; ++iy)
      for (int ix = 0; ix <= nx + 1; ++ix) {
        int e[12][6];
        double a[12];
        edge_pairs(ix, iy, iz, e);
        for (int q = 0; q < 12; ++q) a[q] = vdiff2(o, e[q][0], e[q][1], e[q][2], e[q][3], e[q][4], e[q][5]);
        double pv = A(o->p_visc, ix, iy, iz);
        pv = dmax(pv, -A(o->alpha1, ix, iy, iz) * sqrt(a[0]));
        pv = dmax(pv, -A(o->alpha2, ix, iy, iz) * sqrt(a[1]));
        pv = dmax(pv, -A(o->alpha3, ix, iy, iz) * sqrt(a[8]));
        A(o->p_visc, ix, iy, iz) = pv;
        A(o->visc_heat, ix, iy, iz) = edge_heat(o, ix, iy, iz, a);
      }
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) {
        int ixm = ix - 1, ixp = ix + 1, iym = iy - 1, iyp = iy + 1, izm = iz - 1, izp = iz + 1;
        double a1 = A(o->alpha1, ix, iyp, izp) * DYC(iy) * DZC(iz);
        double a2 = A(o->alpha1, ixp, iyp, izp) * DYC(iy) * DZC(iz);
        double a3 = A(o->alpha2, ix, iy, izp) * DXC(ix) * DZC(iz);
        double a4 = A(o->alpha2, ix, iyp, izp) * DXC(ix) * DZC(iz);
        double a5 = A(o->alpha3, ix, iy, iz) * DXC(ix) * DYC(iy);
        double a6 = A(o->alpha3, ix, iy, izp) * DXC(ix) * DYC(iy);
        const double *vv[3] = {o->vx, o->vy, o->vz};
        double *ff[3] = {o->fx_visc, o->fy_visc, o->fz_visc};
        for (int c = 0; c < 3; ++c) {
          const double *u = vv[c];
          double u0 = A(u, ix, iy, iz);
          A(ff[c], ix, iy, iz) = (a1 * (u0 - A(u, ixm, iy, iz)) + a2 * (u0 - A(u, ixp, iy, iz)) + a3 * (u0 - A(u, ix, iym, iz)) +
                                  a4 * (u0 - A(u, ix, iyp, iz)) + a5 * (u0 - A(u, ix, iy, izm)) + a6 * (u0 - A(u, ix, iy, izp))) /
                                 A(o->cv_vl, ix, iy, iz);
        }
      }
  free(cs);
  free(cs_v);
}

/* lagran.F90:861-937 shock_heating */
static void shock_heating(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  for (int iz = 0; iz <= nz + 1; ++iz)
    for (int iy = 0; iy <= ny + 1; ++iy)
      for (int ix = 0; ix <= nx + 1; ++ix) {
        int e[12][6];
        double a[12];
        edge_pairs(ix, iy, iz, e);
        for (int q = 0; q < 12; ++q) a[q] = vdot1(o, e[q][0], e[q][1], e[q][2], e[q][3], e[q][4], e[q][5]);
        A(o->visc_heat, ix, iy, iz) = edge_heat(o, ix, iy, iz, a);
      }
  for (int iz = 0; iz <= nz + 1; ++iz)
    for (int iy = 0; iy <= ny + 1; ++iy)
      for (int ix = 0; ix <= nx + 1; ++ix) A(o->visc_heat, ix, iy, iz) = dmax(A(o->visc_heat, ix, iy, iz), 0.0);
}

/* lagran.F90:1070-1166 set_dt (no cowling resistivity) */
static void set_dt(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  if (o->first_set_dt) {
    o->first_set_dt = 0;
    if (o->p.restart) {
      o->dt = o->p.dt_from_restart;
      return;
    }
  }
  double dt_local = LARGEST, dtr_local = LARGEST, dt3 = LARGEST;
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) A(o->gamma_boris, ix, iy, iz) = 1.0;
  const double va_max2 = o->p.va_max * o->p.va_max; /* setup.F90:80 */
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) {
        double w1 = SQ(A(o->bx, ix, iy, iz)) + SQ(A(o->by, ix, iy, iz)) + SQ(A(o->bz, ix, iy, iz));
        double rho0 = dmax(A(o->rho, ix, iy, iz), NONE_ZERO);
        double w2 = w1 / rho0;
        if (o->p.boris && w2 >= va_max2) A(o->gamma_boris, ix, iy, iz) = 1.0 / (1.0 + w2 / va_max2);
        double cs2 = (o->p.gamma * A(o->pressure, ix, iy, iz) + A(o->gamma_boris, ix, iy, iz) * w1) / rho0;
        double c_visc2 = A(o->p_visc, ix, iy, iz) / rho0;
        double length = min3(DXB(ix), DYB(iy), DZB(iz));
        double dt1 = length / (sqrt(c_visc2) + sqrt(cs2 + c_visc2));
        dt_local = dmin(dt_local, dt1);
        double dt2 = min3(DXB(ix) / dmax(fabs(A(o->vx, ix, iy, iz)), NONE_ZERO), DYB(iy) / dmax(fabs(A(o->vy, ix, iy, iz)), NONE_ZERO),
                          DZB(iz) / dmax(fabs(A(o->vz, ix, iy, iz)), NONE_ZERO));
        dt_local = dmin(dt_local, dt2);
        double dxlocal = 1.0 / (1.0 / SQ(DXB(ix)) + 1.0 / SQ(DYB(iy)) + 1.0 / SQ(DZB(iz)));
        if (o->p.cowling_resistivity)
          dt3 = 0.2 * dxlocal / dmax(dmax(A(o->eta, ix, iy, iz), A(o->eta_perp, ix, iy, iz)), NONE_ZERO);
        else if (o->p.resistive_mhd)
          dt3 = 0.2 * dxlocal / dmax(A(o->eta, ix, iy, iz), NONE_ZERO);
        dtr_local = dmin(dtr_local, dt3);
      }
  o->dt = DT_MULT * dt_local;
  o->dtr = DT_MULT * dtr_local;
  if (o->p.conduction) {
    calc_s_stages(o);
    int ss_limit = 60;
    if (o->n_s_stages >= ss_limit) {
      double f = (double)(2 * ss_limit * ss_limit - 9) / (double)(2 * o->n_s_stages * o->n_s_stages - 9);
      o->dt = o->dt * f;
      o->dtr = o->dtr * f;
    }
  }
  o->time = o->time + o->dt;
}

/* lagran.F90:208-478 predictor_corrector_step (non-CAUCHY) */
/* b_field_and_cv1_update with CAUCHY (lagran.F90:945-1048): cv1, and the
 * cell-centred B advanced by the Cauchy solution
 * B1 = (B1 + (B1 . grad) v dt/2) / (1 + div v dt/2); no predictor remap */
static void cauchy_b_update(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double *vx = o->vx, *vy = o->vy, *vz = o->vz;
  for (int iz = -1; iz <= nz + 2; ++iz) {
    int izm = iz - 1;
    for (int iy = -1; iy <= ny + 2; ++iy) {
      int iym = iy - 1;
      for (int ix = -1; ix <= nx + 2; ++ix) {
        int ixm = ix - 1;
        /* the normal derivatives at the faces of B, as without CAUCHY */
        double vxb = (A(vx, ix, iy, iz) + A(vx, ix, iym, iz) + A(vx, ix, iy, izm) + A(vx, ix, iym, izm)) * 0.25;
        double vxbm = (A(vx, ixm, iy, iz) + A(vx, ixm, iym, iz) + A(vx, ixm, iy, izm) + A(vx, ixm, iym, izm)) * 0.25;
        double vyb = (A(vy, ix, iy, iz) + A(vy, ixm, iy, iz) + A(vy, ix, iy, izm) + A(vy, ixm, iy, izm)) * 0.25;
        double vybm = (A(vy, ix, iym, iz) + A(vy, ixm, iym, iz) + A(vy, ix, iym, izm) + A(vy, ixm, iym, izm)) * 0.25;
        double vzb = (A(vz, ix, iy, iz) + A(vz, ixm, iy, iz) + A(vz, ix, iym, iz) + A(vz, ixm, iym, iz)) * 0.25;
        double vzbm = (A(vz, ix, iy, izm) + A(vz, ixm, iy, izm) + A(vz, ix, iym, izm) + A(vz, ixm, iym, izm)) * 0.25;
        double dvxdx = (vxb - vxbm) / DXB(ix);
        double dvydy = (vyb - vybm) / DYB(iy);
        double dvzdz = (vzb - vzbm) / DZB(iz);
        double dv = (dvxdx + dvydy + dvzdz) * o->dt2;
        A(o->cv1, ix, iy, iz) = A(o->cv, ix, iy, iz) * (1.0 + dv);
        /* the transverse derivatives (:989-1033) */
        vxb = (A(vx, ix, iy, iz) + A(vx, ixm, iy, iz) + A(vx, ix, iy, izm) + A(vx, ixm, iy, izm)) * 0.25;
        vxbm = (A(vx, ix, iym, iz) + A(vx, ixm, iym, iz) + A(vx, ix, iym, izm) + A(vx, ixm, iym, izm)) * 0.25;
        vyb = (A(vy, ix, iy, iz) + A(vy, ix, iym, iz) + A(vy, ix, iy, izm) + A(vy, ix, iym, izm)) * 0.25;
        vybm = (A(vy, ixm, iy, iz) + A(vy, ixm, iym, iz) + A(vy, ixm, iy, izm) + A(vy, ixm, iym, izm)) * 0.25;
        double dvxdy = (vxb - vxbm) / DYB(iy);
        double dvydx = (vyb - vybm) / DXB(ix);
        vxb = (A(vx, ix, iy, iz) + A(vx, ixm, iy, iz) + A(vx, ix, iym, iz) + A(vx, ixm, iym, iz)) * 0.25;
        vxbm = (A(vx, ix, iy, izm) + A(vx, ixm, iy, izm) + A(vx, ix, iym, izm) + A(vx, ixm, iym, izm)) * 0.25;
        vzb = (A(vz, ix, iy, iz) + A(vz, ix, iym, iz) + A(vz, ix, iy, izm) + A(vz, ix, iym, izm)) * 0.25;
        vzbm = (A(vz, ixm, iy, iz) + A(vz, ixm, iym, iz) + A(vz, ixm, iy, izm) + A(vz, ixm, iym, izm)) * 0.25;
        double dvxdz = (vxb - vxbm) / DZB(iz);
        double dvzdx = (vzb - vzbm) / DXB(ix);
        vyb = (A(vy, ix, iy, iz) + A(vy, ixm, iy, iz) + A(vy, ix, iym, iz) + A(vy, ixm, iym, iz)) * 0.25;
        vybm = (A(vy, ix, iy, izm) + A(vy, ixm, iy, izm) + A(vy, ix, iym, izm) + A(vy, ixm, iym, izm)) * 0.25;
        vzb = (A(vz, ix, iy, iz) + A(vz, ixm, iy, iz) + A(vz, ix, iy, izm) + A(vz, ixm, iy, izm)) * 0.25;
        vzbm = (A(vz, ix, iym, iz) + A(vz, ixm, iym, iz) + A(vz, ix, iym, izm) + A(vz, ixm, iym, izm)) * 0.25;
        double dvydz = (vyb - vybm) / DZB(iz);
        double dvzdy = (vzb - vzbm) / DYB(iy);
        double b1x = A(o->bx1, ix, iy, iz), b1y = A(o->by1, ix, iy, iz), b1z = A(o->bz1, ix, iy, iz);
        double w3 = b1x * dvxdx + b1y * dvxdy + b1z * dvxdz;
        double w4 = b1x * dvydx + b1y * dvydy + b1z * dvydz;
        double w5 = b1x * dvzdx + b1y * dvzdy + b1z * dvzdz;
        A(o->bx1, ix, iy, iz) = (b1x + w3 * o->dt2) / (1.0 + dv);
        A(o->by1, ix, iy, iz) = (b1y + w4 * o->dt2) / (1.0 + dv);
        A(o->bz1, ix, iy, iz) = (b1z + w5 * o->dt2) / (1.0 + dv);
      }
    }
  }
}

/* J x B of the CAUCHY predictor at node (ix, iy, iz) (lagran.F90:285-348):
 * currents and B averaged from the cell-centred B1 cv1 of the eight cells
 * around the node */
static void cauchy_jb(orc_state *o, int ix, int iy, int iz, double *jx, double *jy, double *jz, double *bxv,
                      double *byv, double *bzv) {
  int ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
  double *c1 = o->cv1, *X = o->bx1, *Y = o->by1, *Z = o->bz1;
  double cvx = A(c1, ix, iy, iz) + A(c1, ix, iyp, iz) + A(c1, ix, iy, izp) + A(c1, ix, iyp, izp);
  double cvxp = A(c1, ixp, iy, iz) + A(c1, ixp, iyp, iz) + A(c1, ixp, iy, izp) + A(c1, ixp, iyp, izp);
  double cvy = A(c1, ix, iy, iz) + A(c1, ixp, iy, iz) + A(c1, ix, iy, izp) + A(c1, ixp, iy, izp);
  double cvyp = A(c1, ix, iyp, iz) + A(c1, ixp, iyp, iz) + A(c1, ix, iyp, izp) + A(c1, ixp, iyp, izp);
  double cvz = A(c1, ix, iy, iz) + A(c1, ixp, iy, iz) + A(c1, ix, iyp, iz) + A(c1, ixp, iyp, iz);
  double cvzp = A(c1, ix, iy, izp) + A(c1, ixp, iy, izp) + A(c1, ix, iyp, izp) + A(c1, ixp, iyp, izp);
  double w1 = (A(Z, ix, iy, iz) + A(Z, ixp, iy, iz) + A(Z, ix, iy, izp) + A(Z, ixp, iy, izp)) / cvy;
  double w2 = (A(Z, ix, iyp, iz) + A(Z, ixp, iyp, iz) + A(Z, ix, iyp, izp) + A(Z, ixp, iyp, izp)) / cvyp;
  double x = (w2 - w1) / DYC(iy);
  w1 = (A(Y, ix, iy, iz) + A(Y, ixp, iy, iz) + A(Y, ix, iyp, iz) + A(Y, ixp, iyp, iz)) / cvz;
  w2 = (A(Y, ix, iy, izp) + A(Y, ixp, iy, izp) + A(Y, ix, iyp, izp) + A(Y, ixp, iyp, izp)) / cvzp;
  *jx = x - (w2 - w1) / DZC(iz);
  w1 = (A(Z, ix, iy, iz) + A(Z, ix, iyp, iz) + A(Z, ix, iy, izp) + A(Z, ix, iyp, izp)) / cvx;
  w2 = (A(Z, ixp, iy, iz) + A(Z, ixp, iyp, iz) + A(Z, ixp, iy, izp) + A(Z, ixp, iyp, izp)) / cvxp;
  double y = -(w2 - w1) / DXC(ix);
  w1 = (A(X, ix, iy, iz) + A(X, ixp, iy, iz) + A(X, ix, iyp, iz) + A(X, ixp, iyp, iz)) / cvz;
  w2 = (A(X, ix, iy, izp) + A(X, ixp, iy, izp) + A(X, ix, iyp, izp) + A(X, ixp, iyp, izp)) / cvzp;
  *jy = y + (w2 - w1) / DZC(iz);
  w1 = (A(Y, ix, iy, iz) + A(Y, ix, iyp, iz) + A(Y, ix, iy, izp) + A(Y, ix, iyp, izp)) / cvx;
  w2 = (A(Y, ixp, iy, iz) + A(Y, ixp, iyp, iz) + A(Y, ixp, iy, izp) + A(Y, ixp, iyp, izp)) / cvxp;
  double z = (w2 - w1) / DXC(ix);
  w1 = (A(X, ix, iy, iz) + A(X, ixp, iy, iz) + A(X, ix, iy, izp) + A(X, ixp, iy, izp)) / cvy;
  w2 = (A(X, ix, iyp, iz) + A(X, ixp, iyp, iz) + A(X, ix, iyp, izp) + A(X, ixp, iyp, izp)) / cvyp;
  *jz = z - (w2 - w1) / DYC(iy);
  *bxv = (A(X, ix, iy, iz) + A(X, ixp, iy, iz) + A(X, ix, iyp, iz) + A(X, ixp, iyp, iz) + A(X, ix, iy, izp) +
          A(X, ixp, iy, izp) + A(X, ix, iyp, izp) + A(X, ixp, iyp, izp)) / (cvx + cvxp);
  *byv = (A(Y, ix, iy, iz) + A(Y, ixp, iy, iz) + A(Y, ix, iyp, iz) + A(Y, ixp, iyp, iz) + A(Y, ix, iy, izp) +
          A(Y, ixp, iy, izp) + A(Y, ix, iyp, izp) + A(Y, ixp, iyp, izp)) / (cvx + cvxp);
  *bzv = (A(Z, ix, iy, iz) + A(Z, ixp, iy, iz) + A(Z, ix, iyp, iz) + A(Z, ixp, iyp, iz) + A(Z, ix, iy, izp) +
          A(Z, ixp, iy, izp) + A(Z, ix, iyp, izp) + A(Z, ixp, iyp, izp)) / (cvx + cvxp);
}

static void predictor_corrector_step(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double gm1 = o->p.gamma - 1.0;
  const int cauchy = o->p.cauchy;
  double *bx0 = NULL, *by0 = NULL, *bz0 = NULL;
  if (!cauchy) {
    bx0 = (double *)malloc(sizeof(double) * o->ntot);
    by0 = (double *)malloc(sizeof(double) * o->ntot);
    bz0 = (double *)malloc(sizeof(double) * o->ntot);
    copy_all(o, bx0, o->bx);
    copy_all(o, by0, o->by);
    copy_all(o, bz0, o->bz);
  }

  /* b_field_and_cv1_update (:945-1062): cv1 over the cell box, then the
   * B-only predictor remap with dt/2 (its cv1 supersedes this one) */
  if (cauchy) {
    cauchy_b_update(o);
    /* bx1 = bx1 * cv1 (:230-232) */
    for (int iz = -1; iz <= nz + 2; ++iz)
      for (int iy = -1; iy <= ny + 2; ++iy)
        for (int ix = -1; ix <= nx + 2; ++ix) {
          A(o->bx1, ix, iy, iz) = A(o->bx1, ix, iy, iz) * A(o->cv1, ix, iy, iz);
          A(o->by1, ix, iy, iz) = A(o->by1, ix, iy, iz) * A(o->cv1, ix, iy, iz);
          A(o->bz1, ix, iy, iz) = A(o->bz1, ix, iy, iz) * A(o->cv1, ix, iy, iz);
        }
  } else
  for (int iz = -1; iz <= nz + 2; ++iz) {
    int izm = iz - 1;
    for (int iy = -1; iy <= ny + 2; ++iy) {
      int iym = iy - 1;
      for (int ix = -1; ix <= nx + 2; ++ix) {
        int ixm = ix - 1;
        double vxb = (A(o->vx, ix, iy, iz) + A(o->vx, ix, iym, iz) + A(o->vx, ix, iy, izm) + A(o->vx, ix, iym, izm)) * 0.25;
        double vxbm = (A(o->vx, ixm, iy, iz) + A(o->vx, ixm, iym, iz) + A(o->vx, ixm, iy, izm) + A(o->vx, ixm, iym, izm)) * 0.25;
        double vyb = (A(o->vy, ix, iy, iz) + A(o->vy, ixm, iy, iz) + A(o->vy, ix, iy, izm) + A(o->vy, ixm, iy, izm)) * 0.25;
        double vybm = (A(o->vy, ix, iym, iz) + A(o->vy, ixm, iym, iz) + A(o->vy, ix, iym, izm) + A(o->vy, ixm, iym, izm)) * 0.25;
        double vzb = (A(o->vz, ix, iy, iz) + A(o->vz, ixm, iy, iz) + A(o->vz, ix, iym, iz) + A(o->vz, ixm, iym, iz)) * 0.25;
        double vzbm = (A(o->vz, ix, iy, izm) + A(o->vz, ixm, iy, izm) + A(o->vz, ix, iym, izm) + A(o->vz, ixm, iym, izm)) * 0.25;
        double dv = ((vxb - vxbm) / DXB(ix) + (vyb - vybm) / DYB(iy) + (vzb - vzbm) / DZB(iz)) * o->dt2;
        A(o->cv1, ix, iy, iz) = A(o->cv, ix, iy, iz) * (1.0 + dv);
      }
    }
  }
  if (!cauchy) {
    copy_all(o, o->vx1, o->vx);
    copy_all(o, o->vy1, o->vy);
    copy_all(o, o->vz1, o->vz);
    o->predictor_step = 1;
    o->dt = 0.5 * o->dt;
    orc_eulerian_remap(o, o->step);
    o->dt = 2.0 * o->dt;
    o->predictor_step = 0;
    /* (bx1 = bx1 * cv1, :230-232, is not read again without CAUCHY) */
  }

  for (int iz = 0; iz <= nz + 1; ++iz)
    for (int iy = 0; iy <= ny + 1; ++iy)
      for (int ix = 0; ix <= nx + 1; ++ix) {
        double r = A(o->rho, ix, iy, iz);
        double dv = A(o->cv1, ix, iy, iz) / A(o->cv, ix, iy, iz) - 1.0;
        double e1 = A(o->energy, ix, iy, iz) - A(o->pressure, ix, iy, iz) * dv / r;
        e1 = e1 + A(o->visc_heat, ix, iy, iz) * o->dt2 / r;
        A(o->pressure, ix, iy, iz) = (e1 - (1.0 - xi_at(o, IDX(ix, iy, iz))) * o->ip) * gm1 * r * A(o->cv, ix, iy, iz) / A(o->cv1, ix, iy, iz);
      }

  for (int iz = 0; iz <= nz; ++iz) {
    int izp = iz + 1;
    for (int iy = 0; iy <= ny; ++iy) {
      int iyp = iy + 1;
      for (int ix = 0; ix <= nx; ++ix) {
        int ixp = ix + 1;
        double *P = o->pressure;
        double pp = A(P, ix, iy, iz), ppx = A(P, ixp, iy, iz), ppy = A(P, ix, iyp, iz), ppz = A(P, ix, iy, izp);
        double ppxy = A(P, ixp, iyp, iz), ppxz = A(P, ixp, iy, izp), ppyz = A(P, ix, iyp, izp), ppxyz = A(P, ixp, iyp, izp);
        double w1 = (pp + ppy + ppz + ppyz) * 0.25;
        double w2 = (ppx + ppxy + ppxz + ppxyz) * 0.25;
        double fx = -(w2 - w1) / DXC(ix);
        w1 = (pp + ppx + ppz + ppxz) * 0.25;
        w2 = (ppy + ppxy + ppyz + ppxyz) * 0.25;
        double fy = -(w2 - w1) / DYC(iy);
        w1 = (pp + ppx + ppy + ppxy) * 0.25;
        w2 = (ppz + ppxz + ppyz + ppxyz) * 0.25;
        double fz = -(w2 - w1) / DZC(iz);
        double bxv, byv, bzv, jx, jy, jz;
        if (cauchy) {
          cauchy_jb(o, ix, iy, iz, &jx, &jy, &jz, &bxv, &byv, &bzv);
        } else {
          bxv = 0.25 * (A(o->bx, ix, iy, iz) + A(o->bx, ix, iyp, iz) + A(o->bx, ix, iy, izp) + A(o->bx, ix, iyp, izp));
          byv = 0.25 * (A(o->by, ix, iy, iz) + A(o->by, ixp, iy, iz) + A(o->by, ix, iy, izp) + A(o->by, ixp, iy, izp));
          bzv = 0.25 * (A(o->bz, ix, iy, iz) + A(o->bz, ix, iyp, iz) + A(o->bz, ixp, iy, iz) + A(o->bz, ixp, iyp, iz));
          vertex_j(o, ix, iy, iz, &jx, &jy, &jz);
        }
        const double gb = A(o->gamma_boris, ix, iy, iz);
        fx = fx + gb * (jy * bzv - jz * byv);
        fy = fy + gb * (jz * bxv - jx * bzv);
        fz = fz + gb * (jx * byv - jy * bxv);
        double rv = A(o->rho_vl, ix, iy, iz);
        fz = fz - rv * GRAV(iz);
        A(o->vx1, ix, iy, iz) = A(o->vx, ix, iy, iz) + o->dt2 * (A(o->fx_visc, ix, iy, iz) + fx) / rv;
        A(o->vy1, ix, iy, iz) = A(o->vy, ix, iy, iz) + o->dt2 * (A(o->fy_visc, ix, iy, iz) + fy) / rv;
        A(o->vz1, ix, iy, iz) = A(o->vz, ix, iy, iz) + o->dt2 * (A(o->fz_visc, ix, iy, iz) + fz) / rv;
      }
    }
  }
  if (!cauchy) {
    copy_all(o, o->bx, bx0);
    copy_all(o, o->by, by0);
    copy_all(o, o->bz, bz0);
    free(bx0); free(by0); free(bz0);
  }

  remap_v_bcs(o);
  shock_heating(o);
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) {
        A(o->vx, ix, iy, iz) = 2.0 * A(o->vx1, ix, iy, iz) - A(o->vx, ix, iy, iz);
        A(o->vy, ix, iy, iz) = 2.0 * A(o->vy1, ix, iy, iz) - A(o->vy, ix, iy, iz);
        A(o->vz, ix, iy, iz) = 2.0 * A(o->vz1, ix, iy, iz) - A(o->vz, ix, iy, iz);
      }
  velocity_bcs(o);
  if (o->any_open) open_bcs(o);

  for (int iz = 1; iz <= nz; ++iz) {
    int izm = iz - 1;
    for (int iy = 1; iy <= ny; ++iy) {
      int iym = iy - 1;
      for (int ix = 1; ix <= nx; ++ix) {
        int ixm = ix - 1;
        double *u = o->vx1, *v = o->vy1, *w = o->vz1;
        double vxb = (A(u, ix, iy, iz) + A(u, ix, iym, iz) + A(u, ix, iy, izm) + A(u, ix, iym, izm)) * 0.25;
        double vxbm = (A(u, ixm, iy, iz) + A(u, ixm, iym, iz) + A(u, ixm, iy, izm) + A(u, ixm, iym, izm)) * 0.25;
        double vyb = (A(v, ix, iy, iz) + A(v, ixm, iy, iz) + A(v, ix, iy, izm) + A(v, ixm, iy, izm)) * 0.25;
        double vybm = (A(v, ix, iym, iz) + A(v, ixm, iym, iz) + A(v, ix, iym, izm) + A(v, ixm, iym, izm)) * 0.25;
        double vzb = (A(w, ix, iy, iz) + A(w, ixm, iy, iz) + A(w, ix, iym, iz) + A(w, ixm, iym, iz)) * 0.25;
        double vzbm = (A(w, ix, iy, izm) + A(w, ixm, iy, izm) + A(w, ix, iym, izm) + A(w, ixm, iym, izm)) * 0.25;
        double dv = ((vxb - vxbm) / DXB(ix) + (vyb - vybm) / DYB(iy) + (vzb - vzbm) / DZB(iz)) * o->dt;
        A(o->cv1, ix, iy, iz) = A(o->cv, ix, iy, iz) * (1.0 + dv);
        double vh = A(o->visc_heat, ix, iy, iz);
        double r = A(o->rho, ix, iy, iz);
        A(o->energy, ix, iy, iz) = A(o->energy, ix, iy, iz) + (o->dt * vh - dv * A(o->pressure, ix, iy, iz)) / r;
        A(o->visc_dep, ix, iy, iz) = A(o->visc_dep, ix, iy, iz) + o->dt * vh;
        A(o->rho, ix, iy, iz) = r / (1.0 + dv);
        o->tvh = o->tvh + o->dt * vh * A(o->cv, ix, iy, iz);
      }
    }
  }
  if (o->p.cooling_term) {
    /* lagran.F90:464-476: a running average of the viscous heating is
     * removed again; energy = MAX(energy, 0) over the whole array */
    const double aav = o->p.alpha_av;
    for (int iz = 1; iz <= nz; ++iz)
      for (int iy = 1; iy <= ny; ++iy)
        for (int ix = 1; ix <= nx; ++ix) {
          A(o->cool_term_v, ix, iy, iz) = aav * o->dt * A(o->visc_heat, ix, iy, iz) / A(o->rho, ix, iy, iz) +
                                          (1.0 - aav) * A(o->cool_term_v, ix, iy, iz);
          A(o->energy, ix, iy, iz) = A(o->energy, ix, iy, iz) - A(o->cool_term_v, ix, iy, iz);
        }
    for (int iz = -1; iz <= nz + 2; ++iz)
      for (int iy = -1; iy <= ny + 2; ++iy)
        for (int ix = -1; ix <= nx + 2; ++ix) A(o->energy, ix, iy, iz) = dmax(A(o->energy, ix, iy, iz), 0.0);
  }
}

/* lagran.F90:50-200 lagrangian_step */
void orc_lagrangian_step(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double gm1 = o->p.gamma - 1.0;
  o->bx1 = cellbox(o); o->by1 = cellbox(o); o->bz1 = cellbox(o);
  o->alpha1 = box(o, 0, nx + 1, 0, ny + 2, 0, nz + 2);
  o->alpha2 = box(o, -1, nx + 1, 0, ny + 1, 0, nz + 2);
  o->alpha3 = box(o, -1, nx + 1, -1, ny + 1, 0, nz + 1);
  o->visc_heat = box(o, 0, nx + 1, 0, ny + 1, 0, nz + 1);
  o->pressure = cellbox(o);
  o->rho_vl = box(o, -1, nx + 1, -1, ny + 1, -1, nz + 1);
  o->cv_vl = box(o, -1, nx + 1, -1, ny + 1, -1, nz + 1);
  o->fx_visc = box(o, 0, nx, 0, ny, 0, nz);
  o->fy_visc = box(o, 0, nx, 0, ny, 0, nz);
  o->fz_visc = box(o, 0, nx, 0, ny, 0, nz);
  double *fxr = box(o, 0, nx, 0, ny, 0, nz), *fyr = box(o, 0, nx, 0, ny, 0, nz), *fzr = box(o, 0, nx, 0, ny, 0, nz);
  double *curlb = box(o, 0, nx, 0, ny, 0, nz);

  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) {
        A(o->bx1, ix, iy, iz) = (A(o->bx, ix, iy, iz) + A(o->bx, ix - 1, iy, iz)) * 0.5;
        A(o->by1, ix, iy, iz) = (A(o->by, ix, iy, iz) + A(o->by, ix, iy - 1, iz)) * 0.5;
        A(o->bz1, ix, iy, iz) = (A(o->bz, ix, iy, iz) + A(o->bz, ix, iy, iz - 1)) * 0.5;
        A(o->pressure, ix, iy, iz) = gm1 * A(o->rho, ix, iy, iz) * (A(o->energy, ix, iy, iz) - (1.0 - xi_at(o, IDX(ix, iy, iz))) * o->ip);
      }
  for (int iz = -1; iz <= nz + 1; ++iz)
    for (int iy = -1; iy <= ny + 1; ++iy)
      for (int ix = -1; ix <= nx + 1; ++ix) {
        int ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
        double *r = o->rho, *c = o->cv;
        double s = A(r, ix, iy, iz) * A(c, ix, iy, iz) + A(r, ixp, iy, iz) * A(c, ixp, iy, iz) + A(r, ix, iyp, iz) * A(c, ix, iyp, iz) +
                   A(r, ixp, iyp, iz) * A(c, ixp, iyp, iz) + A(r, ix, iy, izp) * A(c, ix, iy, izp) + A(r, ixp, iy, izp) * A(c, ixp, iy, izp) +
                   A(r, ix, iyp, izp) * A(c, ix, iyp, izp) + A(r, ixp, iyp, izp) * A(c, ixp, iyp, izp);
        double cvv = sum8(o, c, ix, iy, iz);
        A(o->rho_vl, ix, iy, iz) = s / cvv;
        A(o->cv_vl, ix, iy, iz) = 0.125 * cvv;
      }

  if (o->p.use_viscous_damping) viscous_damping(o);
  shock_viscosity(o);
  set_dt(o);
  o->dt2 = 0.5 * o->dt;
  o->last_substeps = 0;

  if (o->p.resistive_mhd) {
    double dt_sub = o->dtr;
    int substeps = (int)(o->dt / dt_sub) + 1;
    if (substeps > o->peak_substeps) o->peak_substeps = substeps;
    o->last_substeps = substeps;
    double actual_dt = o->dt;
    o->dt = o->dt / (double)substeps;
    for (int sc = 1; sc <= substeps; ++sc) {
      orc_eta_calc(o);
      if (o->pi_eos) neutral_fraction(o);
      if (o->p.cowling_resistivity) perpendicular_resistivity(o);
      resistive_effects(o, fxr, fyr, fzr, curlb);
    }
    /* bx1.. recomputed from the new B (:152-163); only the CAUCHY predictor
     * reads them again */
    for (int iz = -1; iz <= nz + 2; ++iz)
      for (int iy = -1; iy <= ny + 2; ++iy)
        for (int ix = -1; ix <= nx + 2; ++ix) {
          A(o->bx1, ix, iy, iz) = (A(o->bx, ix, iy, iz) + A(o->bx, ix - 1, iy, iz)) * 0.5;
          A(o->by1, ix, iy, iz) = (A(o->by, ix, iy, iz) + A(o->by, ix, iy - 1, iz)) * 0.5;
          A(o->bz1, ix, iy, iz) = (A(o->bz, ix, iy, iz) + A(o->bz, ix, iy, iz - 1)) * 0.5;
        }
    o->dt = actual_dt;
  }

  double *energy0 = (double *)malloc(sizeof(double) * o->ntot);
  copy_all(o, energy0, o->energy);
  if (o->p.conduction) {
    orc_conduct_heat(o);
    for (int iz = -1; iz <= nz + 2; ++iz)
      for (int iy = -1; iy <= ny + 2; ++iy)
        for (int ix = -1; ix <= nx + 2; ++ix) {
          double d = A(o->energy, ix, iy, iz) - A(energy0, ix, iy, iz);
          A(o->energy, ix, iy, iz) = A(energy0, ix, iy, iz) + d;
        }
  }
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) A(o->energy, ix, iy, iz) = dmax(A(o->energy, ix, iy, iz), 0.0);
  free(energy0);

  predictor_corrector_step(o);

  drop(&o->bx1); drop(&o->by1); drop(&o->bz1);
  drop(&o->alpha1); drop(&o->alpha2); drop(&o->alpha3);
  drop(&o->visc_heat); drop(&o->pressure); drop(&o->rho_vl); drop(&o->cv_vl);
  drop(&o->fx_visc); drop(&o->fy_visc); drop(&o->fz_visc);
  free(fxr); free(fyr); free(fzr); free(curlb);

  cell_bcs(o, o->energy);
  cell_bcs(o, o->rho);
  velocity_bcs(o);
}

/* diagnostics.F90:517-609 energy_account (one rank, do_sum false) */
void orc_energy_account(orc_state *o, double en[3]) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  double eb = 0.0, eke = 0.0, eint = 0.0;
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        double w1 = (SQ(A(o->bx, ix, iy, iz)) + SQ(A(o->bx, ix - 1, iy, iz))) * 0.5;
        double w2 = (SQ(A(o->by, ix, iy, iz)) + SQ(A(o->by, ix, iy - 1, iz))) * 0.5;
        double w3 = (SQ(A(o->bz, ix, iy, iz)) + SQ(A(o->bz, ix, iy, iz - 1))) * 0.5;
        w1 = (w1 + w2 + w3) * 0.5;
        eb = eb + w1 * A(o->cv, ix, iy, iz);
        eint = eint + A(o->energy, ix, iy, iz) * A(o->rho, ix, iy, iz) * A(o->cv, ix, iy, iz);
      }
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) {
        int ixp = ix + 1, iyp = iy + 1, izp = iz + 1;
        double *r = o->rho, *c = o->cv;
        double rho_v = A(r, ix, iy, iz) * A(c, ix, iy, iz) + A(r, ixp, iy, iz) * A(c, ixp, iy, iz) + A(r, ix, iyp, iz) * A(c, ix, iyp, iz) +
                       A(r, ixp, iyp, iz) * A(c, ixp, iyp, iz) + A(r, ix, iy, izp) * A(c, ix, iy, izp) + A(r, ixp, iy, izp) * A(c, ixp, iy, izp) +
                       A(r, ix, iyp, izp) * A(c, ix, iyp, izp) + A(r, ixp, iyp, izp) * A(c, ixp, iyp, izp);
        double cv_v = sum8(o, c, ix, iy, iz);
        rho_v = rho_v / cv_v;
        cv_v = cv_v * 0.125;
        double w1 = rho_v * cv_v * (SQ(A(o->vx, ix, iy, iz)) + SQ(A(o->vy, ix, iy, iz)) + SQ(A(o->vz, ix, iy, iz)));
        if (ix == 0 || ix == nx) w1 = w1 * 0.5;
        if (iy == 0 || iy == ny) w1 = w1 * 0.5;
        if (iz == 0 || iz == nz) w1 = w1 * 0.5;
        eke = eke + w1 * 0.5;
      }
  en[0] = eb;
  en[1] = eke;
  en[2] = eint;
}

double orc_divb(orc_state *o) {
  double m = 0.0;
  for (int iz = 1; iz <= o->nz; ++iz)
    for (int iy = 1; iy <= o->ny; ++iy)
      for (int ix = 1; ix <= o->nx; ++ix) {
        double d = (A(o->bx, ix, iy, iz) - A(o->bx, ix - 1, iy, iz)) / DXB(ix) + (A(o->by, ix, iy, iz) - A(o->by, ix, iy - 1, iz)) / DYB(iy) +
                   (A(o->bz, ix, iy, iz) - A(o->bz, ix, iy, iz - 1)) / DZB(iz);
        d = fabs(d) * min3(DXB(ix), DYB(iy), DZB(iz));
        if (d > m) m = d;
      }
  return m;
}

/* diagnostics.F90:613-629 energy_correction: the kinetic energy lost by
 * the last remap (where delta_ke < 0) added to the internal energy */
void orc_energy_correction(orc_state *o) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  for (int iz = -1; iz <= nz + 2; ++iz)
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) {
        double d = -A(o->delta_ke, ix, iy, iz);
        if (d < 0.0) d = 0.0;
        A(o->delta_ke, ix, iy, iz) = d / (A(o->rho, ix, iy, iz) * A(o->cv, ix, iy, iz));
      }
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) A(o->energy, ix, iy, iz) = A(o->energy, ix, iy, iz) + A(o->delta_ke, ix, iy, iz);
  cell_bcs(o, o->energy);
}

static void xfer(orc_state *o, double *a, double *f, int lx, int ly, int lz, int to_box);
/* delta_ke(-1:n+2)^3 as it stands (rke on), Fortran order */
int orc_get_delta_ke(orc_state *o, double *out) {
  if (!o->delta_ke) return -1;
  xfer(o, o->delta_ke, out, -1, -1, -1, 0);
  return 0;
}

/* lare3d.f90:82-87 loop body */
void orc_step(orc_state *o, int step, double *dt, double *time) {
  o->step = step;
  if (o->pi_eos) neutral_fraction(o);
  orc_lagrangian_step(o);
  orc_eulerian_remap(o, step);
  if (o->p.rke) orc_energy_correction(o);
  orc_eta_calc(o);
  if (dt) *dt = o->dt;
  if (time) *time = o->time;
}

/* ------------------------------------------------------------------------ */
/* potential_field (initial_conditions.f90:114-296)                          */
/* ------------------------------------------------------------------------ */
static void xfer(orc_state *o, double *a, double *f, int lx, int ly, int lz, int to_box);
/* phi_mpi (:238-294): exchanges on periodic axes, then the flux patch on the
 * lower z boundary, phi = 0 above the upper one, and zero-gradient x and y
 * walls, in that order. */
static void phi_bcs(orc_state *o, double *phi, const double *flux0) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  for (int ax = 0; ax < 3; ++ax) exchange(o, phi, ax, 0);
  if (!o->per[2]) {
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        A(phi, ix, iy, 0) = A(phi, ix, iy, 1) + flux0[(ix - 1) + (long)nx * (iy - 1)];
        A(phi, ix, iy, -1) = A(phi, ix, iy, 0);
      }
    for (int iy = -1; iy <= ny + 2; ++iy)
      for (int ix = -1; ix <= nx + 2; ++ix) {
        A(phi, ix, iy, nz + 1) = 0.0;
        A(phi, ix, iy, nz + 2) = 0.0;
      }
  }
  if (!o->per[0]) {
    copy_plane(o, phi, 0, 0, 1); copy_plane(o, phi, 0, -1, 1);
    copy_plane(o, phi, 0, nx + 1, nx); copy_plane(o, phi, 0, nx + 2, nx);
  }
  if (!o->per[1]) {
    copy_plane(o, phi, 1, 0, 1); copy_plane(o, phi, 1, -1, 1);
    copy_plane(o, phi, 1, ny + 1, ny); copy_plane(o, phi, 1, ny + 2, ny);
  }
}

int orc_potential_field(orc_state *o, const lare_potential *pf, int *iterations, double *residual, double *bz_max,
                        double *phi_out) {
  int nx = o->nx, ny = o->ny, nz = o->nz;
  /* bz(1:nx, 1:nz, 0) below is inside the array only for nz <= ny + 2 */
  if (nz > ny + 2 || o->per[2]) return -1;
  double *phi = cellbox(o);
  const double dx1 = pf->dx1, dy1 = pf->dy1, dz1 = pf->dz1, w = pf->omega;
  double error = 0.0;
  int loop, done = 0;
  for (loop = 1; loop <= pf->max_iterations; ++loop) {
    double errmax = 0.0;
    int z1 = 1;
    for (int redblack = 1; redblack <= 2; ++redblack) {
      for (int iz = 1; iz <= nz; ++iz) {
        int x1 = z1;
        for (int iy = 1; iy <= ny; ++iy) {
          for (int ix = x1; ix <= nx; ix += 2) {
            double r = (((A(phi, ix + 1, iy, iz) - 2.0 * A(phi, ix, iy, iz)) + A(phi, ix - 1, iy, iz)) * dx1 +
                        ((A(phi, ix, iy + 1, iz) - 2.0 * A(phi, ix, iy, iz)) + A(phi, ix, iy - 1, iz)) * dy1 +
                        ((A(phi, ix, iy, iz + 1) - 2.0 * A(phi, ix, iy, iz)) + A(phi, ix, iy, iz - 1)) * dz1) /
                       2.0 / (dx1 + dy1 + dz1);
            A(phi, ix, iy, iz) = A(phi, ix, iy, iz) + w * r;
            errmax = dmax(errmax, fabs(r));
          }
          x1 = 3 - x1;
        }
      }
      phi_bcs(o, phi, pf->flux0);
      z1 = 3 - z1;
    }
    error = errmax;
    if (error < pf->fractional_error) { done = 1; break; }
  }
  *iterations = done ? loop : pf->max_iterations;
  *residual = error;
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 0; ix <= nx; ++ix) A(o->bx, ix, iy, iz) = -(A(phi, ix + 1, iy, iz) - A(phi, ix, iy, iz)) / DXC(ix);
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 0; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) A(o->by, ix, iy, iz) = -(A(phi, ix, iy + 1, iz) - A(phi, ix, iy, iz)) / DYC(iy);
  for (int iz = 0; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) A(o->bz, ix, iy, iz) = -(A(phi, ix, iy, iz + 1) - A(phi, ix, iy, iz)) / DZC(iz);
  bfield_bcs(o);
  /* MAXVAL(bz(1:nx,1:nz,0)): the y range is 1:nz in the reference (:221) */
  double bm = -LARGEST;
  for (int iy = 1; iy <= nz; ++iy)
    for (int ix = 1; ix <= nx; ++ix) bm = dmax(bm, A(o->bz, ix, iy, 0));
  *bz_max = bm;
  double *b[3] = {o->bx, o->by, o->bz};
  int lo[3][3] = {{-2, -1, -1}, {-1, -2, -1}, {-1, -1, -2}};
  for (int c = 0; c < 3; ++c)
    for (int iz = lo[c][2]; iz <= nz + 2; ++iz)
      for (int iy = lo[c][1]; iy <= ny + 2; ++iy)
        for (int ix = lo[c][0]; ix <= nx + 2; ++ix) A(b[c], ix, iy, iz) = A(b[c], ix, iy, iz) / bm;
  bfield_bcs(o);
  /* phi(-1:nx+2, -1:ny+2, -1:nz+2), Fortran order, for the tests */
  if (phi_out) xfer(o, phi, phi_out, -1, -1, -1, 0);
  drop(&phi);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* state management                                                          */
/* ------------------------------------------------------------------------ */
orc_state *orc_create(const lare_grid *g, const lare_params *p) {
  orc_state *o = (orc_state *)calloc(1, sizeof(orc_state));
  o->nx = g->nx; o->ny = g->ny; o->nz = g->nz;
  o->ex = g->nx + 5; o->ey = g->ny + 5; o->ez = g->nz + 5;
  o->ntot = o->ex * o->ey * o->ez;
  int bc[6] = {g->xbc_min, g->xbc_max, g->ybc_min, g->ybc_max, g->zbc_min, g->zbc_max};
  memcpy(o->bc, bc, sizeof(bc));
  for (int a = 0; a < 3; ++a) o->per[a] = (bc[2 * a] == LARE_BC_PERIODIC && bc[2 * a + 1] == LARE_BC_PERIODIC);
  o->p = *p;
  int n[3] = {g->nx, g->ny, g->nz};
  const double *src[7] = {g->dxb, g->dxc, g->dyb, g->dyc, g->dzb, g->dzc, g->grav};
  for (int q = 0; q < 7; ++q) {
    int ax = q < 6 ? q / 2 : 2;
    int len = n[ax] + 5;
    o->m[q] = (double *)malloc(sizeof(double) * len);
    for (int i = 0; i < len; ++i) o->m[q][i] = NAN;
    int lo = -1, hi = (q < 6 && (q % 2) == 1) ? n[ax] + 1 : n[ax] + 2;
    for (int i = lo; i <= hi; ++i) o->m[q][i + 2] = src[q] ? src[q][i - lo] : 0.0;
  }
  o->rho = cellbox(o); o->energy = cellbox(o);
  o->vx = nodebox(o); o->vy = nodebox(o); o->vz = nodebox(o);
  o->vx1 = nodebox(o); o->vy1 = nodebox(o); o->vz1 = nodebox(o);
  o->bx = box(o, -2, o->nx + 2, -1, o->ny + 2, -1, o->nz + 2);
  o->by = box(o, -1, o->nx + 2, -2, o->ny + 2, -1, o->nz + 2);
  o->bz = box(o, -1, o->nx + 2, -1, o->ny + 2, -2, o->nz + 2);
  o->eta = cellbox(o); o->cv = cellbox(o); o->cv1 = cellbox(o);
  o->visc3 = nodebox(o); o->visc_dep = cellbox(o); o->ohmic_dep = cellbox(o);
  o->p_visc = cellbox(o); o->temperature = cellbox(o);
  o->larsen = box(o, 0, o->nx, 0, o->ny, 0, o->nz);
  o->gamma_boris = cellbox(o);
  for (long q = 0; q < o->ntot; ++q)
    if (o->gamma_boris[q] == 0.0) o->gamma_boris[q] = 1.0;
  o->cool_term_v = box(o, 1, o->nx, 1, o->ny, 1, o->nz);
  o->cool_term_b = box(o, 1, o->nx, 1, o->ny, 1, o->nz);
  if (o->p.rke) o->delta_ke = cellbox(o);
  o->pi_eos = o->p.eos_number == LARE_EOS_ION || o->p.eos_number == LARE_EOS_PI;
  for (int q = 0; q < 6; ++q) o->any_open |= bc[q] == LARE_BC_OPEN;
  o->ip = o->p.eos_number == LARE_EOS_ION ? o->p.ionise_pot : 0.0;
  if (o->pi_eos) o->xi_n = cellbox(o);
  if (o->p.cowling_resistivity) o->eta_perp = box(o, 0, o->nx, 0, o->ny, 0, o->nz);
  for (int iz = -1; iz <= o->nz + 2; ++iz)
    for (int iy = -1; iy <= o->ny + 2; ++iy)
      for (int ix = -1; ix <= o->nx + 2; ++ix) A(o->cv, ix, iy, iz) = DXB(ix) * DYB(iy) * DZB(iz);
  o->first_set_dt = 1;
  o->first_shock = 1;
  return o;
}

void orc_destroy(orc_state *o) {
  if (!o) return;
  double **all[] = {&o->rho, &o->energy, &o->vx, &o->vy, &o->vz, &o->vx1, &o->vy1, &o->vz1, &o->bx, &o->by, &o->bz, &o->eta,
                    &o->cv, &o->cv1, &o->visc3, &o->visc_dep, &o->ohmic_dep, &o->p_visc, &o->temperature, &o->larsen,
                    &o->gamma_boris, &o->cool_term_v, &o->cool_term_b, &o->delta_ke, &o->xi_n, &o->eta_perp};
  for (unsigned q = 0; q < sizeof(all) / sizeof(all[0]); ++q) drop(all[q]);
  for (int q = 0; q < 7; ++q) free(o->m[q]);
  free(o);
}

/* copy between a Fortran array with bounds lo..hi and the oracle box */
static void xfer(orc_state *o, double *a, double *f, int lx, int ly, int lz, int to_box) {
  int hx = o->nx + 2, hy = o->ny + 2, hz = o->nz + 2;
  long nxf = hx - lx + 1, nyf = hy - ly + 1;
  for (int k = lz; k <= hz; ++k)
    for (int j = ly; j <= hy; ++j)
      for (int i = lx; i <= hx; ++i) {
        long q = ((long)(k - lz) * nyf + (j - ly)) * nxf + (i - lx);
        if (to_box) A(a, i, j, k) = f[q];
        else f[q] = A(a, i, j, k);
      }
}

int orc_upload(orc_state *o, const lare_fields *f) {
  xfer(o, o->rho, f->rho, -1, -1, -1, 1);
  xfer(o, o->energy, f->energy, -1, -1, -1, 1);
  xfer(o, o->vx, f->vx, -2, -2, -2, 1);
  xfer(o, o->vy, f->vy, -2, -2, -2, 1);
  xfer(o, o->vz, f->vz, -2, -2, -2, 1);
  xfer(o, o->bx, f->bx, -2, -1, -1, 1);
  xfer(o, o->by, f->by, -1, -2, -1, 1);
  xfer(o, o->bz, f->bz, -1, -1, -2, 1);
  if (f->visc3) xfer(o, o->visc3, f->visc3, -2, -2, -2, 1);
  if (f->eta) xfer(o, o->eta, f->eta, -1, -1, -1, 1);
  if (f->visc_dep) xfer(o, o->visc_dep, f->visc_dep, -1, -1, -1, 1);
  if (f->ohmic_dep) xfer(o, o->ohmic_dep, f->ohmic_dep, -1, -1, -1, 1);
  return 0;
}

int orc_download(orc_state *o, lare_fields *f) {
  xfer(o, o->rho, f->rho, -1, -1, -1, 0);
  xfer(o, o->energy, f->energy, -1, -1, -1, 0);
  xfer(o, o->vx, f->vx, -2, -2, -2, 0);
  xfer(o, o->vy, f->vy, -2, -2, -2, 0);
  xfer(o, o->vz, f->vz, -2, -2, -2, 0);
  xfer(o, o->bx, f->bx, -2, -1, -1, 0);
  xfer(o, o->by, f->by, -1, -2, -1, 0);
  xfer(o, o->bz, f->bz, -1, -1, -2, 0);
  if (f->eta) xfer(o, o->eta, f->eta, -1, -1, -1, 0);
  if (f->visc_dep) xfer(o, o->visc_dep, f->visc_dep, -1, -1, -1, 0);
  if (f->ohmic_dep) xfer(o, o->ohmic_dep, f->ohmic_dep, -1, -1, -1, 0);
  return 0;
}

/* xi_n and eta_perp (0..n, zero elsewhere) in the cell layout */
int orc_get_neutral(orc_state *o, double *xi, double *ep) {
  if ((xi && !o->xi_n) || (ep && !o->eta_perp)) return LARE_E_STATE;
  if (xi) xfer(o, o->xi_n, xi, -1, -1, -1, 0);
  if (ep) {
    int hx = o->nx + 2, hy = o->ny + 2;
    long nxf = hx + 2, nyf = hy + 2;
    for (long q = 0; q < nxf * nyf * (o->nz + 4); ++q) ep[q] = 0.0;
    for (int k = 0; k <= o->nz; ++k)
      for (int j = 0; j <= o->ny; ++j)
        for (int i = 0; i <= o->nx; ++i) ep[((long)(k + 1) * nyf + (j + 1)) * nxf + (i + 1)] = A(o->eta_perp, i, j, k);
  }
  return 0;
}

int orc_get_fault(orc_state *o) { return o->fault; }

int orc_get_scalars(orc_state *o, lare_scalars *s) {
  s->dt = o->dt; s->dtr = o->dtr; s->time = o->time;
  s->total_visc_heating = o->tvh; s->total_ohmic_heating = o->toh;
  s->n_s_stages = o->n_s_stages; s->peak_substeps = o->peak_substeps;
  s->last_substeps = o->last_substeps; s->first_set_dt = o->first_set_dt;
  s->step = o->step;
  return 0;
}

int orc_set_scalars(orc_state *o, const lare_scalars *s) {
  o->dt = s->dt; o->dtr = s->dtr; o->time = s->time;
  o->tvh = s->total_visc_heating; o->toh = s->total_ohmic_heating;
  o->n_s_stages = s->n_s_stages; o->peak_substeps = s->peak_substeps;
  o->last_substeps = s->last_substeps; o->first_set_dt = s->first_set_dt;
  o->step = s->step;
  return 0;
}
