// remap.hpp -- Van Leer (1997) mass-coordinate remap with Evans-Hawley
// constrained transport of the face fluxes (remap.f90, xremap.f90,
// yremap.f90, zremap.f90).
//
// One template per kernel, instantiated for the sweep axis AX; the index
// arithmetic differs only in strides.  Loop ranges, clamps and the
// direction-specific CT update boxes are those of the reference (they are
// part of the numerical contract in predictor mode, where ghost faces carry
// half-updated values from one sweep to the next).
//
// Differences in form, not in arithmetic:
//  * dxc1 (xremap.f90:107-115) is recomputed from dxb1 where it is needed.
//  * the in-place averaging of the sweep velocity (xremap.f90:257-265) is
//    done where the momentum flux reads it; later sweeps only read that
//    component multiplied by a zero pass flag.
//  * vertex-averaged cv1/cv2/dm go to separate arrays instead of being
//    written over the cell values (xremap.f90:233-280).
//  * rho is double-buffered (rho1 = the old buffer) instead of copied.
#pragma once
#include "common.hpp"

#define SIXTH (1.0 / 6.0)

template <int AX>
__device__ __forceinline__ int sstride(const Geo &g) {
  return AX == 0 ? 1 : (AX == 1 ? g.sy : g.sz);
}
template <int AX>
__device__ __forceinline__ int coord(int i, int j, int k) {
  return AX == 0 ? i : (AX == 1 ? j : k);
}

// Van Leer limited slope Di (xremap.f90:390-396).
__device__ __forceinline__ double vl_di(double dfi, double dfu, double sign_v, double dci, double dcu, double dbu,
                                        double phi) {
  double Da = (2.0 - phi) * fabs(dfi) / dci + (1.0 + phi) * fabs(dfu) / dcu;
  Da = Da * SIXTH;
  const double ss = 0.5 * (copysign(1.0, dfi) + copysign(1.0, dfu));
  return sign_v * ss * dmin(dmin(fabs(Da) * dbu, fabs(dfi)), fabs(dfu));
}

// Lagrangian geometry of a sweep (xremap.f90:57-105 and analogues).
// Velocities are the time-centred ones (v in predictor mode).
template <int AX>
__global__ __launch_bounds__(BX * BY) void k_geom(Geo g, Box b, const double *__restrict__ ux, const double *__restrict__ uy,
                       const double *__restrict__ uz, double *__restrict__ cv1, double *__restrict__ cv2,
                       double *__restrict__ db1, double dt, double xpass, double ypass, double zpass, int write_cv1,
                       int write_cv2) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int X = 1, Y = g.sy, Z = g.sz;
  const double vxb = (ux[c] + ux[c - Y] + ux[c - Z] + ux[c - Y - Z]) * 0.25;
  const double vxbm = (ux[c - X] + ux[c - X - Y] + ux[c - X - Z] + ux[c - X - Y - Z]) * 0.25;
  const double vyb = (uy[c] + uy[c - X] + uy[c - Z] + uy[c - X - Z]) * 0.25;
  const double vybm = (uy[c - Y] + uy[c - X - Y] + uy[c - Y - Z] + uy[c - X - Y - Z]) * 0.25;
  const double vzb = (uz[c] + uz[c - X] + uz[c - Y] + uz[c - X - Y]) * 0.25;
  const double vzbm = (uz[c - Z] + uz[c - X - Z] + uz[c - Y - Z] + uz[c - X - Y - Z]) * 0.25;
  const double cv = g.cv(i, j, k);
  double dv1, dv2, d1;
  if (AX == 0) {
    const double tA = ypass * (vyb - vybm) / g.dyb[j];
    const double tB = zpass * (vzb - vzbm) / g.dzb[k];
    dv1 = (tA + tB + (vxb - vxbm) / g.dxb[i]) * dt;
    dv2 = (tA + tB) * dt;
    d1 = g.dxb[i] + (vxb - vxbm) * dt;
  } else if (AX == 1) {
    const double tA = xpass * (vxb - vxbm) / g.dxb[i];
    const double tB = zpass * (vzb - vzbm) / g.dzb[k];
    dv1 = (tA + tB + (vyb - vybm) / g.dyb[j]) * dt;
    dv2 = (tA + tB) * dt;
    d1 = g.dyb[j] + (vyb - vybm) * dt;
  } else {
    const double tA = xpass * (vxb - vxbm) / g.dxb[i];
    const double tB = ypass * (vyb - vybm) / g.dyb[j];
    dv1 = (tA + tB + (vzb - vzbm) / g.dzb[k]) * dt;
    dv2 = (tA + tB) * dt;
    d1 = g.dzb[k] + (vzb - vzbm) * dt;
  }
  if (write_cv1) cv1[c] = cv * (1.0 + dv1);
  if (write_cv2) cv2[c] = cv * (1.0 + dv2);
  db1[c] = d1;
}

// dxc1 at sweep coordinate s (index c), clamped as xremap.f90:110
template <int AX>
__device__ __forceinline__ double dc1(const Geo &g, const double *__restrict__ db1, int c, int s) {
  const int S = sstride<AX>(g);
  const int n = g.n(AX);
  return 0.5 * (db1[c] + db1[s + 1 <= n + 2 ? c + S : c]);
}

// CT flux of a transverse face field along AX (vx_by_flux etc.).
// VS: stride of the axis the velocity is averaged along (v(c) + v(c-VS));
// GS: stride of the axis d?b1 is averaged along (d(c) + d(c+GS)).
template <int AX>
__global__ __launch_bounds__(BX * BY) void k_ct_flux(Geo g, Box b, const double *__restrict__ fld, const double *__restrict__ vel,
                          const double *__restrict__ db1, double *__restrict__ F, int VS, int GS, double dt) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int S = sstride<AX>(g);
  const int n = g.n(AX);
  const int s = coord<AX>(i, j, k);
  const int sm = s - 1 > -1 ? s - 1 : -1;
  const int sp2 = s + 2 < n + 2 ? s + 2 : n + 2;
  const int cm = c + (sm - s) * S, cp = c + S, cp2 = c + (sp2 - s) * S;
  const double v_advect = (vel[c] + vel[c - VS]) * 0.5;
  const double db = (db1[c] + db1[c + GS]) * 0.5;
  const double dbp = (db1[cp] + db1[cp + GS]) * 0.5;
  const double dbp2 = (db1[cp2] + db1[cp2 + GS]) * 0.5;
  const double dbm = (db1[cm] + db1[cm + GS]) * 0.5;
  const double fm = fld[cm] / dbm;
  const double fi = fld[c] / db;
  const double fp = fld[cp] / dbp;
  const double fp2 = fld[cp2] / dbp2;
  const double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
  const double sign_v = copysign(1.0, v_advect);
  const double vad_p = (sign_v + 1.0) * 0.5;
  const double vad_m = 1.0 - vad_p;
  const double fu = fi * vad_p + fp * vad_m;
  const double dfu = dfm * vad_p + dfp * vad_m;
  const double dci = dc1<AX>(g, db1, c, s);
  const double dcu = dc1<AX>(g, db1, cm, sm) * vad_p + dc1<AX>(g, db1, cp, s + 1) * vad_m;
  const double dbu = db1[c] * vad_p + db1[cp] * vad_m;
  const double du = db * vad_p + dbp * vad_m;
  const double phi = fabs(v_advect) * dt / du;
  const double Di = vl_di(dfi, dfu, sign_v, dci, dcu, dbu, phi);
  F[c] = (fu + Di * (1.0 - phi)) * v_advect * dt;
}

// CT update pair: t = t - F + F(c - TS) on box bt; a = a + F - F(c - AS) on ba.
__global__ __launch_bounds__(BX * BY) void k_ct_update(Geo g, Box b, const double *__restrict__ F, double *__restrict__ t,
                            double *__restrict__ a, int TS, int AS, Box bt, Box ba) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  if (i >= bt.x0 && i <= bt.x1 && j >= bt.y0 && j <= bt.y1 && k >= bt.z0 && k <= bt.z1)
    t[c] = t[c] - F[c] + F[c - TS];
  if (i >= ba.x0 && i <= ba.x1 && j >= ba.y0 && j <= ba.y1 && k >= ba.z0 && k <= ba.z1)
    a[c] = a[c] + F[c] - F[c - AS];
}

// Face-averaged sweep velocity of the mass/energy fluxes (xremap.f90:492-493)
template <int AX>
__device__ __forceinline__ double face_velocity(const Geo &g, const double *__restrict__ vel, int c) {
  const int X = 1, Y = g.sy, Z = g.sz;
  if (AX == 0) return (vel[c] + vel[c - Y] + vel[c - Z] + vel[c - Y - Z]) * 0.25;
  if (AX == 1) return (vel[c] + vel[c - X] + vel[c - Z] + vel[c - X - Z]) * 0.25;
  return (vel[c] + vel[c - X] + vel[c - Y] + vel[c - X - Y]) * 0.25;
}
template <int AX>
__device__ __forceinline__ double face_area(const Geo &g, int i, int j, int k) {
  if (AX == 0) return g.dyb[j] * g.dzb[k];
  if (AX == 1) return g.dxb[i] * g.dzb[k];
  return g.dxb[i] * g.dyb[j];
}

// x_mass_flux (xremap.f90:474-533): sweep 0..n, transverse 0..n+1
template <int AX>
__global__ __launch_bounds__(BX * BY) void k_mass_flux(Geo g, Box b, const double *__restrict__ rho, const double *__restrict__ vel,
                            const double *__restrict__ db1, double *__restrict__ dm, double dt) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int S = sstride<AX>(g);
  const int s = coord<AX>(i, j, k);
  const double area = face_area<AX>(g, i, j, k);
  const double v_advect = face_velocity<AX>(g, vel, c);
  const double fm = rho[c - S], fi = rho[c], fp = rho[c + S], fp2 = rho[c + 2 * S];
  const double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
  const double sign_v = copysign(1.0, v_advect);
  const double vad_p = (sign_v + 1.0) * 0.5;
  const double vad_m = 1.0 - vad_p;
  const double fu = fi * vad_p + fp * vad_m;
  const double dfu = dfm * vad_p + dfp * vad_m;
  const double dci = dc1<AX>(g, db1, c, s);
  const double dcu = dc1<AX>(g, db1, c - S, s - 1) * vad_p + dc1<AX>(g, db1, c + S, s + 1) * vad_m;
  const double dbu = db1[c] * vad_p + db1[c + S] * vad_m;
  const double phi = fabs(v_advect) * dt / dbu;
  const double Di = vl_di(dfi, dfu, sign_v, dci, dcu, dbu, phi);
  dm[c] = (fu + Di * (1.0 - phi)) * v_advect * dt * area;
}

// x_energy_flux (xremap.f90:539-601): box 0..n
template <int AX>
__global__ __launch_bounds__(BX * BY) void k_energy_flux(Geo g, Box b, const double *__restrict__ energy, const double *__restrict__ rho1,
                              const double *__restrict__ vel, const double *__restrict__ db1,
                              const double *__restrict__ dm, double *__restrict__ F, double dt) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int S = sstride<AX>(g);
  const int s = coord<AX>(i, j, k);
  const double area = face_area<AX>(g, i, j, k);
  const double v_advect = face_velocity<AX>(g, vel, c);
  const double fm = energy[c - S], fi = energy[c], fp = energy[c + S], fp2 = energy[c + 2 * S];
  const double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
  const double sign_v = copysign(1.0, v_advect);
  const double vad_p = (sign_v + 1.0) * 0.5;
  const double vad_m = 1.0 - vad_p;
  const double fu = fi * vad_p + fp * vad_m;
  const double dfu = dfm * vad_p + dfp * vad_m;
  const double dci = dc1<AX>(g, db1, c, s);
  const double dcu = dc1<AX>(g, db1, c - S, s - 1) * vad_p + dc1<AX>(g, db1, c + S, s + 1) * vad_m;
  const double dbu = db1[c] * vad_p + db1[c + S] * vad_m;
  const double phi = fabs(v_advect) * dt / dbu;
  const double Di = vl_di(dfi, dfu, sign_v, dci, dcu, dbu, phi);
  const double rhou = rho1[c] * vad_p + rho1[c + S] * vad_m;
  const double d = dm[c];
  const double dmu = fabs(d) / area / dbu / rhou;
  F[c] = (fu + Di * (1.0 - dmu)) * d;
}

// density and energy updates of xremap.f90:174-196 on cells 1..n
template <int AX>
__global__ __launch_bounds__(BX * BY) void k_rho_energy_update(Geo g, Box b, const double *__restrict__ rho1, double *__restrict__ rho_new,
                                    double *__restrict__ energy, const double *__restrict__ cv1,
                                    const double *__restrict__ cv2, const double *__restrict__ dm,
                                    const double *__restrict__ F) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int S = sstride<AX>(g);
  const double r1 = rho1[c];
  const double rn = (r1 * cv1[c] + dm[c - S] - dm[c]) / cv2[c];
  rho_new[c] = rn;
  energy[c] = (energy[c] * cv1[c] * r1 + F[c - S] - F[c]) / (cv2[c] * rn);
}

// vertex quantities of xremap.f90:201-280: rho_v over sweep -1..n+1 and
// transverse 0..n; cv1v, cv2v over 0..n; dm_v over sweep -1..n.
template <int AX>
__global__ __launch_bounds__(BX * BY) void k_vertex(Geo g, Box b, const double *__restrict__ rho1, const double *__restrict__ cv1,
                         const double *__restrict__ cv2, const double *__restrict__ dm, double *__restrict__ rho_v,
                         double *__restrict__ cv1v, double *__restrict__ cv2v, double *__restrict__ dmv) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int X = 1, Y = g.sy, Z = g.sz;
  const int cc[8] = {c, c + X, c + Y, c + X + Y, c + Z, c + X + Z, c + Y + Z, c + X + Y + Z};
  double s = rho1[cc[0]] * cv1[cc[0]];
#pragma unroll
  for (int q = 1; q < 8; ++q) s = s + rho1[cc[q]] * cv1[cc[q]];
  double s1 = cv1[cc[0]];
#pragma unroll
  for (int q = 1; q < 8; ++q) s1 = s1 + cv1[cc[q]];
  rho_v[c] = s / s1;
  const int sc = coord<AX>(i, j, k);
  const int n = g.n(AX);
  if (sc >= 0 && sc <= n) {
    cv1v[c] = s1 * 0.125;
    double s2 = cv2[cc[0]];
#pragma unroll
    for (int q = 1; q < 8; ++q) s2 = s2 + cv2[cc[q]];
    cv2v[c] = s2 * 0.125;
  }
  if (sc <= n) {
    double sd = dm[cc[0]];
#pragma unroll
    for (int q = 1; q < 8; ++q) sd = sd + dm[cc[q]];
    dmv[c] = sd * 0.125;
  }
}

template <int AX>
__device__ __forceinline__ double mom_area(const Geo &g, int i, int j, int k) {
  if (AX == 0) return g.dyc[j] * g.dzc[k];
  if (AX == 1) return g.dxc[i] * g.dzc[k];
  return g.dxc[i] * g.dyc[j];
}

// x_momx/y/z_flux (xremap.f90:605-895, rke off): sweep -1..n, transverse
// 0..n; the three velocity components share the upwind switch and geometry.
template <int AX>
__global__ __launch_bounds__(BX * BY) void k_mom_flux(Geo g, Box b, const double *__restrict__ vx, const double *__restrict__ vy,
                           const double *__restrict__ vz, const double *__restrict__ vel1,
                           const double *__restrict__ db1, const double *__restrict__ rho_v,
                           const double *__restrict__ dmv, double *__restrict__ Fx, double *__restrict__ Fy,
                           double *__restrict__ Fz, double dt) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int S = sstride<AX>(g);
  const int s = coord<AX>(i, j, k);
  const double area = mom_area<AX>(g, i, j, k);
  const double v_advect = (vel1[c] + vel1[c + S]) * 0.5;  // averaged vx1 (xremap.f90:261)
  const double sign_v = copysign(1.0, v_advect);
  const double vad_p = (sign_v + 1.0) * 0.5;
  const double vad_m = 1.0 - vad_p;
  const double dci = db1[c + S];
  const double dcu = db1[c] * vad_p + db1[c + 2 * S] * vad_m;
  const double dbu = dc1<AX>(g, db1, c, s) * vad_p + dc1<AX>(g, db1, c + S, s + 1) * vad_m;
  const double phi = fabs(v_advect) * dt / dbu;
  const double rhou = rho_v[c] * vad_p + rho_v[c + S] * vad_m;
  const double d = dmv[c];
  const double dmu = fabs(d) / area / dbu / rhou;
  const double *V[3] = {vx, vy, vz};
  double *FF[3] = {Fx, Fy, Fz};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double *u = V[q];
    const double fm = u[c - S], fi = u[c], fp = u[c + S], fp2 = u[c + 2 * S];
    const double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
    const double fu = fi * vad_p + fp * vad_m;
    const double dfu = dfm * vad_p + dfp * vad_m;
    const double Di = vl_di(dfi, dfu, sign_v, dci, dcu, dbu, phi);
    const double f = fu + Di * (1.0 - dmu);
    FF[q][c] = f * d;
  }
}

// momentum updates (xremap.f90:282-329) on nodes 0..n
template <int AX>
__global__ __launch_bounds__(BX * BY) void k_mom_update(Geo g, Box b, double *__restrict__ vx, double *__restrict__ vy, double *__restrict__ vz,
                             const double *__restrict__ rho_v, const double *__restrict__ cv1v,
                             const double *__restrict__ cv2v, const double *__restrict__ dmv,
                             const double *__restrict__ Fx, const double *__restrict__ Fy,
                             const double *__restrict__ Fz) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int S = sstride<AX>(g);
  const double rv = rho_v[c], c1 = cv1v[c], c2 = cv2v[c];
  const double rv1 = (rv * c1 + dmv[c - S] - dmv[c]) / c2;
  vx[c] = (rv * vx[c] * c1 + Fx[c - S] - Fx[c]) / (c2 * rv1);
  vy[c] = (rv * vy[c] * c1 + Fy[c - S] - Fy[c]) / (c2 * rv1);
  vz[c] = (rv * vz[c] * c1 + Fz[c - S] - Fz[c]) / (c2 * rv1);
}

// B <-> face flux scaling of remap.f90:43-69 / :101-127 over the whole
// face arrays.  dir > 0 multiplies (out may differ from in: the predictor
// remap works on a copy); dir < 0 divides in the reference's order.
__global__ __launch_bounds__(BX * BY) void k_bscale(Geo g, Box b, const double *__restrict__ bxi, const double *__restrict__ byi,
                         const double *__restrict__ bzi, double *__restrict__ bxo, double *__restrict__ byo,
                         double *__restrict__ bzo, int dir) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const bool okx = j >= -1 && k >= -1;  // bx(-2:nx+2,-1:ny+2,-1:nz+2)
  const bool oky = i >= -1 && k >= -1;
  const bool okz = i >= -1 && j >= -1;
  if (dir > 0) {
    if (okx) bxo[c] = bxi[c] * g.dyb[j] * g.dzb[k];
    if (oky) byo[c] = byi[c] * g.dxb[i] * g.dzb[k];
    if (okz) bzo[c] = bzi[c] * g.dxb[i] * g.dyb[j];
  } else {
    if (okx) bxo[c] = bxi[c] / g.dyb[j] / g.dzb[k];
    if (oky) byo[c] = byi[c] / g.dxb[i] / g.dzb[k];
    if (okz) bzo[c] = bzi[c] / g.dxb[i] / g.dyb[j];
  }
}
