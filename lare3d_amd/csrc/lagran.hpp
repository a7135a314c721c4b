// lagran.hpp -- Lagrangian predictor-corrector step (lagran.F90).
//
// Kernel plan of lagrangian_step (lagran.F90:50-200), non-CAUCHY, boris
// off (gamma_boris = 1), cooling off:
//   k_vertex_prep     rho_v, cs_v at nodes -1..n+1 (:99-125, :551-581);
//                     cs and cv_v recomputed from the cell state / metrics
//   k_alpha           Caramana edge viscosities alpha1..3 (:583-678)
//   k_visc_dt         p_visc + visc_heat over cells 0..n+1 (:680-737), the
//                     start-of-step pressure, and the CFL minimum of set_dt
//                     (:1094-1143) fused in the same pass
//   k_eta / k_rkstep / k_bstep / k_ohmic   resistive sub-cycle (:1176-1635)
//   k_pred_pressure   predictor energy/pressure over 0..n+1 (:234-249)
//   k_force           f_visc (:745-789) + grad p + J x B (half-step B)
//                     + gravity -> v1 at nodes 0..n (:251-387)
//   k_shock_corr      shock_heating (:861-937) fused with the corrector
//                     (:415-462) over cells 1..n
//   k_vupdate         v = 2 v1 - v (:399-407)

#define NONE_ZERO 2.2250738585072014e-308
#define LARGEST 1.7976931348623157e+308

struct LagArgs {
  double gamma, visc1, visc2_norm, dt, dt2;
  double eta_bg, eta0, j_max;
  int resistive;
};

// 8-cell sum around node (i,j,k) in the reference order (lagran.F90:105-117)
#define SUM8(F, c, sy, sz)                                                                            \
  ((F(c)) + (F((c) + 1)) + (F((c) + (sy))) + (F((c) + 1 + (sy))) + (F((c) + (sz))) + (F((c) + 1 + (sz))) + \
   (F((c) + (sy) + (sz))) + (F((c) + 1 + (sy) + (sz))))

__device__ __forceinline__ double cv_at(const Geo &g, int i, int j, int k) { return g.cv(i, j, k); }

// sum of cv over the 8 cells of node (i,j,k), reference order
__device__ __forceinline__ double cv_sum8(const Geo &g, int i, int j, int k) {
  return g.cv(i, j, k) + g.cv(i + 1, j, k) + g.cv(i, j + 1, k) + g.cv(i + 1, j + 1, k) + g.cv(i, j, k + 1) +
         g.cv(i + 1, j, k + 1) + g.cv(i, j + 1, k + 1) + g.cv(i + 1, j + 1, k + 1);
}

// fast magnetosonic speed of shock_viscosity (:554-556) in cell c
__device__ __forceinline__ double cell_cs(const Geo &g, const real *__restrict__ rho,
                                          const real *__restrict__ energy, const double *__restrict__ bx,
                                          const double *__restrict__ by, const double *__restrict__ bz, int c,
                                          double gamma) {
  const double r = rho[c];
  const double rmin = dmax(r, NONE_ZERO);
  const double b1x = (bx[c] + bx[c - 1]) * 0.5;
  const double b1y = (by[c] + by[c - g.sy]) * 0.5;
  const double b1z = (bz[c] + bz[c - g.sz]) * 0.5;
  const double b2 = b1x * b1x + b1y * b1y + b1z * b1z;
  const double p = (gamma - 1.0) * r * (energy[c] - 0.0);
  return sqrt((gamma * p + 1.0 * b2) / rmin);
}

__global__ __launch_bounds__(BX * BY) void k_vertex_prep(Geo g, Box b, const real *__restrict__ rho, const real *__restrict__ energy,
                              const double *__restrict__ bx, const double *__restrict__ by,
                              const double *__restrict__ bz, real *__restrict__ rho_v, real *__restrict__ cs_v,
                              double gamma) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int sy = g.sy, sz = g.sz;
  const int cc[8] = {c, c + 1, c + sy, c + 1 + sy, c + sz, c + 1 + sz, c + sy + sz, c + 1 + sy + sz};
  const int di[8] = {0, 1, 0, 1, 0, 1, 0, 1}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1}, dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  double cvs[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) cvs[q] = g.cv(i + di[q], j + dj[q], k + dk[q]);
  double s = rho[cc[0]] * cvs[0];
#pragma unroll
  for (int q = 1; q < 8; ++q) s = s + rho[cc[q]] * cvs[q];
  double sc = cvs[0];
#pragma unroll
  for (int q = 1; q < 8; ++q) sc = sc + cvs[q];
  rho_v[c] = s / sc;
  const double cvv = 0.125 * sc;
  double scs = cell_cs(g, rho, energy, bx, by, bz, cc[0], gamma) * cvs[0];
#pragma unroll
  for (int q = 1; q < 8; ++q) scs = scs + cell_cs(g, rho, energy, bx, by, bz, cc[q], gamma) * cvs[q];
  cs_v[c] = 0.125 * scs / cvv;
}

// edge_viscosity (lagran.F90:795-855) between nodes n1 and n2 (linear idx)
__device__ __forceinline__ double edge_visc(const real *__restrict__ vx, const real *__restrict__ vy,
                                            const real *__restrict__ vz, const real *__restrict__ rho_v,
                                            const real *__restrict__ cs_v, int n1, int n2, double dvdots,
                                            double dx, const LagArgs &a) {
  dvdots = -fabs(dvdots);
  const double r1 = rho_v[n1], r2 = rho_v[n2];
  const double rho_edge = 2.0 * r1 * r2 / (r1 + r2);
  const double cs_edge = dmin(cs_v[n1], cs_v[n2]);
  const double dvx = vx[n1] - vx[n2];
  const double dvy = vy[n1] - vy[n2];
  const double dvz = vz[n1] - vz[n2];
  const double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
  const double dv = sqrt(dv2);
  const double psi = 0.0;
  if (dv * a.dt / dx < 1.0e-14) dvdots = 0.0;
  else dvdots = dvdots / dv;
  const double vn = a.visc2_norm;
  const double vc = a.visc1 * cs_edge;
  const double q_k_bar = rho_edge * (vn * dv + sqrt(vn * vn * dv2 + vc * vc));
  return q_k_bar * (1.0 - psi) * dvdots;
}

__global__ __launch_bounds__(BX * BY) void k_alpha(Geo g, Box b, const real *__restrict__ vx, const real *__restrict__ vy,
                        const real *__restrict__ vz, const real *__restrict__ rho_v,
                        const real *__restrict__ cs_v, real *__restrict__ al1, real *__restrict__ al2,
                        real *__restrict__ al3, LagArgs a) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i >= 0 && i <= nx + 1 && j >= 0 && j <= ny + 2 && k >= 0 && k <= nz + 2) {
    const int n1 = g.idx(i - 1, j - 1, k - 1), n2 = g.idx(i, j - 1, k - 1);
    al1[c] = edge_visc(vx, vy, vz, rho_v, cs_v, n1, n2, -(vx[n1] - vx[n2]), g.dxb[i], a);
  }
  if (i >= -1 && i <= nx + 1 && j >= 0 && j <= ny + 1 && k >= 0 && k <= nz + 2) {
    const int n1 = g.idx(i, j - 1, k - 1), n2 = g.idx(i, j, k - 1);
    al2[c] = edge_visc(vx, vy, vz, rho_v, cs_v, n1, n2, -(vy[n1] - vy[n2]), g.dyb[j], a);
  }
  if (i >= -1 && i <= nx + 1 && j >= -1 && j <= ny + 1 && k >= 0 && k <= nz + 1) {
    const int n1 = g.idx(i, j, k - 1), n2 = g.idx(i, j, k);
    al3[c] = edge_visc(vx, vy, vz, rho_v, cs_v, n1, n2, -(vz[n1] - vz[n2]), g.dzb[k], a);
  }
}

// node pairs of the 12 cell edges a1..a12 (lagran.F90:690-715), as linear
// offsets relative to node (i,j,k) given sy, sz
__device__ __forceinline__ void edge_nodes(int c, int sy, int sz, int e[12][2]) {
  const int m = c - 1 - sy - sz;  // (ixm, iym, izm)
  const int X = 1, Y = sy, Z = sz;
  // (ixm,iym,izm)-(ix,iym,izm)
  e[0][0] = m;             e[0][1] = m + X;
  e[1][0] = m + X;         e[1][1] = m + X + Y;          // (ix,iym,izm)-(ix,iy,izm)
  e[2][0] = m + X + Y;     e[2][1] = m + Y;              // (ix,iy,izm)-(ixm,iy,izm)
  e[3][0] = m + Y;         e[3][1] = m;                  // (ixm,iy,izm)-(ixm,iym,izm)
  e[4][0] = m + Z;         e[4][1] = m + X + Z;          // (ixm,iym,iz)-(ix,iym,iz)
  e[5][0] = m + X + Z;     e[5][1] = m + X + Y + Z;      // (ix,iym,iz)-(ix,iy,iz)
  e[6][0] = m + X + Y + Z; e[6][1] = m + Y + Z;          // (ix,iy,iz)-(ixm,iy,iz)
  e[7][0] = m + Y + Z;     e[7][1] = m + Z;              // (ixm,iy,iz)-(ixm,iym,iz)
  e[8][0] = m + X + Y;     e[8][1] = m + X + Y + Z;      // (ix,iy,izm)-(ix,iy,iz)
  e[9][0] = m + Y;         e[9][1] = m + Y + Z;          // (ixm,iy,izm)-(ixm,iy,iz)
  e[10][0] = m;            e[10][1] = m + Z;             // (ixm,iym,izm)-(ixm,iym,iz)
  e[11][0] = m + X;        e[11][1] = m + X + Z;         // (ix,iym,izm)-(ix,iym,iz)
}

// 12-edge heating sum of lagran.F90:722-736 (also :916-928), divided by cv
__device__ __forceinline__ double edge_heat(const Geo &g, int i, int j, int k, const real *__restrict__ al1,
                                            const real *__restrict__ al2, const real *__restrict__ al3,
                                            const double aa[12]) {
  const int c = g.idx(i, j, k);
  const int sy = g.sy, sz = g.sz;
  const double dxb = g.dxb[i], dyb = g.dyb[j], dzb = g.dzb[k];
  double h = -0.25 * dyb * dzb * al1[c] * aa[0];
  h = h - 0.25 * dxb * dzb * al2[c] * aa[1];
  h = h - 0.25 * dyb * dzb * al1[c + sy] * aa[2];
  h = h - 0.25 * dxb * dzb * al2[c - 1] * aa[3];
  h = h - 0.25 * dyb * dzb * al1[c + sz] * aa[4];
  h = h - 0.25 * dxb * dzb * al2[c + sz] * aa[5];
  h = h - 0.25 * dyb * dzb * al1[c + sy + sz] * aa[6];
  h = h - 0.25 * dxb * dzb * al2[c - 1 + sz] * aa[7];
  h = h - 0.25 * dxb * dyb * al3[c] * aa[8];
  h = h - 0.25 * dxb * dyb * al3[c - 1] * aa[9];
  h = h - 0.25 * dxb * dyb * al3[c - 1 - sy] * aa[10];
  h = h - 0.25 * dxb * dyb * al3[c - sy] * aa[11];
  return h / g.cv(i, j, k);
}

// p_visc + visc_heat over cells 0..n+1; pressure p0 of lagran.F90:93 stored
// for the predictor; set_dt partial minima over 0..n.
__global__ __launch_bounds__(BX * BY) void k_visc_dt(Geo g, const real *__restrict__ vx, const real *__restrict__ vy,
                          const real *__restrict__ vz, const real *__restrict__ rho,
                          const real *__restrict__ energy, const double *__restrict__ bx,
                          const double *__restrict__ by, const double *__restrict__ bz,
                          const real *__restrict__ eta, const real *__restrict__ al1,
                          const real *__restrict__ al2, const real *__restrict__ al3,
                          real *__restrict__ visc_heat, real *__restrict__ p0, double *__restrict__ partial,
                          LagArgs a) {
  const int i = (int)(blockIdx.x * BX + threadIdx.x);
  const int j = (int)(blockIdx.y * BY + threadIdx.y);
  const int k = (int)blockIdx.z;
  double v[2] = {LARGEST, LARGEST};
  if (i <= g.nx + 1 && j <= g.ny + 1 && k <= g.nz + 1) {
    const int c = g.idx(i, j, k);
    int e[12][2];
    edge_nodes(c, g.sy, g.sz, e);
    double aa[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      const double dx = vx[e[q][0]] - vx[e[q][1]];
      const double dy = vy[e[q][0]] - vy[e[q][1]];
      const double dz = vz[e[q][0]] - vz[e[q][1]];
      aa[q] = dx * dx + dy * dy + dz * dz;
    }
    double pv = 0.0;
    pv = dmax(pv, -al1[c] * sqrt(aa[0]));
    pv = dmax(pv, -al2[c] * sqrt(aa[1]));
    pv = dmax(pv, -al3[c] * sqrt(aa[8]));
    visc_heat[c] = edge_heat(g, i, j, k, al1, al2, al3, aa);
    const double r = rho[c];
    const double p = (a.gamma - 1.0) * r * (energy[c] - 0.0);
    p0[c] = p;
    if (i <= g.nx && j <= g.ny && k <= g.nz) {
      const double w1 = bx[c] * bx[c] + by[c] * by[c] + bz[c] * bz[c];
      const double rho0 = dmax(r, NONE_ZERO);
      const double cs2 = (a.gamma * p + 1.0 * w1) / rho0;
      const double c_visc2 = pv / rho0;
      const double dxb = g.dxb[i], dyb = g.dyb[j], dzb = g.dzb[k];
      const double length = dmin(dmin(dxb, dyb), dzb);
      const double dt1 = length / (sqrt(c_visc2) + sqrt(cs2 + c_visc2));
      const double dt2 = dmin(dmin(dxb / dmax(fabs(vx[c]), NONE_ZERO), dyb / dmax(fabs(vy[c]), NONE_ZERO)),
                              dzb / dmax(fabs(vz[c]), NONE_ZERO));
      v[0] = dmin(dt1, dt2);
      if (a.resistive) {
        const double dxlocal = 1.0 / (1.0 / (dxb * dxb) + 1.0 / (dyb * dyb) + 1.0 / (dzb * dzb));
        v[1] = 0.2 * dxlocal / dmax(eta[c], NONE_ZERO);
      }
    }
  }
  const int ops[2] = {0, 0};
  block_reduce_store<2>(v, ops, partial);
}

// viscous_damping (lagran.F90:483-524), two passes: q then v += dt q
__global__ __launch_bounds__(BX * BY) void k_damp_q(Geo g, Box b, const real *__restrict__ vx, const real *__restrict__ vy,
                         const real *__restrict__ vz, const real *__restrict__ visc3, real *__restrict__ qx,
                         real *__restrict__ qy, real *__restrict__ qz) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const real *V[3] = {vx, vy, vz};
  real *Q[3] = {qx, qy, qz};
  const double dxbp = g.dxb[i + 1], dxb = g.dxb[i], dxc = g.dxc[i];
  const double dybp = g.dyb[j + 1], dyb = g.dyb[j], dyc = g.dyc[j];
  const double dzbp = g.dzb[k + 1], dzb = g.dzb[k], dzc = g.dzc[k];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const real *u = V[q];
    const double u0 = u[c];
    Q[q][c] = visc3[c] * (((u[c + 1] - u0) / dxbp - (u0 - u[c - 1]) / dxb) / dxc +
                          ((u[c + g.sy] - u0) / dybp - (u0 - u[c - g.sy]) / dyb) / dyc +
                          ((u[c + g.sz] - u0) / dzbp - (u0 - u[c - g.sz]) / dzb) / dzc);
  }
}
__global__ __launch_bounds__(BX * BY) void k_damp_apply(Geo g, Box b, real *__restrict__ vx, real *__restrict__ vy, real *__restrict__ vz,
                             const real *__restrict__ qx, const real *__restrict__ qy,
                             const real *__restrict__ qz, double dt) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  vx[c] = vx[c] + dt * qx[c];
  vy[c] = vy[c] + dt * qy[c];
  vz[c] = vz[c] + dt * qz[c];
}

// Face B of one plane needed by the vertex current: bx at (i, j), (i, j+1)
// and by at (i, j), (i+1, j).
struct BPlane {
  double bx0, bxy, by0, byx;
};
__device__ __forceinline__ BPlane bplane(const double *__restrict__ bx, const double *__restrict__ by, int c, int Y) {
  BPlane p = {bx[c], bx[c + Y], by[c], by[c + 1]};
  return p;
}

// vertex current of lagran.F90:1479-1494 (also :358-373, :1424-1439) from
// bx, by of planes k (lo) and k+1 (hi) and bz at (i,j), (i+1,j), (i,j+1),
// (i+1,j+1) of plane k
__device__ __forceinline__ void vertex_current_vals(double dxc, double dyc, double dzc, const BPlane &lo,
                                                    const BPlane &hi, double bz0, double bzx, double bzy, double bzxy,
                                                    double &jx, double &jy, double &jz) {
  const double jx1 = (bzy - bz0) / dyc - (hi.by0 - lo.by0) / dzc;
  const double jx2 = (bzxy - bzx) / dyc - (hi.byx - lo.byx) / dzc;
  const double jy1 = (hi.bx0 - lo.bx0) / dzc - (bzx - bz0) / dxc;
  const double jy2 = (hi.bxy - lo.bxy) / dzc - (bzxy - bzy) / dxc;
  const double jz1 = (lo.byx - lo.by0) / dxc - (lo.bxy - lo.bx0) / dyc;
  const double jz2 = (hi.byx - hi.by0) / dxc - (hi.bxy - hi.bx0) / dyc;
  jx = (jx1 + jx2) * 0.5;
  jy = (jy1 + jy2) * 0.5;
  jz = (jz1 + jz2) * 0.5;
}

// the same vertex current, read from memory at node (i, j, k)
__device__ __forceinline__ void vertex_current(const Geo &g, const double *__restrict__ bx,
                                               const double *__restrict__ by, const double *__restrict__ bz, int i,
                                               int j, int k, double &jx, double &jy, double &jz) {
  const int c = g.idx(i, j, k);
  const int Y = g.sy;
  const BPlane lo = bplane(bx, by, c, Y), hi = bplane(bx, by, c + g.sz, Y);
  vertex_current_vals(g.dxc[i], g.dyc[j], g.dzc[k], lo, hi, bz[c], bz[c + 1], bz[c + Y], bz[c + 1 + Y], jx, jy, jz);
}

// eta_calc (lagran.F90:1176-1230) over -1..n+1, marching in z with the face
// B of plane k+1 kept for the next plane
__global__ __launch_bounds__(BX * BY) void k_eta(Geo g, Box b, const double *__restrict__ bx, const double *__restrict__ by,
                      const double *__restrict__ bz, real *__restrict__ eta, double eta_bg, double eta0,
                      double j_max) {
  ZCOLUMN(b, i, j, k0, k1);
  const int Y = g.sy, Z = g.sz;
  BPlane lo = bplane(bx, by, g.idx(i, j, k0), Y);
  for (int k = k0; k <= k1; ++k) {
    const int c = g.idx(i, j, k);
    const BPlane hi = bplane(bx, by, c + Z, Y);
    double jx, jy, jz;
    vertex_current_vals(g.dxc[i], g.dyc[j], g.dzc[k], lo, hi, bz[c], bz[c + 1], bz[c + Y], bz[c + 1 + Y], jx, jy, jz);
    eta[c] = sqrt(jx * jx + jy * jy + jz * jz) > j_max ? eta_bg + eta0 : eta_bg;
    lo = hi;
  }
}

__global__ __launch_bounds__(BX * BY) void k_fill(Geo g, Box b, real *__restrict__ a, double v) {
  BOX_INDEX(b, i, j, k);
  a[g.idx(i, j, k)] = v;
}

// rkstep (lagran.F90:1472-1503) over nodes 0..n, with the
// total_ohmic_heating increments of :1309-1328 reduced in the same pass;
// marches in z (box nodes 0..n).
__global__ __launch_bounds__(BX * BY) void k_rkstep(Geo g, Box b, const double *__restrict__ bx, const double *__restrict__ by,
                         const double *__restrict__ bz, const real *__restrict__ eta, double *__restrict__ fxr,
                         double *__restrict__ fyr, double *__restrict__ fzr, real *__restrict__ curlb,
                         double *__restrict__ partial, double dt) {
  const int i = b.x0 + (int)(blockIdx.x * BX + threadIdx.x);
  const int j = b.y0 + (int)(blockIdx.y * BY + threadIdx.y);
  const int k0 = b.z0 + (int)blockIdx.z * ZCHUNK;
  const int k1 = k0 + ZCHUNK - 1 < b.z1 ? k0 + ZCHUNK - 1 : b.z1;
  double v[1] = {0.0};
  if (i <= b.x1 && j <= b.y1) {
    const int Y = g.sy, Z = g.sz;
    BPlane lo = bplane(bx, by, g.idx(i, j, k0), Y);
    for (int k = k0; k <= k1; ++k) {
      const int c = g.idx(i, j, k);
      const BPlane hi = bplane(bx, by, c + Z, Y);
      double jx, jy, jz;
      vertex_current_vals(g.dxc[i], g.dyc[j], g.dzc[k], lo, hi, bz[c], bz[c + 1], bz[c + Y], bz[c + 1 + Y], jx, jy,
                          jz);
      lo = hi;
      const double e = eta[c];
      fxr[c] = -jx * e * g.dxc[i] * 0.5;
      fyr[c] = -jy * e * g.dyc[j] * 0.5;
      fzr[c] = -jz * e * g.dzc[k] * 0.5;
      const double cb = e * (jx * jx + jy * jy + jz * jz);
      curlb[c] = cb;
      double w1 = dt * g.dxc[i] * g.dyc[j] * g.dzc[k] * cb;
      if (i == 0 || i == g.nx) w1 = w1 * 0.5;
      if (j == 0 || j == g.ny) w1 = w1 * 0.5;
      if (k == 0 || k == g.nz) w1 = w1 * 0.5;
      v[0] += w1;
    }
  }
  const int ops[1] = {1};
  block_reduce_store<1>(v, ops, partial);
}

// bstep (lagran.F90:1582-1631); B updated in place (bx1 = bx at the same
// point, :1259-1261).  Marches in z with the EMFs of plane k-1 in registers.
struct EPlane {
  double kx0, kxm, ky0, kym, kz0, kzm, kzx;  // kx at (i,j),(i-1,j); ky at (i,j),(i,j-1); kz at (i,j),(i,j-1),(i-1,j)
};
__device__ __forceinline__ EPlane eplane(const double *__restrict__ kx, const double *__restrict__ ky,
                                         const double *__restrict__ kz, int c, int Y) {
  EPlane p = {kx[c], kx[c - 1], ky[c], ky[c - Y], kz[c], kz[c - Y], kz[c - 1]};
  return p;
}
__global__ __launch_bounds__(BX * BY) void k_bstep(Geo g, Box b, double *__restrict__ bx, double *__restrict__ by, double *__restrict__ bz,
                        const double *__restrict__ kx, const double *__restrict__ ky, const double *__restrict__ kz,
                        double dt) {
  ZCOLUMN(b, i, j, k0, k1);
  const int Y = g.sy, Z = g.sz;
  EPlane m = eplane(kx, ky, kz, g.idx(i, j, k0) - Z, Y);  // plane k-1
  for (int k = k0; k <= k1; ++k) {
    const int c = g.idx(i, j, k);
    const EPlane p = eplane(kx, ky, kz, c, Y);
    if (j >= 1 && k >= 1) {
      const double area = g.dyb[j] * g.dzb[k];
      bx[c] = bx[c] + (p.kz0 - p.kzm + m.kz0 - m.kzm - p.ky0 + m.ky0 - p.kym + m.kym) * dt / area;
    }
    if (i >= 1 && k >= 1) {
      const double area = g.dxb[i] * g.dzb[k];
      by[c] = by[c] + (p.kx0 - m.kx0 + p.kxm - m.kxm - p.kz0 + p.kzx - m.kz0 + m.kzx) * dt / area;
    }
    if (i >= 1 && j >= 1) {
      const double area = g.dxb[i] * g.dyb[j];
      bz[c] = bz[c] + (p.ky0 - ky[c - 1] + p.kym - ky[c - 1 - Y] - p.kx0 + kx[c - Y] - p.kxm + kx[c - 1 - Y]) *
                          dt / area;
    }
    m = p;
  }
}

// Ohmic heating (lagran.F90:1270-1288) on cells 1..n, then
// energy = MAX(energy, 0) over the whole cell box (:1305); marches in z with
// the vertex heating of plane k-1 in registers.
__global__ __launch_bounds__(BX * BY) void k_ohmic(Geo g, Box b, real *__restrict__ energy, real *__restrict__ ohmic_dep,
                        const real *__restrict__ rho, const real *__restrict__ curlb, double dt) {
  ZCOLUMN(b, i, j, k0, k1);
  const int Y = g.sy;
  const bool inxy = i >= 1 && i <= g.nx && j >= 1 && j <= g.ny;
  Quad m = quad_back(curlb, g.idx(i, j, k0 - 1), Y);
  for (int k = k0; k <= k1; ++k) {
    const int c = g.idx(i, j, k);
    const Quad p = quad_back(curlb, c, Y);
    double e = energy[c];
    if (inxy && k >= 1 && k <= g.nz) {
      // curlb(c) + (c-X) + (c-Y) + (c-X-Y) + (c-Z) + (c-X-Z) + (c-Y-Z) + (c-X-Y-Z)
      const double s = (double)p.p0 + p.px + p.py + p.pxy + m.p0 + m.px + m.py + m.pxy;
      const double lh = s * dt / (8.0 * rho[c]);
      ohmic_dep[c] = ohmic_dep[c] + lh;
      e = e + lh;
    }
    energy[c] = max0(e);
    m = p;
  }
}

__global__ __launch_bounds__(BX * BY) void k_energy_max(Geo g, Box b, real *__restrict__ energy) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  energy[c] = max0(energy[c]);
}

// predictor energy and pressure (lagran.F90:234-249) over cells 0..n+1;
// p0 is the start-of-step pressure (:93), cv1 the predictor-remap volume
__global__ __launch_bounds__(BX * BY) void k_pred_pressure(Geo g, Box b, const real *__restrict__ energy, const real *__restrict__ rho,
                                const real *__restrict__ cv1, const real *__restrict__ visc_heat,
                                const real *__restrict__ p0, real *__restrict__ pp, double gamma, double dt2) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const double cv = g.cv(i, j, k);
  const double r = rho[c];
  const double dv = cv1[c] / cv - 1.0;
  double e1 = energy[c] - p0[c] * dv / r;
  e1 = e1 + visc_heat[c] * dt2 / r;
  pp[c] = (e1 - 0.0) * (gamma - 1.0) * r * cv / cv1[c];
}

// predictor momentum (lagran.F90:251-387) with fx/fy/fz_visc (:745-789)
// recomputed here from alpha and v; B is the half-step (predictor-remapped)
// field.
__global__ __launch_bounds__(BX * BY) void k_force(Geo g, Box b, const real *__restrict__ vx, const real *__restrict__ vy,
                        const real *__restrict__ vz, const real *__restrict__ pp,
                        const double *__restrict__ bx, const double *__restrict__ by, const double *__restrict__ bz,
                        const real *__restrict__ rho_v, const real *__restrict__ al1,
                        const real *__restrict__ al2, const real *__restrict__ al3, real *__restrict__ vx1,
                        real *__restrict__ vy1, real *__restrict__ vz1, double dt2) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const int X = 1, Y = g.sy, Z = g.sz;
  const double dxc = g.dxc[i], dyc = g.dyc[j], dzc = g.dzc[k];
  // viscous forces
  const double a1 = al1[c + Y + Z] * dyc * dzc;
  const double a2 = al1[c + X + Y + Z] * dyc * dzc;
  const double a3 = al2[c + Z] * dxc * dzc;
  const double a4 = al2[c + Y + Z] * dxc * dzc;
  const double a5 = al3[c] * dxc * dyc;
  const double a6 = al3[c + Z] * dxc * dyc;
  const double cvv = 0.125 * cv_sum8(g, i, j, k);
  const real *V[3] = {vx, vy, vz};
  double fv[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const real *u = V[q];
    const double u0 = u[c];
    fv[q] = (a1 * (u0 - u[c - X]) + a2 * (u0 - u[c + X]) + a3 * (u0 - u[c - Y]) + a4 * (u0 - u[c + Y]) +
             a5 * (u0 - u[c - Z]) + a6 * (u0 - u[c + Z])) /
            cvv;
  }
  const double p = pp[c], ppx = pp[c + X], ppy = pp[c + Y], ppz = pp[c + Z];
  const double ppxy = pp[c + X + Y], ppxz = pp[c + X + Z], ppyz = pp[c + Y + Z], ppxyz = pp[c + X + Y + Z];
  double w1 = (p + ppy + ppz + ppyz) * 0.25;
  double w2 = (ppx + ppxy + ppxz + ppxyz) * 0.25;
  double fx = -(w2 - w1) / dxc;
  w1 = (p + ppx + ppz + ppxz) * 0.25;
  w2 = (ppy + ppxy + ppyz + ppxyz) * 0.25;
  double fy = -(w2 - w1) / dyc;
  w1 = (p + ppx + ppy + ppxy) * 0.25;
  w2 = (ppz + ppxz + ppyz + ppxyz) * 0.25;
  double fz = -(w2 - w1) / dzc;
  const double bxv = 0.25 * (bx[c] + bx[c + Y] + bx[c + Z] + bx[c + Y + Z]);
  const double byv = 0.25 * (by[c] + by[c + X] + by[c + Z] + by[c + X + Z]);
  const double bzv = 0.25 * (bz[c] + bz[c + Y] + bz[c + X] + bz[c + X + Y]);
  double jx, jy, jz;
  vertex_current(g, bx, by, bz, i, j, k, jx, jy, jz);
  fx = fx + 1.0 * (jy * bzv - jz * byv);
  fy = fy + 1.0 * (jz * bxv - jx * bzv);
  fz = fz + 1.0 * (jx * byv - jy * bxv);
  const double rv = rho_v[c];
  fz = fz - rv * g.grav[k];
  vx1[c] = vx[c] + dt2 * (fv[0] + fx) / rv;
  vy1[c] = vy[c] + dt2 * (fv[1] + fy) / rv;
  vz1[c] = vz[c] + dt2 * (fv[2] + fz) / rv;
}

// shock_heating (lagran.F90:861-937) + corrector (:415-462) over cells
// 1..n; total_visc_heating partial sums.
__global__ __launch_bounds__(BX * BY) void k_shock_corr(Geo g, const real *__restrict__ vx, const real *__restrict__ vy,
                             const real *__restrict__ vz, const real *__restrict__ vx1,
                             const real *__restrict__ vy1, const real *__restrict__ vz1,
                             const real *__restrict__ al1, const real *__restrict__ al2,
                             const real *__restrict__ al3, const real *__restrict__ pp,
                             real *__restrict__ energy, real *__restrict__ rho, real *__restrict__ visc_dep,
                             double *__restrict__ partial, double dt) {
  const int i = 1 + (int)(blockIdx.x * BX + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * BY + threadIdx.y);
  const int k = 1 + (int)blockIdx.z;
  double v[1] = {0.0};
  if (i <= g.nx && j <= g.ny && k <= g.nz) {
    const int c = g.idx(i, j, k);
    int e[12][2];
    edge_nodes(c, g.sy, g.sz, e);
    double aa[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      const int p = e[q][0], r = e[q][1];
      aa[q] = (vx[p] - vx[r]) * (vx1[p] - vx1[r]) + (vy[p] - vy[r]) * (vy1[p] - vy1[r]) +
              (vz[p] - vz[r]) * (vz1[p] - vz1[r]);
    }
    const double vh = max0(edge_heat(g, i, j, k, al1, al2, al3, aa));
    const int X = 1, Y = g.sy, Z = g.sz;
    const double vxb = (vx1[c] + vx1[c - Y] + vx1[c - Z] + vx1[c - Y - Z]) * 0.25;
    const double vxbm = (vx1[c - X] + vx1[c - X - Y] + vx1[c - X - Z] + vx1[c - X - Y - Z]) * 0.25;
    const double vyb = (vy1[c] + vy1[c - X] + vy1[c - Z] + vy1[c - X - Z]) * 0.25;
    const double vybm = (vy1[c - Y] + vy1[c - X - Y] + vy1[c - Y - Z] + vy1[c - X - Y - Z]) * 0.25;
    const double vzb = (vz1[c] + vz1[c - X] + vz1[c - Y] + vz1[c - X - Y]) * 0.25;
    const double vzbm = (vz1[c - Z] + vz1[c - X - Z] + vz1[c - Y - Z] + vz1[c - X - Y - Z]) * 0.25;
    const double dv = ((vxb - vxbm) / g.dxb[i] + (vyb - vybm) / g.dyb[j] + (vzb - vzbm) / g.dzb[k]) * dt;
    const double r = rho[c];
    energy[c] = energy[c] + (dt * vh - dv * pp[c]) / r;
    visc_dep[c] = visc_dep[c] + dt * vh;
    rho[c] = r / (1.0 + dv);
    v[0] = dt * vh * g.cv(i, j, k);
  }
  const int ops[1] = {1};
  block_reduce_store<1>(v, ops, partial);
}

__global__ __launch_bounds__(BX * BY) void k_vupdate(Geo g, Box b, real *__restrict__ vx, real *__restrict__ vy, real *__restrict__ vz,
                          const real *__restrict__ vx1, const real *__restrict__ vy1,
                          const real *__restrict__ vz1) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  vx[c] = 2.0 * vx1[c] - vx[c];
  vy[c] = 2.0 * vy1[c] - vy[c];
  vz[c] = 2.0 * vz1[c] - vz[c];
}

// energy_account (diagnostics.F90:517-609) + max |div B| * min(dx):
// partial[0..3] = (en_b, en_ke, en_int, -max divb) per block over box 0..n
__global__ __launch_bounds__(BX * BY) void k_energy(Geo g, const real *__restrict__ rho, const real *__restrict__ energy,
                         const real *__restrict__ vx, const real *__restrict__ vy, const real *__restrict__ vz,
                         const double *__restrict__ bx, const double *__restrict__ by,
                         const double *__restrict__ bz, double *__restrict__ partial) {
  const int i = (int)(blockIdx.x * BX + threadIdx.x);
  const int j = (int)(blockIdx.y * BY + threadIdx.y);
  const int k = (int)blockIdx.z;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (i <= g.nx && j <= g.ny && k <= g.nz) {
    const int c = g.idx(i, j, k);
    const int X = 1, Y = g.sy, Z = g.sz;
    if (i >= 1 && j >= 1 && k >= 1) {
      double w1 = (bx[c] * bx[c] + bx[c - X] * bx[c - X]) * 0.5;
      const double w2 = (by[c] * by[c] + by[c - Y] * by[c - Y]) * 0.5;
      const double w3 = (bz[c] * bz[c] + bz[c - Z] * bz[c - Z]) * 0.5;
      w1 = (w1 + w2 + w3) * 0.5;
      const double cv = g.cv(i, j, k);
      v[0] = w1 * cv;
      v[2] = energy[c] * rho[c] * cv;
      const double dx = g.dxb[i], dy = g.dyb[j], dz = g.dzb[k];
      const double d = (bx[c] - bx[c - X]) / dx + (by[c] - by[c - Y]) / dy + (bz[c] - bz[c - Z]) / dz;
      v[3] = -(fabs(d) * dmin(dmin(dx, dy), dz));
    }
    double cvs[8];
    const int di[8] = {0, 1, 0, 1, 0, 1, 0, 1}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1}, dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
    const int cc[8] = {c, c + X, c + Y, c + X + Y, c + Z, c + X + Z, c + Y + Z, c + X + Y + Z};
#pragma unroll
    for (int q = 0; q < 8; ++q) cvs[q] = g.cv(i + di[q], j + dj[q], k + dk[q]);
    double rv = rho[cc[0]] * cvs[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) rv = rv + rho[cc[q]] * cvs[q];
    double cvv = cvs[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) cvv = cvv + cvs[q];
    rv = rv / cvv;
    cvv = cvv * 0.125;
    double w1 = rv * cvv * (vx[c] * vx[c] + vy[c] * vy[c] + vz[c] * vz[c]);
    if (i == 0 || i == g.nx) w1 = w1 * 0.5;
    if (j == 0 || j == g.ny) w1 = w1 * 0.5;
    if (k == 0 || k == g.nz) w1 = w1 * 0.5;
    v[1] = w1 * 0.5;
  }
  const int ops[4] = {1, 1, 1, 0};
  block_reduce_store<4>(v, ops, partial);
}
