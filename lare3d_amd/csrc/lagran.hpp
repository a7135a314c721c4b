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
//   k_resist          resistive sub-cycle (:1176-1635): eta_calc, rkstep,
//                     bstep and the ohmic heating in one pass
//   k_eta             eta_calc after the remap (lare3d.f90:87)
//   predictor energy/pressure (:234-249): in the predictor remap's last
//                     geometry pass (k_geom<AX, 1>, remap.hpp)
//   k_force           f_visc (:745-789) + grad p + J x B (half-step B)
//                     + gravity -> v1 at nodes 0..n (:251-387)
//   k_shock_corr      shock_heating (:861-937) fused with the corrector
//                     (:415-462) over cells 1..n, and v = 2 v1 - v
//                     (:399-407) at nodes 0..n into a second buffer

#define NONE_ZERO 2.2250738585072014e-308
#define LARGEST 1.7976931348623157e+308
// TINY of the arithmetic type (the smallest normal float in the FP32 mode)
#define ACC_NONE_ZERO ((acc)(sizeof(acc) == sizeof(double) ? NONE_ZERO : 1.1754943508222875e-38))

struct LagArgs {
  double gamma, visc1, visc2_norm, dt, dt2;
  double eta_bg, eta0, j_max;
  double va_max2;   // boris: va_max^2 (setup.F90:80)
  double alpha_av;  // cooling_term weight
  int resistive;
  int boris;        // compute gamma_boris in set_dt (lagran.F90:1104-1108)
  int write_gb;     // 0 when set_dt returns early (first step of a restart)
  int limiter;      // the SHOCKLIMITER variant of edge_viscosity
  int compression;  // the SHOCKCOMPRESSION variant
  int cowling;      // Cowling resistivity: set_dt takes max(eta, eta_perp)
  const real *eta_perp;
};

// Partially ionised equation of state: the neutral fraction per cell and
// the ionisation potential (zero unless EOS_ION, normalise.f90:48).  The
// pressure is (gamma - 1) rho (e - (1 - xi_n) ionise_pot)
// (lagran.F90:93-94, :244).
struct EosArgs {
  const real *xi;
  double ip;
};
// e - (1 - xi_n) ionise_pot of cell c; without PI, ionise_pot is 0
template <bool PI>
__device__ __forceinline__ acc eos_e(acc e, const EosArgs &q, int c) {
  if constexpr (PI) return e - ((acc)1.0 - (acc)q.xi[c]) * (acc)q.ip;
  else return e - (acc)0.0;
}

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

// fast magnetosonic speed of shock_viscosity (:554-556) in cell c; with GB,
// gb is gamma_boris of the previous set_dt, else 1 (boris off) and gb unused
template <bool GB, bool PI>
__device__ __forceinline__ acc cell_cs(const Geo &g, const real *__restrict__ rho, const real *__restrict__ energy,
                                       const double *__restrict__ bx, const double *__restrict__ by,
                                       const double *__restrict__ bz, const real *__restrict__ gb, int c,
                                       double gamma, const EosArgs &q) {
  const acc r = rho[c];
  const acc rmin = dmax(r, ACC_NONE_ZERO);
  const acc b1x = (bx[c] + bx[c - 1]) * 0.5;
  const acc b1y = (by[c] + by[c - g.sy]) * 0.5;
  const acc b1z = (bz[c] + bz[c - g.sz]) * 0.5;
  const acc b2 = b1x * b1x + b1y * b1y + b1z * b1z;
  const acc ga = (acc)gamma;
  const acc p = (ga - (acc)1.0) * r * eos_e<PI>(energy[c], q, c);
  const acc g_b = GB ? (acc)gb[c] : (acc)1.0;
  return sqrt((ga * p + g_b * b2) / rmin);
}

// rho and the fast speed of the four cells (i,j), (i+1,j), (i,j+1), (i+1,j+1)
// of one plane, starting at cell index c
struct CellQuad {
  acc r[4], cs[4];
};
template <bool GB, bool PI>
__device__ __forceinline__ CellQuad cell_quad(const Geo &g, const real *__restrict__ rho,
                                              const real *__restrict__ energy, const double *__restrict__ bx,
                                              const double *__restrict__ by, const double *__restrict__ bz,
                                              const real *__restrict__ gb, int c, double gamma, const EosArgs &eq) {
  CellQuad q;
  const int off[4] = {0, 1, g.sy, 1 + g.sy};
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    q.r[p] = rho[c + off[p]];
    q.cs[p] = cell_cs<GB, PI>(g, rho, energy, bx, by, bz, gb, c + off[p], gamma, eq);
  }
  return q;
}

// vertex density rho_v and fast speed cs_v (lagran.F90:105-117, :561-581)
// over nodes -1..n+1; marches in z, keeping the cell values of plane k+1 for
// the next node plane, so each cell's fast speed is computed once per
// column chunk.  Sums run over the 8 cells in the order (i,j,k), (i+1,j,k),
// (i,j+1,k), (i+1,j+1,k), then the same four at k+1.  GB: boris on.  PI:
// partially ionised pressure.
template <bool GB, bool PI>
__global__ __launch_bounds__(BX * BY) void k_vertex_prep(Geo g, Box b, const real *__restrict__ rho, const real *__restrict__ energy,
                              const double *__restrict__ bx, const double *__restrict__ by,
                              const double *__restrict__ bz, const real *__restrict__ gb, real *__restrict__ rho_v,
                              real *__restrict__ cs_v, double gamma, EosArgs eq) {
  ZCOLUMN(b, i, j, k0, k1);
  const int Z = g.sz;
  CellQuad lo = cell_quad<GB, PI>(g, rho, energy, bx, by, bz, gb, g.idx(i, j, k0), gamma, eq);
  for (int k = k0; k <= k1; ++k) {
    const int c = g.idx(i, j, k);
    const CellQuad hi = cell_quad<GB, PI>(g, rho, energy, bx, by, bz, gb, c + Z, gamma, eq);
    acc cvs[8];
    const int di[8] = {0, 1, 0, 1, 0, 1, 0, 1}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1}, dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
#pragma unroll
    for (int q = 0; q < 8; ++q) cvs[q] = (acc)g.cv(i + di[q], j + dj[q], k + dk[q]);
    acc s = lo.r[0] * cvs[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) s = s + (q < 4 ? lo.r[q] : hi.r[q - 4]) * cvs[q];
    acc sc = cvs[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) sc = sc + cvs[q];
    rho_v[c] = s / sc;
    const acc cvv = (acc)0.125 * sc;
    acc scs = lo.cs[0] * cvs[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) scs = scs + (q < 4 ? lo.cs[q] : hi.cs[q - 4]) * cvs[q];
    cs_v[c] = (acc)0.125 * scs / cvv;
    lo = hi;
  }
}

// node values used by the edge viscosity
struct NodeV {
  real vx, vy, vz, r, cs;
};
__device__ __forceinline__ NodeV node_v(const real *__restrict__ vx, const real *__restrict__ vy,
                                        const real *__restrict__ vz, const real *__restrict__ rho_v,
                                        const real *__restrict__ cs_v, int c) {
  NodeV n = {vx[c], vy[c], vz[c], rho_v[c], cs_v[c]};
  return n;
}

// Velocity of a node (the outer neighbours of an edge for the limiter).
struct Vel3 {
  acc x, y, z;
};
__device__ __forceinline__ Vel3 vel3(const real *__restrict__ vx, const real *__restrict__ vy,
                                     const real *__restrict__ vz, int c) {
  Vel3 v = {vx[c], vy[c], vz[c]};
  return v;
}

// edge_viscosity (lagran.F90:795-855) between nodes 1 and 2.  With the
// SHOCKLIMITER variant (a.limiter) the velocities of node 0 (before node 1
// along the edge) and node 3 (after node 2) and the widths dxm, dxp of the
// neighbouring cells give the limiter psi (:831-845); with SHOCKCOMPRESSION
// (a.compression) an expanding edge gets no viscosity (:807-808).
__device__ __forceinline__ acc edge_visc(const NodeV &n1, const NodeV &n2, acc dvdots, double dx_,
                                         const LagArgs &a, const Vel3 &v0 = Vel3{0, 0, 0},
                                         const Vel3 &v3 = Vel3{0, 0, 0}, double dxm_ = 1.0, double dxp_ = 1.0) {
  const acc dx = (acc)dx_, dxm = (acc)dxm_, dxp = (acc)dxp_, dt = (acc)a.dt;
  if (a.compression) dvdots = dmin((acc)0.0, dvdots);
  else dvdots = -fabs(dvdots);
  const acc r1 = n1.r, r2 = n2.r;
  const acc rho_edge = (acc)2.0 * r1 * r2 / (r1 + r2);
  const acc cs_edge = dmin((acc)n1.cs, (acc)n2.cs);
  const acc dvx = n1.vx - n2.vx;
  const acc dvy = n1.vy - n2.vy;
  const acc dvz = n1.vz - n2.vz;
  const acc dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
  const acc dv = sqrt(dv2);
  acc psi = 0.0;
  if (dv * dt / dx < (acc)1.0e-14) dvdots = 0.0;
  else dvdots = dvdots / dv;
  if (a.limiter) {
    const acc dvxm = v0.x - (acc)n1.vx, dvxp = (acc)n2.vx - v3.x;
    const acc dvym = v0.y - (acc)n1.vy, dvyp = (acc)n2.vy - v3.y;
    const acc dvzm = v0.z - (acc)n1.vz, dvzp = (acc)n2.vz - v3.z;
    acc rl, rr;
    if (dv * dt / dx < (acc)1.0e-14) {
      rl = 1.0;
      rr = 1.0;
    } else {
      rl = (dvxp * dvx + dvyp * dvy + dvzp * dvz) * dx / (dxp * dv2);
      rr = (dvxm * dvx + dvym * dvy + dvzm * dvz) * dx / (dxm * dv2);
    }
    psi = dmin(dmin(dmin((acc)0.5 * (rr + rl), (acc)2.0 * rl), (acc)2.0 * rr), (acc)1.0);
    psi = dmax((acc)0.0, psi);
  }
  const acc vn = (acc)a.visc2_norm;
  const acc vc = (acc)a.visc1 * cs_edge;
  const acc q_k_bar = rho_edge * (vn * dv + sqrt(vn * vn * dv2 + vc * vc));
  return q_k_bar * ((acc)1.0 - psi) * dvdots;
}

// Caramana edge viscosities alpha1..3 (lagran.F90:583-678) on their edge
// ranges.  Marches in z: the nodes of plane k become the plane k-1 nodes of
// the next edge plane.
__global__ __launch_bounds__(BX * BY) void k_alpha(Geo g, Box b, const real *__restrict__ vx, const real *__restrict__ vy,
                        const real *__restrict__ vz, const real *__restrict__ rho_v,
                        const real *__restrict__ cs_v, real *__restrict__ al1, real *__restrict__ al2,
                        real *__restrict__ al3, LagArgs a) {
  ZCOLUMN(b, i, j, k0, k1);
  const int Y = g.sy;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int cb = g.idx(i, j, k0 - 1);
  // plane k-1 nodes (i-1, j-1), (i, j-1), (i, j)
  NodeV m1 = node_v(vx, vy, vz, rho_v, cs_v, cb - 1 - Y), m2 = node_v(vx, vy, vz, rho_v, cs_v, cb - Y),
        m3 = node_v(vx, vy, vz, rho_v, cs_v, cb);
  for (int k = k0; k <= k1; ++k) {
    const int c = g.idx(i, j, k);
    const NodeV p3 = node_v(vx, vy, vz, rho_v, cs_v, c);
    const int Z = g.sz;
    if (i >= 0 && i <= nx + 1 && j >= 0 && j <= ny + 2 && k >= 0 && k <= nz + 2) {
      if (a.limiter)
        al1[c] = edge_visc(m1, m2, -(m1.vx - m2.vx), g.dxb[i], a, vel3(vx, vy, vz, c - Z - Y - 2),
                           vel3(vx, vy, vz, c - Z - Y + 1), g.dxb[i - 1], g.dxb[i + 1]);
      else
        al1[c] = edge_visc(m1, m2, -(m1.vx - m2.vx), g.dxb[i], a);
    }
    if (i >= -1 && i <= nx + 1 && j >= 0 && j <= ny + 1 && k >= 0 && k <= nz + 2) {
      if (a.limiter)
        al2[c] = edge_visc(m2, m3, -(m2.vy - m3.vy), g.dyb[j], a, vel3(vx, vy, vz, c - Z - 2 * Y),
                           vel3(vx, vy, vz, c - Z + Y), g.dyb[j - 1], g.dyb[j + 1]);
      else
        al2[c] = edge_visc(m2, m3, -(m2.vy - m3.vy), g.dyb[j], a);
    }
    if (i >= -1 && i <= nx + 1 && j >= -1 && j <= ny + 1 && k >= 0 && k <= nz + 1) {
      if (a.limiter)
        al3[c] = edge_visc(m3, p3, -(m3.vz - p3.vz), g.dzb[k], a, vel3(vx, vy, vz, c - 2 * Z), vel3(vx, vy, vz, c + Z),
                           g.dzb[k - 1], g.dzb[k + 1]);
      else
        al3[c] = edge_visc(m3, p3, -(m3.vz - p3.vz), g.dzb[k], a);
    }
    m1 = node_v(vx, vy, vz, rho_v, cs_v, c - 1 - Y);
    m2 = node_v(vx, vy, vz, rho_v, cs_v, c - Y);
    m3 = p3;
  }
}

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

// Node values of one plane at the four nodes of a cell column: A = (i-1,j-1),
// B = (i,j-1), C = (i,j), D = (i-1,j), for one field.
struct Node4 {
  real a, b, c, d;
};
__device__ __forceinline__ Node4 node4(const real *__restrict__ u, int c, int Y) {
  Node4 n = {u[c - 1 - Y], u[c - Y], u[c], u[c - 1]};
  return n;
}
// the 12 edges of cell (i,j,k) in the order of edge_nodes, as (plane, node)
// of their first and second node: plane l = k-1, h = k; nodes a..d as Node4
#define EDGES(F)                                                                                       \
  F(0, l, a, l, b) F(1, l, b, l, c) F(2, l, c, l, d) F(3, l, d, l, a) F(4, h, a, h, b) F(5, h, b, h, c) \
  F(6, h, c, h, d) F(7, h, d, h, a) F(8, l, c, h, c) F(9, l, d, h, d) F(10, l, a, h, a) F(11, l, b, h, b)

// alpha values around a cell: al1 at (i,j), (i,j+1) and al2 at (i,j), (i-1,j)
// of one plane
struct Al12 {
  real a1, a1y, a2, a2x;
};
__device__ __forceinline__ Al12 al12(const real *__restrict__ al1, const real *__restrict__ al2, int c, int Y) {
  Al12 v = {al1[c], al1[c + Y], al2[c], al2[c - 1]};
  return v;
}

// edge_heat from values: al1/al2 of planes k (lo) and k+1 (hi), al3 at
// (i,j), (i-1,j), (i-1,j-1), (i,j-1) of plane k; same terms and order
__device__ __forceinline__ acc edge_heat_v(acc dxb, acc dyb, acc dzb, acc cv, const Al12 &lo, const Al12 &hi, acc a30,
                                           acc a3x, acc a3xy, acc a3y, const acc aa[12]) {
  acc h = -(acc)0.25 * dyb * dzb * lo.a1 * aa[0];
  h = h - (acc)0.25 * dxb * dzb * lo.a2 * aa[1];
  h = h - (acc)0.25 * dyb * dzb * lo.a1y * aa[2];
  h = h - (acc)0.25 * dxb * dzb * lo.a2x * aa[3];
  h = h - (acc)0.25 * dyb * dzb * hi.a1 * aa[4];
  h = h - (acc)0.25 * dxb * dzb * hi.a2 * aa[5];
  h = h - (acc)0.25 * dyb * dzb * hi.a1y * aa[6];
  h = h - (acc)0.25 * dxb * dzb * hi.a2x * aa[7];
  h = h - (acc)0.25 * dxb * dyb * a30 * aa[8];
  h = h - (acc)0.25 * dxb * dyb * a3x * aa[9];
  h = h - (acc)0.25 * dxb * dyb * a3xy * aa[10];
  h = h - (acc)0.25 * dxb * dyb * a3y * aa[11];
  return h / cv;
}

// p_visc + visc_heat over cells 0..n+1; pressure p0 of lagran.F90:93 stored
// for the predictor; set_dt partial minima over 0..n.  PI: partially
// ionised pressure.
template <bool PI>
__global__ __launch_bounds__(BX * BY) void k_visc_dt(Geo g, Box bx0, const real *__restrict__ vx, const real *__restrict__ vy,
                          const real *__restrict__ vz, const real *__restrict__ rho,
                          const real *__restrict__ energy, const double *__restrict__ bx,
                          const double *__restrict__ by, const double *__restrict__ bz,
                          const real *__restrict__ eta, const real *__restrict__ al1,
                          const real *__restrict__ al2, const real *__restrict__ al3,
                          real *__restrict__ visc_heat, real *__restrict__ p0, real *__restrict__ gbo,
                          double *__restrict__ partial, LagArgs a, EosArgs eq) {
  // cells 0..n+1, marching in z: node velocities and al1/al2 of plane k
  // become those of plane k-1 / k for the next cell
  PLANE_POS(bx0, i, j)
  ZRANGE(bx0, k0, k1)
  // v[2]: state check of the cells and nodes 1..n (no reference
  // counterpart): -((field + 1) 2^32 + cell) for a non-finite field value
  // or rho <= 0, fields numbered rho, energy, vx, vy, vz, bx, by, bz
  double v[3] = {LARGEST, LARGEST, LARGEST};
  if (j <= bx0.y1) {
    const int Y = g.sy, Z = g.sz;
    const int cb = g.idx(i, j, k0);
    Node4 lx = node4(vx, cb - Z, Y), ly = node4(vy, cb - Z, Y), lz = node4(vz, cb - Z, Y);
    Al12 alo = al12(al1, al2, cb, Y);
    for (int k = k0; k <= k1; ++k) {
      const int c = g.idx(i, j, k);
      const Node4 hx = node4(vx, c, Y), hy = node4(vy, c, Y), hz = node4(vz, c, Y);
      const Al12 ahi = al12(al1, al2, c + Z, Y);
      acc aa[12];
      // |v(first) - v(second)|^2 per edge
#define AAV(q, P1, N1, P2, N2)                                                                    \
  {                                                                                               \
    const acc dx = P1##x.N1 - P2##x.N2, dy = P1##y.N1 - P2##y.N2, dz = P1##z.N1 - P2##z.N2;    \
    aa[q] = dx * dx + dy * dy + dz * dz;                                                          \
  }
      EDGES(AAV)
#undef AAV
      const acc a30 = al3[c], a3x = al3[c - 1], a3xy = al3[c - 1 - Y], a3y = al3[c - Y];
      acc pv = 0.0;
      pv = dmax(pv, -(acc)alo.a1 * sqrt(aa[0]));
      pv = dmax(pv, -(acc)alo.a2 * sqrt(aa[1]));
      pv = dmax(pv, -a30 * sqrt(aa[8]));
      visc_heat[c] = edge_heat_v((acc)g.dxb[i], (acc)g.dyb[j], (acc)g.dzb[k], (acc)g.cv(i, j, k), alo, ahi, a30, a3x,
                                 a3xy, a3y, aa);
      const acc r = rho[c];
      const acc p = ((acc)a.gamma - (acc)1.0) * r * eos_e<PI>(energy[c], eq, c);
      p0[c] = p;
      if (i <= g.nx && j <= g.ny && k <= g.nz) {
        const double w1 = bx[c] * bx[c] + by[c] * by[c] + bz[c] * bz[c];
        const double rho0 = dmax((double)r, NONE_ZERO);
        double g_b = 1.0;
        if (a.boris) {
          const double w2 = w1 / rho0;
          if (w2 >= a.va_max2) g_b = 1.0 / (1.0 + w2 / a.va_max2);
          if (a.write_gb) gbo[c] = (real)g_b;
          g_b = (real)g_b;
        }
        const double cs2 = (a.gamma * (double)p + g_b * w1) / rho0;
        const double c_visc2 = pv / rho0;
        const double dxb = g.dxb[i], dyb = g.dyb[j], dzb = g.dzb[k];
        const double length = dmin(dmin(dxb, dyb), dzb);
        const double dt1 = length / (sqrt(c_visc2) + sqrt(cs2 + c_visc2));
        const double dt2 = dmin(dmin(dxb / dmax(fabs((double)hx.c), NONE_ZERO), dyb / dmax(fabs((double)hy.c), NONE_ZERO)),
                                dzb / dmax(fabs((double)hz.c), NONE_ZERO));
        v[0] = dmin(v[0], dmin(dt1, dt2));
        if (i >= 1 && j >= 1 && k >= 1) {
          const double vals[8] = {r, (double)energy[c], (double)hx.c, (double)hy.c, (double)hz.c,
                                  bx[c], by[c], bz[c]};
          const double cell = (double)((i - 1) + g.nx * ((j - 1) + g.ny * (k - 1)));
#pragma unroll
          for (int q = 0; q < 8; ++q)
            if (!__builtin_isfinite(vals[q]) || (q == 0 && !(r > 0.0)))
              v[2] = dmin(v[2], -((double)(q + 1) * 4294967296.0 + cell));
        }
        if (a.cowling) {
          const double dxlocal = 1.0 / (1.0 / (dxb * dxb) + 1.0 / (dyb * dyb) + 1.0 / (dzb * dzb));
          v[1] = dmin(v[1], 0.2 * dxlocal / dmax(dmax((double)eta[c], (double)a.eta_perp[c]), NONE_ZERO));
        } else if (a.resistive) {
          const double dxlocal = 1.0 / (1.0 / (dxb * dxb) + 1.0 / (dyb * dyb) + 1.0 / (dzb * dzb));
          v[1] = dmin(v[1], 0.2 * dxlocal / dmax((double)eta[c], NONE_ZERO));
        }
      }
      lx = hx;
      ly = hy;
      lz = hz;
      alo = ahi;
    }
  }
  const int ops[3] = {0, 0, 0};
  block_reduce_store<3>(v, ops, partial);
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

// The same vertex current in the arithmetic type A: the differences of
// the FP64 face field are taken in FP64 and divided by the widths in A (for
// A = double, the operations of vertex_current_vals).
template <class A>
__device__ __forceinline__ void vertex_current_acc(double dxc, double dyc, double dzc, const BPlane &lo,
                                                   const BPlane &hi, double bz0, double bzx, double bzy, double bzxy,
                                                   A &jx, A &jy, A &jz) {
  const A ax = (A)dxc, ay = (A)dyc, az = (A)dzc;
  const A jx1 = (A)(bzy - bz0) / ay - (A)(hi.by0 - lo.by0) / az;
  const A jx2 = (A)(bzxy - bzx) / ay - (A)(hi.byx - lo.byx) / az;
  const A jy1 = (A)(hi.bx0 - lo.bx0) / az - (A)(bzx - bz0) / ax;
  const A jy2 = (A)(hi.bxy - lo.bxy) / az - (A)(bzxy - bzy) / ax;
  const A jz1 = (A)(lo.byx - lo.by0) / ax - (A)(lo.bxy - lo.bx0) / ay;
  const A jz2 = (A)(hi.byx - hi.by0) / ax - (A)(hi.bxy - hi.bx0) / ay;
  jx = (jx1 + jx2) * (A)0.5;
  jy = (jy1 + jy2) * (A)0.5;
  jz = (jz1 + jz2) * (A)0.5;
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
    acc jx, jy, jz;
    vertex_current_acc<acc>(g.dxc[i], g.dyc[j], g.dzc[k], lo, hi, bz[c], bz[c + 1], bz[c + Y], bz[c + 1 + Y], jx, jy,
                            jz);
    eta[c] = sqrt(jx * jx + jy * jy + jz * jz) > (acc)j_max ? eta_bg + eta0 : eta_bg;
    lo = hi;
  }
}

__global__ __launch_bounds__(BX * BY) void k_fill(Geo g, Box b, real *__restrict__ a, double v) {
  BOX_INDEX(b, i, j, k);
  a[g.idx(i, j, k)] = v;
}

// rkstep's EMFs and curlb at one node with Cowling resistivity
// (lagran.F90:1504-1577): the vertex current j splits into j_par along the
// vertex field and j_perp = j - j_par; the EMF is -((j_par + j_perp) eta +
// j_perp eta_perp) d/2 per component and curlb = eta |j_par|^2 + (eta_perp
// + eta) |j_perp|^2.  lo, hi: face bx, by of planes k and k+1; bz0..bzxy: bz
// at (i,j), (i+1,j), (i,j+1), (i+1,j+1) of plane k.
__device__ __forceinline__ void cowling_emf(const BPlane &lo, const BPlane &hi, double bz0, double bzx, double bzy,
                                            double bzxy, acc jx, acc jy, acc jz, acc e, acc ep, acc dxc, acc dyc,
                                            acc dzc, acc &kx, acc &ky, acc &kz, acc &cb) {
  const acc bxv = ((acc)lo.bx0 + (acc)lo.bxy + (acc)hi.bx0 + (acc)hi.bxy) * (acc)0.25;
  const acc byv = ((acc)lo.by0 + (acc)lo.byx + (acc)hi.by0 + (acc)hi.byx) * (acc)0.25;
  const acc bzv = ((acc)bz0 + (acc)bzx + (acc)bzy + (acc)bzxy) * (acc)0.25;
  const acc magn_b = bxv * bxv + byv * byv + bzv * bzv;
  acc jpx = 0.0, jpy = 0.0, jpz = 0.0;
  if (magn_b > ACC_NONE_ZERO) {
    jpx = (jx * bxv + jy * byv + jz * bzv) * bxv / magn_b;
    jpy = (jx * bxv + jy * byv + jz * bzv) * byv / magn_b;
    jpz = (jx * bxv + jy * byv + jz * bzv) * bzv / magn_b;
  }
  const acc jqx = jx - jpx, jqy = jy - jpy, jqz = jz - jpz;
  const acc mpar = sqrt(jpx * jpx + jpy * jpy + jpz * jpz);
  const acc mperp = sqrt(jqx * jqx + jqy * jqy + jqz * jqz);
  cb = e * (mpar * mpar) + (ep + e) * (mperp * mperp);
  kx = -((jpx + jqx) * e + jqx * ep) * dxc * (acc)0.5;
  ky = -((jpy + jqy) * e + jqy * ep) * dyc * (acc)0.5;
  kz = -((jpz + jqz) * e + jqz * ep) * dzc * (acc)0.5;
}

// EMFs around a node in one plane, named as in bstep (lagran.F90:1582-1631):
// kx at (i,j), (i-1,j); ky at (i,j), (i,j-1); kz at (i,j), (i,j-1), (i-1,j)
struct EPlane {
  acc kx0, kxm, ky0, kym, kz0, kzm, kzx;
};

// One resistive substep in a single pass: at every node of the box
// -1..n+1 the vertex current and eta of the substep's eta_calc
// (lagran.F90:1176-1230), at nodes 0..n the rkstep EMFs and curlb
// (:1472-1503) with the total_ohmic_heating increments (:1309-1328), and
// bstep (:1582-1631), which takes the EMFs of the node and of its (i-1),
// (j-1) neighbours in planes k and k-1.  The ohmic heating of cell (i,j,k)
// (:1270-1288, cells 1..n) takes curlb of the same eight nodes, rounded to
// the storage type as stored values are, and energy = MAX(energy, 0)
// (:1305) follows over the cell box -1..n+2.  A block is a tile of RS_ROWS rows of
// 64 nodes marching in z over ZCHUNK planes (plus plane k0-1 for the EMFs of
// plane k-1); lane 0 and row 0 only supply neighbour EMFs, which pass through
// LDS, so a tile writes RS_OUT_X x RS_OUT_Y nodes.  Neighbouring tiles still
// read the old B, so the new B goes to bxo..bzo at every node of the launch
// box (the padded box): updated where bstep updates it, copied elsewhere.
// COWL: rkstep's Cowling form with eta_perp at the nodes.
#define RS_ROWS 8
#define RS_OUT_X 63
#define RS_OUT_Y (RS_ROWS - 1)
template <bool COWL>
__global__ __launch_bounds__(64 * RS_ROWS) void k_resist(Geo g, Box b, int tiles_x, const double *__restrict__ bx,
                                                         const double *__restrict__ by, const double *__restrict__ bz,
                                                         double *__restrict__ bxo, double *__restrict__ byo,
                                                         double *__restrict__ bzo, real *__restrict__ eta,
                                                         real *__restrict__ energy, real *__restrict__ ohmic_dep,
                                                         const real *__restrict__ rho, real *__restrict__ curlb_out,
                                                         double *__restrict__ partial, double dt, double eta_bg,
                                                         double eta0, double j_max, const real *__restrict__ eta_perp) {
  // EMFs kx, ky, kz and curlb of the tile's nodes, by plane parity
  __shared__ acc sk[2][4][RS_ROWS][64];
  __shared__ double red[RS_ROWS];
  const int l = (int)threadIdx.x, w = (int)threadIdx.y;
  const int t = (int)xcd_tile(blockIdx.x, gridDim.x);
  const int ty = t / tiles_x, tx = t - ty * tiles_x;
  const int i = b.x0 + tx * RS_OUT_X + l - 1;
  const int j = b.y0 + ty * RS_OUT_Y + w - 1;
  ZRANGE(b, k0, k1)
  const bool out = l >= 1 && w >= 1 && i <= b.x1 && j <= b.y1;
  const int Y = g.sy, Z = g.sz;
  const bool eij = i >= -1 && i <= g.nx + 1 && j >= -1 && j <= g.ny + 1;  // eta box, in the plane
  const bool kij = i >= 0 && i <= g.nx && j >= 0 && j <= g.ny;            // EMF box, in the plane
  // cell box of the energy clamp, in the plane
  const bool cij = i >= -1 && i <= g.nx + 2 && j >= -1 && j <= g.ny + 2;
  const bool hij = i >= 1 && i <= g.nx && j >= 1 && j <= g.ny;  // heated cells, in the plane
  double v = 0.0;
  EPlane m = {0, 0, 0, 0, 0, 0, 0};  // plane k-1
  acc mq[4] = {0, 0, 0, 0};          // curlb of plane k-1 at (i,j), (i-1,j), (i,j-1), (i-1,j-1)
  BPlane lo = {0, 0, 0, 0};
  bool have_lo = false;
  for (int k = k0 - 1; k <= k1; ++k) {
    const int c = g.idx(i, j, k);
    acc kx = 0.0, ky = 0.0, kz = 0.0, cbr = 0.0;
    if (eij && k >= -1 && k <= g.nz + 1) {
      if (!have_lo) lo = bplane(bx, by, c, Y);
      const BPlane hi = bplane(bx, by, c + Z, Y);
      acc jx, jy, jz;
      vertex_current_acc<acc>(g.dxc[i], g.dyc[j], g.dzc[k], lo, hi, bz[c], bz[c + 1], bz[c + Y], bz[c + 1 + Y], jx,
                              jy, jz);
      const BPlane plo = lo;
      lo = hi;
      have_lo = true;
      const real es = sqrt(jx * jx + jy * jy + jz * jz) > (acc)j_max ? eta_bg + eta0 : eta_bg;
      if (out && k >= k0) eta[c] = es;
      if (kij && k >= 0 && k <= g.nz) {
        const acc e = es;
        acc cb;
        if constexpr (COWL) {
          cowling_emf(plo, hi, bz[c], bz[c + 1], bz[c + Y], bz[c + 1 + Y], jx, jy, jz, e, (acc)eta_perp[c],
                      (acc)g.dxc[i], (acc)g.dyc[j], (acc)g.dzc[k], kx, ky, kz, cb);
        } else {
          kx = -jx * e * (acc)g.dxc[i] * (acc)0.5;
          ky = -jy * e * (acc)g.dyc[j] * (acc)0.5;
          kz = -jz * e * (acc)g.dzc[k] * (acc)0.5;
          cb = e * (jx * jx + jy * jy + jz * jz);
        }
        cbr = (real)cb;
        if (out && k >= k0) {
          if (curlb_out) curlb_out[c] = cbr;
          double w1 = dt * g.dxc[i] * g.dyc[j] * g.dzc[k] * cb;
          if (i == 0 || i == g.nx) w1 = w1 * 0.5;
          if (j == 0 || j == g.ny) w1 = w1 * 0.5;
          if (k == 0 || k == g.nz) w1 = w1 * 0.5;
          v += w1;
        }
      }
    }
    acc(&s)[4][RS_ROWS][64] = sk[k & 1];
    s[0][w][l] = kx;
    s[1][w][l] = ky;
    s[2][w][l] = kz;
    s[3][w][l] = cbr;
    __syncthreads();
    if (out) {
      // EMFs at (i,j), (i-1,j), (i,j-1) of this plane, as EPlane names them
      const EPlane p = {kx, s[0][w][l - 1], ky, s[1][w - 1][l], kz, s[2][w - 1][l], s[2][w][l - 1]};
      if (k >= k0) {
        const bool upd = kij && k >= 0 && k <= g.nz;
        double nbx = bx[c], nby = by[c], nbz = bz[c];
        if (upd && j >= 1 && k >= 1) {
          const double area = g.dyb[j] * g.dzb[k];
          nbx = nbx + ((double)p.kz0 - p.kzm + m.kz0 - m.kzm - p.ky0 + m.ky0 - p.kym + m.kym) * dt / area;
        }
        if (upd && i >= 1 && k >= 1) {
          const double area = g.dxb[i] * g.dzb[k];
          nby = nby + ((double)p.kx0 - m.kx0 + p.kxm - m.kxm - p.kz0 + p.kzx - m.kz0 + m.kzx) * dt / area;
        }
        if (upd && i >= 1 && j >= 1) {
          const double area = g.dxb[i] * g.dyb[j];
          nbz = nbz + ((double)p.ky0 - s[1][w][l - 1] + p.kym - s[1][w - 1][l - 1] - p.kx0 + s[0][w - 1][l] - p.kxm +
                       s[0][w - 1][l - 1]) *
                          dt / area;
        }
        bxo[c] = nbx;
        byo[c] = nby;
        bzo[c] = nbz;
      }
      const acc pq[4] = {cbr, s[3][w][l - 1], s[3][w - 1][l], s[3][w - 1][l - 1]};
      if (k >= k0 && cij && k >= -1 && k <= g.nz + 2) {
        acc e = energy[c];
        if (hij && k >= 1 && k <= g.nz) {
          // curlb(c) + (c-X) + (c-Y) + (c-X-Y) + (c-Z) + (c-X-Z) + (c-Y-Z) + (c-X-Y-Z)
          const acc sq = pq[0] + pq[1] + pq[2] + pq[3] + mq[0] + mq[1] + mq[2] + mq[3];
          const acc lh = sq * (acc)dt / ((acc)8.0 * rho[c]);
          ohmic_dep[c] = ohmic_dep[c] + lh;
          e = e + lh;
        }
        // with the cooling_term, k_cool_b takes the cooling off first and
        // then clamps
        energy[c] = curlb_out ? e : max0(e);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) mq[q] = pq[q];
      m = p;
    }
  }
  v = wave_sum(v);
  if (l == 0) red[w] = v;
  __syncthreads();
  if (l == 0 && w == 0) {
    double r = red[0];
    for (int q = 1; q < RS_ROWS; ++q) r = r + red[q];
    partial[(size_t)blockIdx.z * gridDim.x + blockIdx.x] = r;
  }
}

__global__ __launch_bounds__(BX * BY) void k_energy_max(Geo g, Box b, real *__restrict__ energy) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  energy[c] = max0(energy[c]);
}

// FOURTHORDER resistive substep (lagran.F90:1329-1415): four rksteps, each
// followed by a bstep from the substep's starting field B1.  k_rk4_emf is
// rkstep (:1472-1503) at nodes 0..n from the current B and this substep's
// eta: the EMFs and curlb of the stage go to k, and the weighted sum
// k1 + 2 k2 + 2 k3 + k4 builds up in ks in the reference's order (w = 0
// starts it with k1, then w = 2, 2, 1).
struct RkK {
  real *k[4];  // kx, ky, kz, curlb
};
template <bool COWL>
__global__ __launch_bounds__(BX * BY) void k_rk4_emf(Geo g, Box b, const double *__restrict__ bx,
                                                     const double *__restrict__ by, const double *__restrict__ bz,
                                                     const real *__restrict__ eta, RkK k, RkK ks, int w,
                                                     const real *__restrict__ eta_perp) {
  BOX_INDEX(b, i, j, kk);
  const int Y = g.sy, Z = g.sz;
  const int c = g.idx(i, j, kk);
  const BPlane lo = bplane(bx, by, c, Y), hi = bplane(bx, by, c + Z, Y);
  acc jx, jy, jz;
  vertex_current_acc<acc>(g.dxc[i], g.dyc[j], g.dzc[kk], lo, hi, bz[c], bz[c + 1], bz[c + Y], bz[c + 1 + Y], jx, jy,
                          jz);
  const acc e = eta[c];
  real v[4];
  if constexpr (COWL) {
    acc kx, ky, kz, cb;
    cowling_emf(lo, hi, bz[c], bz[c + 1], bz[c + Y], bz[c + 1 + Y], jx, jy, jz, e, (acc)eta_perp[c],
                (acc)g.dxc[i], (acc)g.dyc[j], (acc)g.dzc[kk], kx, ky, kz, cb);
    v[0] = kx;
    v[1] = ky;
    v[2] = kz;
    v[3] = cb;
  } else {
    v[0] = -jx * e * (acc)g.dxc[i] * (acc)0.5;
    v[1] = -jy * e * (acc)g.dyc[j] * (acc)0.5;
    v[2] = -jz * e * (acc)g.dzc[kk] * (acc)0.5;
    v[3] = e * (jx * jx + jy * jy + jz * jz);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    k.k[q][c] = v[q];
    ks.k[q][c] = w == 0 ? v[q] : (real)((acc)ks.k[q][c] + (acc)w * (acc)v[q]);
  }
}

// bstep(kx, ky, kz, h) (:1582-1631) of the EMFs k from B1: bx on x = 0..n,
// y, z = 1..n, and likewise by and bz; the other faces are left to the
// bfield_bcs that follows.  B accumulates in FP64.
__global__ __launch_bounds__(BX * BY) void k_rk4_bstep(Geo g, Box b, const double *__restrict__ b1x,
                                                       const double *__restrict__ b1y,
                                                       const double *__restrict__ b1z, RkK k, double *__restrict__ bx,
                                                       double *__restrict__ by, double *__restrict__ bz, double h) {
  BOX_INDEX(b, i, j, kk);
  const int X = 1, Y = g.sy, Z = g.sz;
  const int c = g.idx(i, j, kk);
  const real *kx = k.k[0], *ky = k.k[1], *kz = k.k[2];
  if (j >= 1 && kk >= 1) {
    const double area = g.dyb[j] * g.dzb[kk];
    bx[c] = b1x[c] + ((double)kz[c] - kz[c - Y] + kz[c - Z] - kz[c - Y - Z] - ky[c] + ky[c - Z] - ky[c - Y] +
                      ky[c - Y - Z]) * h / area;
  }
  if (i >= 1 && kk >= 1) {
    const double area = g.dxb[i] * g.dzb[kk];
    by[c] = b1y[c] + ((double)kx[c] - kx[c - Z] + kx[c - X] - kx[c - X - Z] - kz[c] + kz[c - X] - kz[c - Z] +
                      kz[c - X - Z]) * h / area;
  }
  if (i >= 1 && j >= 1) {
    const double area = g.dxb[i] * g.dyb[j];
    bz[c] = b1z[c] + ((double)ky[c] - ky[c - X] + ky[c - Y] - ky[c - X - Y] - kx[c] + kx[c - Y] - kx[c - X] +
                      kx[c - X - Y]) * h / area;
  }
}

// the heating of the weighted curlb cs on cells 1..n (:1403-1413), with no
// ohmic_dep, cooling or floor in this variant
__global__ __launch_bounds__(BX * BY) void k_rk4_heat(Geo g, Box b, const real *__restrict__ cs,
                                                      const real *__restrict__ rho, real *__restrict__ energy,
                                                      double dt6) {
  BOX_INDEX(b, i, j, k);
  const int X = 1, Y = g.sy, Z = g.sz;
  const int c = g.idx(i, j, k);
  const acc sq = (acc)cs[c] + cs[c - X] + cs[c - Y] + cs[c - X - Y] + cs[c - Z] + cs[c - X - Z] + cs[c - Y - Z] +
                 cs[c - X - Y - Z];
  energy[c] = energy[c] + sq * (acc)dt6 / ((acc)8.0 * rho[c]);
}

// total_ohmic_heating increments dt dxc dyc dzc curlb over nodes 0..n,
// halved on each boundary plane (:1309-1328); block partial sums
__global__ __launch_bounds__(BX * BY) void k_ohmic_total(Geo g, Box b, const real *__restrict__ cb,
                                                         double *__restrict__ partial, double dt) {
  PLANE_POS(b, i, j)
  const int k = b.z0 + (int)blockIdx.z;
  double v[1] = {0.0};
  if (j <= b.y1) {
    double w1 = dt * g.dxc[i] * g.dyc[j] * g.dzc[k] * (double)cb[g.idx(i, j, k)];
    if (i == 0 || i == g.nx) w1 = w1 * 0.5;
    if (j == 0 || j == g.ny) w1 = w1 * 0.5;
    if (k == 0 || k == g.nz) w1 = w1 * 0.5;
    v[0] = w1;
  }
  const int ops[1] = {1};
  block_reduce_store<1>(v, ops, partial);
}

// The ohmic part of the cooling_term after a resistive substep
// (lagran.F90:1290-1302) on cells 1..n, then energy = MAX(energy, 0).  The
// reference's loop sets no ixm or iym, which keep the last values of the
// heating loop before it, nx - 1 and ny - 1: the average takes curlb at
// (i, j, k), (nx-1, j, k), (i, ny-1, k) and (nx-1, ny-1, k).
__global__ __launch_bounds__(BX * BY) void k_cool_b(Geo g, Box b, const real *__restrict__ curlb,
                                                    const real *__restrict__ rho, real *__restrict__ cool_b,
                                                    real *__restrict__ energy, double alpha_av, double dt) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const double lh = ((double)curlb[c] + curlb[g.idx(g.nx - 1, j, k)] + curlb[g.idx(i, g.ny - 1, k)] +
                     curlb[g.idx(g.nx - 1, g.ny - 1, k)]) *
                    dt / (4.0 * rho[c]);
  const real ct = alpha_av * lh + (1.0 - alpha_av) * cool_b[c];
  cool_b[c] = ct;
  energy[c] = max0(energy[c] - ct);
}

// b_field_and_cv1_update with CAUCHY (lagran.F90:945-1048) over the cell box
// -1..n+2: cv1, and the cell-centred B1 (the average of each component's two
// faces, :93-95) advanced by the Cauchy solution
// B1 <- (B1 + (B1 . grad) v dt/2) / (1 + div v dt/2), stored times cv1
// (:230-232) in b1x..b1z; on cells 0..n+1 also the predictor energy and
// pressure (:234-249), from the start-of-step pressure p0.
__global__ __launch_bounds__(BX * BY) void k_cauchy_b(Geo g, Box b, const real *__restrict__ vx,
                                                      const real *__restrict__ vy, const real *__restrict__ vz,
                                                      const double *__restrict__ bx, const double *__restrict__ by,
                                                      const double *__restrict__ bz, real *__restrict__ cv1,
                                                      real *__restrict__ b1x, real *__restrict__ b1y,
                                                      real *__restrict__ b1z, const real *__restrict__ energy,
                                                      const real *__restrict__ rho,
                                                      const real *__restrict__ visc_heat,
                                                      const real *__restrict__ p0, real *__restrict__ pp,
                                                      double gamma, double dt2, EosArgs eq) {
  BOX_INDEX(b, i, j, k);
  const int X = 1, Y = g.sy, Z = g.sz;
  const int c = g.idx(i, j, k);
  // node velocity at offset (di, dj, dk) in {-1, 0} from node (i, j, k)
  auto at = [&](const real *__restrict__ v, int di, int dj, int dk) -> acc { return v[c + di * X + dj * Y + dk * Z]; };
  const acc q4 = 0.25, h = (acc)dt2;
  const acc dxb = (acc)g.dxb[i], dyb = (acc)g.dyb[j], dzb = (acc)g.dzb[k];
  // normal derivatives at the faces of B (:958-984)
  acc vxb = (at(vx, 0, 0, 0) + at(vx, 0, -1, 0) + at(vx, 0, 0, -1) + at(vx, 0, -1, -1)) * q4;
  acc vxbm = (at(vx, -1, 0, 0) + at(vx, -1, -1, 0) + at(vx, -1, 0, -1) + at(vx, -1, -1, -1)) * q4;
  acc vyb = (at(vy, 0, 0, 0) + at(vy, -1, 0, 0) + at(vy, 0, 0, -1) + at(vy, -1, 0, -1)) * q4;
  acc vybm = (at(vy, 0, -1, 0) + at(vy, -1, -1, 0) + at(vy, 0, -1, -1) + at(vy, -1, -1, -1)) * q4;
  acc vzb = (at(vz, 0, 0, 0) + at(vz, -1, 0, 0) + at(vz, 0, -1, 0) + at(vz, -1, -1, 0)) * q4;
  acc vzbm = (at(vz, 0, 0, -1) + at(vz, -1, 0, -1) + at(vz, 0, -1, -1) + at(vz, -1, -1, -1)) * q4;
  const acc dvxdx = (vxb - vxbm) / dxb;
  const acc dvydy = (vyb - vybm) / dyb;
  const acc dvzdz = (vzb - vzbm) / dzb;
  const acc dv = (dvxdx + dvydy + dvzdz) * h;
  const acc cv = (acc)g.cv(i, j, k);
  const real c1 = cv * ((acc)1.0 + dv);
  cv1[c] = c1;
  // transverse derivatives (:989-1033)
  vxb = (at(vx, 0, 0, 0) + at(vx, -1, 0, 0) + at(vx, 0, 0, -1) + at(vx, -1, 0, -1)) * q4;
  vxbm = (at(vx, 0, -1, 0) + at(vx, -1, -1, 0) + at(vx, 0, -1, -1) + at(vx, -1, -1, -1)) * q4;
  vyb = (at(vy, 0, 0, 0) + at(vy, 0, -1, 0) + at(vy, 0, 0, -1) + at(vy, 0, -1, -1)) * q4;
  vybm = (at(vy, -1, 0, 0) + at(vy, -1, -1, 0) + at(vy, -1, 0, -1) + at(vy, -1, -1, -1)) * q4;
  const acc dvxdy = (vxb - vxbm) / dyb;
  const acc dvydx = (vyb - vybm) / dxb;
  vxb = (at(vx, 0, 0, 0) + at(vx, -1, 0, 0) + at(vx, 0, -1, 0) + at(vx, -1, -1, 0)) * q4;
  vxbm = (at(vx, 0, 0, -1) + at(vx, -1, 0, -1) + at(vx, 0, -1, -1) + at(vx, -1, -1, -1)) * q4;
  acc vzb2 = (at(vz, 0, 0, 0) + at(vz, 0, -1, 0) + at(vz, 0, 0, -1) + at(vz, 0, -1, -1)) * q4;
  acc vzbm2 = (at(vz, -1, 0, 0) + at(vz, -1, -1, 0) + at(vz, -1, 0, -1) + at(vz, -1, -1, -1)) * q4;
  const acc dvxdz = (vxb - vxbm) / dzb;
  const acc dvzdx = (vzb2 - vzbm2) / dxb;
  vyb = (at(vy, 0, 0, 0) + at(vy, -1, 0, 0) + at(vy, 0, -1, 0) + at(vy, -1, -1, 0)) * q4;
  vybm = (at(vy, 0, 0, -1) + at(vy, -1, 0, -1) + at(vy, 0, -1, -1) + at(vy, -1, -1, -1)) * q4;
  vzb2 = (at(vz, 0, 0, 0) + at(vz, -1, 0, 0) + at(vz, 0, 0, -1) + at(vz, -1, 0, -1)) * q4;
  vzbm2 = (at(vz, 0, -1, 0) + at(vz, -1, -1, 0) + at(vz, 0, -1, -1) + at(vz, -1, -1, -1)) * q4;
  const acc dvydz = (vyb - vybm) / dzb;
  const acc dvzdy = (vzb2 - vzbm2) / dyb;
  const acc bx1 = (bx[c] + bx[c - X]) * 0.5, by1 = (by[c] + by[c - Y]) * 0.5, bz1 = (bz[c] + bz[c - Z]) * 0.5;
  const acc w3 = bx1 * dvxdx + by1 * dvxdy + bz1 * dvxdz;
  const acc w4 = bx1 * dvydx + by1 * dvydy + bz1 * dvydz;
  const acc w5 = bx1 * dvzdx + by1 * dvzdy + bz1 * dvzdz;
  const real nx = (bx1 + w3 * h) / ((acc)1.0 + dv);
  const real ny = (by1 + w4 * h) / ((acc)1.0 + dv);
  const real nz = (bz1 + w5 * h) / ((acc)1.0 + dv);
  b1x[c] = (acc)nx * (acc)c1;
  b1y[c] = (acc)ny * (acc)c1;
  b1z[c] = (acc)nz * (acc)c1;
  if (i >= 0 && i <= g.nx + 1 && j >= 0 && j <= g.ny + 1 && k >= 0 && k <= g.nz + 1) {
    // predictor energy and pressure (lagran.F90:234-249)
    const acc r = rho[c];
    const acc dvv = c1 / cv - (acc)1.0;
    acc e1 = energy[c] - p0[c] * dvv / r;
    e1 = e1 + visc_heat[c] * h / r;
    const acc ei = eq.xi ? eos_e<true>(e1, eq, c) : eos_e<false>(e1, eq, c);
    pp[c] = ei * ((acc)gamma - (acc)1.0) * r * cv / c1;
  }
}

// B1 cv1 and cv1 of the four cells (i, j), (i+1, j), (i, j+1), (i+1, j+1)
// of one plane (CAUCHY predictor)
struct CPlane {
  acc v[4], x[4], y[4], z[4];
};
__device__ __forceinline__ CPlane cplane(const real *__restrict__ cv1, const real *__restrict__ b1x,
                                         const real *__restrict__ b1y, const real *__restrict__ b1z, int c, int Y) {
  CPlane p;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int o = c + (q & 1) + (q >> 1) * Y;
    p.v[q] = cv1[o];
    p.x[q] = b1x[o];
    p.y[q] = b1y[o];
    p.z[q] = b1z[o];
  }
  return p;
}
// J and B at node (i, j, k) of the CAUCHY predictor (lagran.F90:285-348)
// from the cells of planes k (lo) and k+1 (hi), in the reference's order
__device__ __forceinline__ void cauchy_jb(acc dxc, acc dyc, acc dzc, const CPlane &lo, const CPlane &hi, acc &jx,
                                          acc &jy, acc &jz, acc &bxv, acc &byv, acc &bzv) {
  const acc cvx = lo.v[0] + lo.v[2] + hi.v[0] + hi.v[2];
  const acc cvxp = lo.v[1] + lo.v[3] + hi.v[1] + hi.v[3];
  const acc cvy = lo.v[0] + lo.v[1] + hi.v[0] + hi.v[1];
  const acc cvyp = lo.v[2] + lo.v[3] + hi.v[2] + hi.v[3];
  const acc cvz = lo.v[0] + lo.v[1] + lo.v[2] + lo.v[3];
  const acc cvzp = hi.v[0] + hi.v[1] + hi.v[2] + hi.v[3];
  acc w1 = (lo.z[0] + lo.z[1] + hi.z[0] + hi.z[1]) / cvy;
  acc w2 = (lo.z[2] + lo.z[3] + hi.z[2] + hi.z[3]) / cvyp;
  jx = (w2 - w1) / dyc;
  w1 = (lo.y[0] + lo.y[1] + lo.y[2] + lo.y[3]) / cvz;
  w2 = (hi.y[0] + hi.y[1] + hi.y[2] + hi.y[3]) / cvzp;
  jx = jx - (w2 - w1) / dzc;
  w1 = (lo.z[0] + lo.z[2] + hi.z[0] + hi.z[2]) / cvx;
  w2 = (lo.z[1] + lo.z[3] + hi.z[1] + hi.z[3]) / cvxp;
  jy = -(w2 - w1) / dxc;
  w1 = (lo.x[0] + lo.x[1] + lo.x[2] + lo.x[3]) / cvz;
  w2 = (hi.x[0] + hi.x[1] + hi.x[2] + hi.x[3]) / cvzp;
  jy = jy + (w2 - w1) / dzc;
  w1 = (lo.y[0] + lo.y[2] + hi.y[0] + hi.y[2]) / cvx;
  w2 = (lo.y[1] + lo.y[3] + hi.y[1] + hi.y[3]) / cvxp;
  jz = (w2 - w1) / dxc;
  w1 = (lo.x[0] + lo.x[1] + hi.x[0] + hi.x[1]) / cvy;
  w2 = (lo.x[2] + lo.x[3] + hi.x[2] + hi.x[3]) / cvyp;
  jz = jz - (w2 - w1) / dyc;
  bxv = (lo.x[0] + lo.x[1] + lo.x[2] + lo.x[3] + hi.x[0] + hi.x[1] + hi.x[2] + hi.x[3]) / (cvx + cvxp);
  byv = (lo.y[0] + lo.y[1] + lo.y[2] + lo.y[3] + hi.y[0] + hi.y[1] + hi.y[2] + hi.y[3]) / (cvx + cvxp);
  bzv = (lo.z[0] + lo.z[1] + lo.z[2] + lo.z[3] + hi.z[0] + hi.z[1] + hi.z[2] + hi.z[3]) / (cvx + cvxp);
}

// predictor momentum (lagran.F90:251-387) with fx/fy/fz_visc (:745-789)
// recomputed here from alpha and v.  Without CAUCHY, B is the half-step
// (predictor-remapped) face field; with CAUCHY, J and B come from the
// cell-centred B1 cv1 (b1x..b1z) and cv1 of k_cauchy_b.
template <bool CAUCHY>
__global__ __launch_bounds__(BX * BY) void k_force(Geo g, Box b, const real *__restrict__ vx, const real *__restrict__ vy,
                        const real *__restrict__ vz, const real *__restrict__ pp,
                        const double *__restrict__ bx, const double *__restrict__ by, const double *__restrict__ bz,
                        const real *__restrict__ rho_v, const real *__restrict__ al1,
                        const real *__restrict__ al2, const real *__restrict__ al3, const real *__restrict__ gb,
                        real *__restrict__ vx1, real *__restrict__ vy1, real *__restrict__ vz1, double dt2,
                        const real *__restrict__ cv1, const real *__restrict__ b1x, const real *__restrict__ b1y,
                        const real *__restrict__ b1z) {
  // gb: gamma_boris of this step's set_dt (lagran.F90:375-377), NULL for 1
  ZCOLUMN(b, i, j, k0, k1);
  const int X = 1, Y = g.sy, Z = g.sz;
  const real *V[3] = {vx, vy, vz};
  // marching windows: node velocity at (i, j) of planes k-1, k; pressure,
  // face B and alpha3 of plane k
  const int cb = g.idx(i, j, k0);
  real um[3], u0[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    um[q] = V[q][cb - Z];
    u0[q] = V[q][cb];
  }
  Quad plo = quad_fwd(pp, cb, Y);
  BPlane blo = {};
  CPlane clo = {};
  if (CAUCHY) clo = cplane(cv1, b1x, b1y, b1z, cb, Y);
  else blo = bplane(bx, by, cb, Y);
  real a3lo = al3[cb];
  for (int k = k0; k <= k1; ++k) {
    const int c = g.idx(i, j, k);
    const acc dxc = (acc)g.dxc[i], dyc = (acc)g.dyc[j], dzc = (acc)g.dzc[k];
    real up[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) up[q] = V[q][c + Z];
    const Quad phi = quad_fwd(pp, c + Z, Y);
    BPlane bhi = {};
    CPlane chi = {};
    if (CAUCHY) chi = cplane(cv1, b1x, b1y, b1z, c + Z, Y);
    else bhi = bplane(bx, by, c + Z, Y);
    const real a3hi = al3[c + Z];
    // viscous forces
    const acc a1 = al1[c + Y + Z] * dyc * dzc;
    const acc a2 = al1[c + X + Y + Z] * dyc * dzc;
    const acc a3 = al2[c + Z] * dxc * dzc;
    const acc a4 = al2[c + Y + Z] * dxc * dzc;
    const acc a5 = a3lo * dxc * dyc;
    const acc a6 = a3hi * dxc * dyc;
    const acc cvv = (acc)0.125 * (acc)cv_sum8(g, i, j, k);
    acc fv[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const real *u = V[q];
      const acc uc = u0[q];
      fv[q] = (a1 * (uc - u[c - X]) + a2 * (uc - u[c + X]) + a3 * (uc - u[c - Y]) + a4 * (uc - u[c + Y]) +
               a5 * (uc - um[q]) + a6 * (uc - up[q])) /
              cvv;
    }
    const acc q4 = 0.25;
    const acc p = plo.p0, ppx = plo.px, ppy = plo.py, ppz = phi.p0;
    const acc ppxy = plo.pxy, ppxz = phi.px, ppyz = phi.py, ppxyz = phi.pxy;
    acc w1 = (p + ppy + ppz + ppyz) * q4;
    acc w2 = (ppx + ppxy + ppxz + ppxyz) * q4;
    acc fx = -(w2 - w1) / dxc;
    w1 = (p + ppx + ppz + ppxz) * q4;
    w2 = (ppy + ppxy + ppyz + ppxyz) * q4;
    acc fy = -(w2 - w1) / dyc;
    w1 = (p + ppx + ppy + ppxy) * q4;
    w2 = (ppz + ppxz + ppyz + ppxyz) * q4;
    acc fz = -(w2 - w1) / dzc;
    acc bxv, byv, bzv, jx, jy, jz;
    if (CAUCHY) {
      cauchy_jb(dxc, dyc, dzc, clo, chi, jx, jy, jz, bxv, byv, bzv);
    } else {
      // bx(c) + bx(c+Y) + bx(c+Z) + bx(c+Y+Z), and likewise for by and bz
      // (the FP64 face field averaged in FP64)
      const double bz0 = bz[c], bzx = bz[c + X], bzy = bz[c + Y], bzxy = bz[c + X + Y];
      bxv = 0.25 * (blo.bx0 + blo.bxy + bhi.bx0 + bhi.bxy);
      byv = 0.25 * (blo.by0 + blo.byx + bhi.by0 + bhi.byx);
      bzv = 0.25 * (bz0 + bzy + bzx + bzxy);
      vertex_current_acc<acc>(g.dxc[i], g.dyc[j], g.dzc[k], blo, bhi, bz0, bzx, bzy, bzxy, jx, jy, jz);
    }
    const acc g_b = gb ? (acc)gb[c] : (acc)1.0;
    fx = fx + g_b * (jy * bzv - jz * byv);
    fy = fy + g_b * (jz * bxv - jx * bzv);
    fz = fz + g_b * (jx * byv - jy * bxv);
    const acc rv = rho_v[c];
    fz = fz - rv * (acc)g.grav[k];
    const acc h = (acc)dt2;
    vx1[c] = u0[0] + h * (fv[0] + fx) / rv;
    vy1[c] = u0[1] + h * (fv[1] + fy) / rv;
    vz1[c] = u0[2] + h * (fv[2] + fz) / rv;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      um[q] = u0[q];
      u0[q] = up[q];
    }
    plo = phi;
    blo = bhi;
    clo = chi;
    a3lo = a3hi;
  }
}

// shock_heating (lagran.F90:861-937) + corrector (:415-462) over cells
// 1..n; total_visc_heating partial sums.  COOL: cooling_term on.
template <bool COOL>
__global__ __launch_bounds__(BX * BY) void k_shock_corr(Geo g, Box b, const real *__restrict__ vx, const real *__restrict__ vy,
                             const real *__restrict__ vz, const real *__restrict__ vx1,
                             const real *__restrict__ vy1, const real *__restrict__ vz1,
                             const real *__restrict__ al1, const real *__restrict__ al2,
                             const real *__restrict__ al3, const real *__restrict__ pp,
                             real *__restrict__ energy, real *__restrict__ rho, real *__restrict__ visc_dep,
                             real *__restrict__ vxo, real *__restrict__ vyo, real *__restrict__ vzo,
                             real *__restrict__ cool_v, double alpha_av, double *__restrict__ partial, double dt) {
  // cool_v: cool_term_v of the cooling_term (lagran.F90:464-476), unused
  // without COOL
  // cells 1..n, marching in z: the node values (v, v1) and al1/al2 of plane
  // k are reused as plane k-1 / k values of the next cell.  The new
  // velocity v = 2 v1 - v of node (i, j, k), the cell's upper corner, goes
  // to vxo..vzo, since neighbouring cells still read the old v.
  PLANE_POS(b, i, j)
  ZRANGE(b, k0, k1)
  double v[1] = {0.0};
  if (j <= b.y1) {
    const int Y = g.sy, Z = g.sz;
    const int cb = g.idx(i, j, k0) - Z;
    Node4 lx = node4(vx, cb, Y), ly = node4(vy, cb, Y), lz = node4(vz, cb, Y);
    Node4 lx1 = node4(vx1, cb, Y), ly1 = node4(vy1, cb, Y), lz1 = node4(vz1, cb, Y);
    Al12 alo = al12(al1, al2, cb + Z, Y);
    for (int k = k0; k <= k1; ++k) {
      const int c = g.idx(i, j, k);
      const Node4 hx = node4(vx, c, Y), hy = node4(vy, c, Y), hz = node4(vz, c, Y);
      const Node4 hx1 = node4(vx1, c, Y), hy1 = node4(vy1, c, Y), hz1 = node4(vz1, c, Y);
      const Al12 ahi = al12(al1, al2, c + Z, Y);
      acc aa[12];
      // (v(first) - v(second)) . (v1(first) - v1(second)) per edge
#define AAS(q, P1, N1, P2, N2)                                                            \
  aa[q] = (P1##x.N1 - P2##x.N2) * (P1##x1.N1 - P2##x1.N2) + (P1##y.N1 - P2##y.N2) * (P1##y1.N1 - P2##y1.N2) + \
          (P1##z.N1 - P2##z.N2) * (P1##z1.N1 - P2##z1.N2);
      EDGES(AAS)
#undef AAS
      const acc a30 = al3[c], a3x = al3[c - 1], a3xy = al3[c - 1 - Y], a3y = al3[c - Y];
      const acc vh = max0(edge_heat_v((acc)g.dxb[i], (acc)g.dyb[j], (acc)g.dzb[k], (acc)g.cv(i, j, k), alo, ahi, a30,
                                      a3x, a3xy, a3y, aa));
      // face-averaged v1 (quad sums in the order c, c-Y, c-Z, c-Y-Z etc.;
      // node C = c, B = c-Y, D = c-X, A = c-X-Y)
      const acc q4 = 0.25;
      const acc vxb = (hx1.c + hx1.b + lx1.c + lx1.b) * q4;
      const acc vxbm = (hx1.d + hx1.a + lx1.d + lx1.a) * q4;
      const acc vyb = (hy1.c + hy1.d + ly1.c + ly1.d) * q4;
      const acc vybm = (hy1.b + hy1.a + ly1.b + ly1.a) * q4;
      const acc vzb = (hz1.c + hz1.d + hz1.b + hz1.a) * q4;
      const acc vzbm = (lz1.c + lz1.d + lz1.b + lz1.a) * q4;
      const acc dta = (acc)dt;
      const acc dv =
          ((vxb - vxbm) / (acc)g.dxb[i] + (vyb - vybm) / (acc)g.dyb[j] + (vzb - vzbm) / (acc)g.dzb[k]) * dta;
      const acc r = rho[c];
      const real e1 = energy[c] + (dta * vh - dv * pp[c]) / r;
      visc_dep[c] = visc_dep[c] + dta * vh;
      const real rn = r / ((acc)1.0 + dv);
      rho[c] = rn;
      if (COOL) {
        // a running average of the viscous heating is removed again, and
        // energy = MAX(energy, 0)
        const acc aav = (acc)alpha_av;
        const real ct = aav * dta * vh / rn + ((acc)1.0 - aav) * cool_v[c];
        cool_v[c] = ct;
        energy[c] = max0((acc)(e1 - ct));
      } else {
        energy[c] = e1;
      }
      v[0] += dt * (double)vh * g.cv(i, j, k);
      // v = 2 v1 - v (:399-407) at node (i, j, k); k_vupdate_low does the
      // nodes with an index 0
      vxo[c] = (real)((acc)2.0 * hx1.c - hx.c);
      vyo[c] = (real)((acc)2.0 * hy1.c - hy.c);
      vzo[c] = (real)((acc)2.0 * hz1.c - hz.c);
      lx = hx;
      ly = hy;
      lz = hz;
      lx1 = hx1;
      ly1 = hy1;
      lz1 = hz1;
      alo = ahi;
    }
  }
  const int ops[1] = {1};
  block_reduce_store<1>(v, ops, partial);
}


// v = 2 v1 - v (:399-407) into vxo..vzo at the nodes 0..n that have an
// index 0 (k_shock_corr does the rest): the plane k = 0 and, above it, the
// faces i = 0 and j = 0.  One thread per node of the (n+1)^2 plane, then
// one per node of the two faces, flattened.
__global__ __launch_bounds__(256) void k_vupdate_low(Geo g, const real *__restrict__ vx, const real *__restrict__ vy,
                                                     const real *__restrict__ vz, const real *__restrict__ vx1,
                                                     const real *__restrict__ vy1, const real *__restrict__ vz1,
                                                     real *__restrict__ vxo, real *__restrict__ vyo,
                                                     real *__restrict__ vzo) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long px = g.nx + 1, py = g.ny + 1;
  const long plane = px * py;                 // k = 0
  const long face = (long)g.nz * (px + py - 1);  // k = 1..n: i = 0 or j = 0
  if (t >= plane + face) return;
  int i, j, k;
  if (t < plane) {
    k = 0;
    j = (int)(t / px);
    i = (int)(t - (long)j * px);
  } else {
    const long u = t - plane;
    const long row = px + py - 1;  // (0, j) for j = 0..ny, then (i, 0) for i = 1..nx
    k = 1 + (int)(u / row);
    const int r = (int)(u - (long)(k - 1) * row);
    if (r < py) {
      i = 0;
      j = r;
    } else {
      i = r - (int)py + 1;
      j = 0;
    }
  }
  const int c = g.idx(i, j, k);
  vxo[c] = (real)(2.0 * vx1[c] - vx[c]);
  vyo[c] = (real)(2.0 * vy1[c] - vy[c]);
  vzo[c] = (real)(2.0 * vz1[c] - vz[c]);
}

// Element offsets of the fields energy_account reads, per staggering.
// PadLayout: the library's padded box (one index expression for all kinds).
struct PadLayout {
  Geo g;
  __device__ __forceinline__ int cell(int i, int j, int k) const { return g.idx(i, j, k); }
  __device__ __forceinline__ int node(int i, int j, int k) const { return g.idx(i, j, k); }
  __device__ __forceinline__ int fx(int i, int j, int k) const { return g.idx(i, j, k); }
  __device__ __forceinline__ int fy(int i, int j, int k) const { return g.idx(i, j, k); }
  __device__ __forceinline__ int fz(int i, int j, int k) const { return g.idx(i, j, k); }
};
// FtnLayout: the caller's Fortran arrays with the bounds of
// mpi_routines.F90:295-334 -- cell (-1:n+2)^3, node (-2:n+2)^3,
// bx (-2:nx+2, -1:ny+2, -1:nz+2) and likewise by, bz.
struct FtnLayout {
  int nx, ny, nz;
  __device__ __forceinline__ int cell(int i, int j, int k) const {
    return (i + 1) + (nx + 4) * ((j + 1) + (ny + 4) * (k + 1));
  }
  __device__ __forceinline__ int node(int i, int j, int k) const {
    return (i + 2) + (nx + 5) * ((j + 2) + (ny + 5) * (k + 2));
  }
  __device__ __forceinline__ int fx(int i, int j, int k) const {
    return (i + 2) + (nx + 5) * ((j + 1) + (ny + 4) * (k + 1));
  }
  __device__ __forceinline__ int fy(int i, int j, int k) const {
    return (i + 1) + (nx + 4) * ((j + 2) + (ny + 5) * (k + 1));
  }
  __device__ __forceinline__ int fz(int i, int j, int k) const {
    return (i + 1) + (nx + 4) * ((j + 1) + (ny + 4) * (k + 2));
  }
};

// energy_account (diagnostics.F90:517-609) plus max |div B| dx: per-block
// partials of (en_b, en_ke, en_int, -|div B| dx) over nodes 0..n; the cell
// terms on cells 1..n.  The layout L says where each staggering keeps
// element (i, j, k); the arithmetic does not depend on it.
template <class L, class R>
__global__ __launch_bounds__(BX * BY) void k_energy(Geo g, L lay, const R *__restrict__ rho,
                                                    const R *__restrict__ energy, const R *__restrict__ vx,
                                                    const R *__restrict__ vy, const R *__restrict__ vz,
                                                    const double *__restrict__ bx, const double *__restrict__ by,
                                                    const double *__restrict__ bz, double *__restrict__ partial) {
  const Box b = {0, g.nx, 0, g.ny, 0, g.nz};
  PLANE_POS(b, i, j)
  const int k = (int)blockIdx.z;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (j <= g.ny && k <= g.nz) {
    if (i >= 1 && j >= 1 && k >= 1) {
      const double bx0 = bx[lay.fx(i, j, k)], bxm = bx[lay.fx(i - 1, j, k)];
      const double by0 = by[lay.fy(i, j, k)], bym = by[lay.fy(i, j - 1, k)];
      const double bz0 = bz[lay.fz(i, j, k)], bzm = bz[lay.fz(i, j, k - 1)];
      double w1 = (bx0 * bx0 + bxm * bxm) * 0.5;
      const double w2 = (by0 * by0 + bym * bym) * 0.5;
      const double w3 = (bz0 * bz0 + bzm * bzm) * 0.5;
      w1 = (w1 + w2 + w3) * 0.5;
      const double cv = g.cv(i, j, k);
      const int c = lay.cell(i, j, k);
      v[0] = w1 * cv;
      v[2] = energy[c] * rho[c] * cv;
      const double dx = g.dxb[i], dy = g.dyb[j], dz = g.dzb[k];
      const double d = (bx0 - bxm) / dx + (by0 - bym) / dy + (bz0 - bzm) / dz;
      v[3] = -(fabs(d) * dmin(dmin(dx, dy), dz));
    }
    double cvs[8];
    const int di[8] = {0, 1, 0, 1, 0, 1, 0, 1}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1}, dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
#pragma unroll
    for (int q = 0; q < 8; ++q) cvs[q] = g.cv(i + di[q], j + dj[q], k + dk[q]);
    double rv = rho[lay.cell(i, j, k)] * cvs[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) rv = rv + rho[lay.cell(i + di[q], j + dj[q], k + dk[q])] * cvs[q];
    double cvv = cvs[0];
#pragma unroll
    for (int q = 1; q < 8; ++q) cvv = cvv + cvs[q];
    rv = rv / cvv;
    cvv = cvv * 0.125;
    const int n = lay.node(i, j, k);
    double w1 = rv * cvv * (vx[n] * vx[n] + vy[n] * vy[n] + vz[n] * vz[n]);
    if (i == 0 || i == g.nx) w1 = w1 * 0.5;
    if (j == 0 || j == g.ny) w1 = w1 * 0.5;
    if (k == 0 || k == g.nz) w1 = w1 * 0.5;
    v[1] = w1 * 0.5;
  }
  const int ops[4] = {1, 1, 1, 0};
  block_reduce_store<4>(v, ops, partial);
}
