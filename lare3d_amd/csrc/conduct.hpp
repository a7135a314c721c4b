// conduct.hpp -- RKL2 super-stepped Braginskii conduction (conduct.f90).
//
// Device form of heat_conduct_sts2 (conduct.f90:299-416):
//  * heat_flux (conduct.f90:152-291) scatters each face flux into both
//    neighbouring cells; here every thread gathers the six face fluxes of
//    its own cell and sums them in the reference's accumulation order
//    ((((fz_in + fy_in) + fx_in) - fx_out) - fy_out) - fz_out, so no
//    atomics are needed and the result is the scatter's sum.
//  * temperature is never materialised: each stage derives T from the
//    stage input on the fly.  Ghost temperatures follow the reference's
//    boundary behaviour without a halo pass: on BC_USER axes the mirror of
//    temperature_bcs is an index map; on periodic axes stage 2 reads the
//    pre-conduction energy ghosts (temperature_bcs exchanges rho, not T,
//    boundary.f90:214, and only the interior of Y(2) was refreshed,
//    conduct.f90:366), while stages >= 3 read the periodic image (the
//    energy_bcs at conduct.f90:403).
//  * the Y(0:3) buffer copies of conduct.f90:400-406 become a rotation of
//    three device buffers.

#define MIN_B 1.0e-10

// Ghost-read modes of the stage kernels.
#define TMODE_DIRECT 0   // read the array as stored (stage 1: Y0 = energy)
#define TMODE_STALE 1    // stage 2: user axes mirrored, periodic ghosts from E0
#define TMODE_WRAP 2     // stages >= 3: user axes mirrored, periodic axes wrapped

__device__ __forceinline__ int mirror_idx(int c, int n) {
  return c < 1 ? 1 - c : 2 * n + 1 - c;  // 0->1, -1->2, n+1->n, n+2->n-1
}

// Energy-like value at (i,j,k) under a ghost-read mode.  A BC_OPEN side
// is never filled (energy_bcs and temperature_bcs leave it alone), so its
// ghosts keep the values stored before the super-step: E0 at the same
// index, in every mode.
__device__ __forceinline__ double ghost_read(const Geo &g, const real *__restrict__ Y,
                                             const real *__restrict__ E0, int i, int j, int k, int mode) {
  if (mode == TMODE_DIRECT) return Y[g.idx(i, j, k)];
  int c[3] = {i, j, k};
  bool stale = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int n = g.n(a);
    if (c[a] < 1 || c[a] > n) {
      if (c[a] < 1 ? g.open_lo[a] : g.open_hi[a]) stale = true;
      else if (g.user[a]) c[a] = mirror_idx(c[a], n);
      else if (mode == TMODE_STALE) stale = true;
      else c[a] = c[a] < 1 ? c[a] + n : c[a] - n;
    }
  }
  return (stale ? E0 : Y)[g.idx(c[0], c[1], c[2])];
}

// temperature**2.5 and **1.5 (conduct.f90:27, :83) correctly rounded: t^2
// and t*sqrt(t) carried in double-double (FMA-exact products, sqrt(t) =
// r + (t - r^2) / (2r)), rounded once.  The oracle computes the same
// operations (a libm pow is not correctly rounded for every argument).
__device__ __forceinline__ double pow25(double t) {
  if (!(t > 0.0) || t == INFINITY) return t == 0.0 ? 0.0 : (t > 0.0 ? t : __builtin_nan(""));
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
__device__ __forceinline__ double pow15(double t) {
  if (!(t > 0.0) || t == INFINITY) return t == 0.0 ? 0.0 : (t > 0.0 ? t : __builtin_nan(""));
  const double r = sqrt(t);
  const double e = fma(-r, r, t);
  const double rl = e / (2.0 * r);
  const double p = t * r;
  if (!(p < INFINITY)) return p;
  const double pl = fma(t, r, -p);
  return p + (pl + t * rl);
}
// The same powers in FP32 (FP32 arithmetic mode): a few ulp, with t^1.5
// formed first so that t^2.5 overflows only where the result does.
__device__ __forceinline__ float pow25(float t) { return t * sqrtf(t) * t; }
__device__ __forceinline__ float pow15(float t) { return t * sqrtf(t); }

struct CondArgs {
  double kappa_0, tfac, dt;
  // partially ionised: T = gm1 / (2 - xi_n) (e - (1 - xi_n) ionise_pot) per
  // cell (conduct.f90:72-73); otherwise T = tfac e with the uniform xi_n
  const real *xi;
  double gm1, ip;
};

// temperature of the energy-like value e of cell c
template <bool PI>
__device__ __forceinline__ acc cond_temp(const CondArgs &a, acc e, int c) {
  if constexpr (PI) {
    const acc x = (acc)a.xi[c];
    return (acc)a.gm1 / ((acc)2.0 - x) * (e - ((acc)1.0 - x) * (acc)a.ip);
  } else {
    return (acc)a.tfac * e;
  }
}

// Larsen flux-limiter factor over cells 0..n and the explicit parabolic
// time-step minimum over 1..n (calc_s_stages, conduct.f90:69-129).
// Temperatures from energy as stored (ghosts valid at call time).  PI:
// per-cell temperatures of the partially ionised gas.
template <bool PI>
__global__ __launch_bounds__(BX * BY) void k_larsen(Geo g, const real *__restrict__ energy, const real *__restrict__ rho,
                         real *__restrict__ larsen, double *__restrict__ partial, double flux_limiter,
                         double kappa_0, CondArgs ca, int limiter_on) {
  const double tfac = ca.tfac;
  const Box b = {0, g.nx, 0, g.ny, 0, g.nz};
  PLANE_POS(b, i, j)
  const int k = (int)blockIdx.z;
  double v[1] = {1.0e300};
  const bool in = j <= g.ny && k <= g.nz;
  if (in) {
    const int c = g.idx(i, j, k);
    auto T_at = [&](int q) -> double {
      if constexpr (PI) return (double)cond_temp<true>(ca, (acc)energy[q], q);
      else return tfac * energy[q];
    };
    const double T = T_at(c);
    double L = 1.0;
    if (limiter_on) {
      const double temp = pow25(T);
      const double q_fs = flux_limiter * 42.85 * rho[c] * pow15(T);
      const double q_fs2 = q_fs * q_fs;
      const double q_spx = -kappa_0 * temp * (T_at(c + 1) - T_at(c - 1)) * 0.5 / g.dxb[i];
      const double q_spy = -kappa_0 * temp * (T_at(c + g.sy) - T_at(c - g.sy)) * 0.5 / g.dyb[j];
      const double q_spz = -kappa_0 * temp * (T_at(c + g.sz) - T_at(c - g.sz)) * 0.5 / g.dzb[k];
      const double q_sp2 = q_spx * q_spx + q_spy * q_spy + q_spz * q_spz;
      L = q_fs / sqrt(q_fs2 + q_sp2);
    }
    larsen[c] = L;
    if (i >= 1 && j >= 1 && k >= 1) {
      const double kappa1 = kappa_0 * L;
      const double tp = pow25(T);
      const double r = rho[c];
      const double dx = g.dxb[i], dy = g.dyb[j], dz = g.dzb[k];
      const double dt1 = r * (dx * dx) / (2.0 * kappa1 * tp);
      const double dt2 = r * (dy * dy) / (2.0 * kappa1 * tp);
      const double dt3 = r * (dz * dz) / (2.0 * kappa1 * tp);
      v[0] = dmin(dt1, dmin(dt2, dt3));
    }
  }
  const int ops[1] = {0};
  block_reduce_store<1>(v, ops, partial);
}

// Face flux fc_sp of conduct.f90:176-207 (x face between cells ix, ix+1),
// given the temperature neighbourhood T[a][b][c] with a,b,c in {0,1,2}
// meaning offsets -1,0,+1 around the *lower* cell... see callers.
// The three face kinds are written out separately, in the reference's
// operation order.

struct Tn {  // 3x3x3 temperature neighbourhood of a cell (corners unused)
  acc t[3][3][3];
};

// x face between (ox,oy,oz)+(0,0,0) and +(1,0,0) relative to centre index
// (1,1,1) of the neighbourhood, lower cell at neighbourhood x index lx.
__device__ __forceinline__ acc flux_x_face(const Geo &g, const Tn &T, int lx, int ix, int iy, int iz,
                                              const double *__restrict__ bx, const double *__restrict__ by,
                                              const double *__restrict__ bz, acc L, acc kappa_0) {
  const int c = g.idx(ix, iy, iz);
  const acc byf = (acc)(0.25 * (by[c] + by[c + 1] + by[c - g.sy] + by[c + 1 - g.sy]));
  const acc bzf = (acc)(0.25 * (bz[c] + bz[c + 1] + bz[c - g.sz] + bz[c + 1 - g.sz]));
  const acc bxc = (acc)bx[c];
  const acc modb = bxc * bxc + byf * byf + bzf * bzf + (acc)MIN_B;
  acc tb = (T.t[lx + 1][1][1] + T.t[lx][1][1]) / (acc)2.0;
  acc tb_p = (T.t[lx + 1][2][1] + T.t[lx][2][1]) / (acc)2.0;
  acc tb_m = (T.t[lx + 1][0][1] + T.t[lx][0][1]) / (acc)2.0;
  const acc tg_a1 = (tb_p - tb_m) / (acc)(g.dyc[iy] + g.dyc[iy - 1]);
  tb_p = (T.t[lx + 1][1][2] + T.t[lx][1][2]) / (acc)2.0;
  tb_m = (T.t[lx + 1][1][0] + T.t[lx][1][0]) / (acc)2.0;
  const acc tg_a2 = (tb_p - tb_m) / (acc)(g.dzc[iz] + g.dzc[iz - 1]);
  const acc tg = (T.t[lx + 1][1][1] - T.t[lx][1][1]) / (acc)g.dxc[ix];
  tb = dmax(tb, (acc)0.0);
  return -L * kappa_0 * pow25(tb) * (bxc * (tg * bxc + tg_a1 * byf + tg_a2 * bzf) + tg * (acc)MIN_B) / modb;
}

__device__ __forceinline__ acc flux_y_face(const Geo &g, const Tn &T, int ly, int ix, int iy, int iz,
                                              const double *__restrict__ bx, const double *__restrict__ by,
                                              const double *__restrict__ bz, acc L, acc kappa_0) {
  const int c = g.idx(ix, iy, iz);
  const acc bxf = (acc)(0.25 * (bx[c] + bx[c + g.sy] + bx[c - 1] + bx[c - 1 + g.sy]));
  const acc bzf = (acc)(0.25 * (bz[c] + bz[c + g.sy] + bz[c - g.sz] + bz[c + g.sy - g.sz]));
  const acc byc = (acc)by[c];
  const acc modb = bxf * bxf + byc * byc + bzf * bzf + (acc)MIN_B;
  acc tb = (T.t[1][ly + 1][1] + T.t[1][ly][1]) / (acc)2.0;
  acc tb_p = (T.t[2][ly][1] + T.t[2][ly + 1][1]) / (acc)2.0;
  acc tb_m = (T.t[0][ly + 1][1] + T.t[0][ly][1]) / (acc)2.0;
  const acc tg_a1 = (tb_p - tb_m) / (acc)(g.dxc[ix] + g.dxc[ix - 1]);
  tb_p = (T.t[1][ly + 1][2] + T.t[1][ly][2]) / (acc)2.0;
  tb_m = (T.t[1][ly + 1][0] + T.t[1][ly][0]) / (acc)2.0;
  const acc tg_a2 = (tb_p - tb_m) / (acc)(g.dzc[iz] + g.dzc[iz - 1]);
  const acc tg = (T.t[1][ly + 1][1] - T.t[1][ly][1]) / (acc)g.dyc[iy];
  tb = dmax(tb, (acc)0.0);
  return -L * kappa_0 * pow25(tb) * (byc * (tg * byc + tg_a1 * bxf + tg_a2 * bzf) + tg * (acc)MIN_B) / modb;
}

__device__ __forceinline__ acc flux_z_face(const Geo &g, const Tn &T, int lz, int ix, int iy, int iz,
                                              const double *__restrict__ bx, const double *__restrict__ by,
                                              const double *__restrict__ bz, acc L, acc kappa_0) {
  const int c = g.idx(ix, iy, iz);
  const acc bxf = (acc)(0.25 * (bx[c] + bx[c + g.sz] + bx[c - 1] + bx[c - 1 + g.sz]));
  const acc byf = (acc)(0.25 * (by[c] + by[c + g.sz] + by[c - g.sy] + by[c - g.sy + g.sz]));
  const acc bzc = (acc)bz[c];
  const acc modb = bxf * bxf + byf * byf + bzc * bzc + (acc)MIN_B;
  acc tb = (T.t[1][1][lz + 1] + T.t[1][1][lz]) / (acc)2.0;
  acc tb_p = (T.t[2][1][lz + 1] + T.t[2][1][lz]) / (acc)2.0;
  acc tb_m = (T.t[0][1][lz + 1] + T.t[0][1][lz]) / (acc)2.0;
  const acc tg_a1 = (tb_p - tb_m) / (acc)(g.dxc[ix] + g.dxc[ix - 1]);
  tb_p = (T.t[1][2][lz + 1] + T.t[1][2][lz]) / (acc)2.0;
  tb_m = (T.t[1][0][lz + 1] + T.t[1][0][lz]) / (acc)2.0;
  const acc tg_a2 = (tb_p - tb_m) / (acc)(g.dyc[iy] + g.dyc[iy - 1]);
  const acc tg = (T.t[1][1][lz + 1] - T.t[1][1][lz]) / (acc)g.dzc[iz];
  tb = dmax(tb, (acc)0.0);
  return -L * kappa_0 * pow25(tb) * (bzc * (tg * bzc + tg_a1 * bxf + tg_a2 * byf) + tg * (acc)MIN_B) / modb;
}

// Ghost mapping of one coordinate under a read mode (the per-axis part of
// ghost_read): the coordinate to read and whether to read E0 instead.
// OPEN: some side is BC_OPEN (its ghosts read E0 as stored).
template <int MODE, bool OPEN>
__device__ __forceinline__ void map_coord(int c, int n, int user, int olo, int ohi, int &m, bool &stale) {
  m = c;
  stale = false;
  if (MODE == TMODE_DIRECT || (c >= 1 && c <= n)) return;
  if (OPEN && (c < 1 ? olo : ohi)) stale = true;
  else if (user) m = mirror_idx(c, n);
  else if (MODE == TMODE_STALE) stale = true;
  else m = c < 1 ? c + n : c - n;
}

// RKL2 stages marching in z (conduct.f90:345-396; FIRST = stage 1).  T of
// planes k-1, k, k+1 (3x3 in-plane each) stays in registers, so each plane
// is read once per column; the ghost mapping of a lane's x and y neighbours
// is worked out once.  The z face shared with the cell below is the
// previous plane's out face, and the x face shared with the left neighbour
// is that lane's out face: waves cover runs of 64 points of the plane
// flattened over x = x0-1..x1, overlapping by one, and lanes at x0-1 only
// provide their out face.  Each shared flux is the same expression on the
// same values as the neighbour's own, and the six are summed in the order
// of cell_heat_flux.
#define RKL_OUT 63
template <int MODE, int FIRST, bool PI, bool OPEN>
__global__ __launch_bounds__(BX * BY) void k_rkl2_z(Geo g, Box b, const real *__restrict__ Y2, const real *__restrict__ Y1,
                                                    const real *__restrict__ E0, const real *__restrict__ rho,
                                                    const double *__restrict__ bx, const double *__restrict__ by,
                                                    const double *__restrict__ bz, const real *__restrict__ larsen,
                                                    real *__restrict__ Lc, real *__restrict__ Yout, CondArgs a,
                                                    double mu, double nu, double one_m_mu_nu, double mu_t,
                                                    double gam_t) {
  const int lane = (int)threadIdx.x;  // BX = 64: threadIdx.y is the wave
  const int wq = b.x1 - b.x0 + 2;     // x0-1 .. x1
  const int npts = wq * (b.y1 - b.y0 + 1);
  const int f = (int)(xcd_tile(blockIdx.x, gridDim.x) * BY + threadIdx.y) * RKL_OUT + lane - 1;
  const bool valid = f >= 0 && f < npts;
  const int ff = valid ? f : 0;  // lanes past the plane read point 0, unused
  const int jr = ff / wq;
  const int i = b.x0 - 1 + (ff - jr * wq);
  const int j = b.y0 + jr;
  const bool out = valid && lane >= 1 && i >= b.x0;
  ZRANGE(b, k0, k1)
  int xm[3], ym[3];
  bool xs[3], ys[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    map_coord<MODE, OPEN>(i + q - 1, g.nx, g.user[0], g.open_lo[0], g.open_hi[0], xm[q], xs[q]);
    map_coord<MODE, OPEN>(j + q - 1, g.ny, g.user[1], g.open_lo[1], g.open_hi[1], ym[q], ys[q]);
  }
  const acc kap = (acc)a.kappa_0, tdt = (acc)a.dt;
  // T.t[p][q][r]: x offset p-1, y offset q-1, plane k-1+r
  Tn T;
#define RKL_LOAD_PLANE(kk, r)                                                    \
  {                                                                              \
    int zm;                                                                      \
    bool zs;                                                                     \
    map_coord<MODE, OPEN>((kk), g.nz, g.user[2], g.open_lo[2], g.open_hi[2], zm, zs); \
    _Pragma("unroll") for (int p = 0; p < 3; ++p)                                \
      _Pragma("unroll") for (int q = 0; q < 3; ++q) {                            \
        const int cq = g.idx(xm[p], ym[q], zm);                                  \
        T.t[p][q][r] = cond_temp<PI>(a, (acc)(((xs[p] || ys[q] || zs) ? E0 : Y2)[cq]), cq); \
      }                                                                          \
  }
  RKL_LOAD_PLANE(k0 - 1, 0)
  RKL_LOAD_PLANE(k0, 1)
  acc fz_in = 0.0;
  for (int k = k0; k <= k1; ++k) {
    RKL_LOAD_PLANE(k + 1, 2)
    const int c0 = g.idx(i, j, k);
    const acc L0 = larsen[c0];
    const acc fx_out = flux_x_face(g, T, 1, i, j, k, bx, by, bz, L0, kap);
    const acc fx_in = __shfl_up(fx_out, 1, 64);
    if (k == k0) fz_in = flux_z_face(g, T, 0, i, j, k - 1, bx, by, bz, (acc)larsen[c0 - g.sz], kap);
    const acc fy_in = flux_y_face(g, T, 0, i, j - 1, k, bx, by, bz, (acc)larsen[c0 - g.sy], kap);
    const acc fy_out = flux_y_face(g, T, 1, i, j, k, bx, by, bz, L0, kap);
    const acc fz_out = flux_z_face(g, T, 1, i, j, k, bx, by, bz, L0, kap);
    if (out) {
      acc fl = (acc)0.0 + fz_in / (acc)g.dzb[k - 1];
      fl = fl + fy_in / (acc)g.dyb[j - 1];
      fl = fl + fx_in / (acc)g.dxb[i - 1];
      fl = fl - fx_out / (acc)g.dxb[i];
      fl = fl - fy_out / (acc)g.dyb[j];
      fl = fl - fz_out / (acc)g.dzb[k];
      if (FIRST) {
        const acc lc = ((acc)1.0 / rho[c0]) * fl;
        Lc[c0] = lc;
        const acc cc = (acc)mu_t * tdt * lc;
        Yout[c0] = E0[c0] + cc;
      } else {
        const acc cc0 = (acc)gam_t * tdt * Lc[c0];
        const acc lj = ((acc)1.0 / rho[c0]) * fl;
        const acc cc1 = (acc)mu_t * tdt * lj;
        acc y3;
        if constexpr (sizeof(acc) == sizeof(double)) {
          y3 = (acc)mu * Y2[c0] + (acc)nu * Y1[c0] + (acc)one_m_mu_nu * E0[c0];
          y3 = y3 + cc1 + cc0;
        } else {
          // FP32 arithmetic: the same combination in increment form,
          // w E0 + mu (Y2 - E0) + nu (Y1 - E0) with w = mu + nu + (1 - mu -
          // nu), so that rounding scales with the stage increments, not
          // with E0.  w rounds to exactly 1 for an RKL2 stage; the
          // trapezoid evaluation (all three weights zero) has w = 0.
          const acc e0 = E0[c0];
          const acc w = (acc)(mu + nu + one_m_mu_nu);
          y3 = w * e0 + (((acc)mu * (Y2[c0] - e0) + (acc)nu * (Y1[c0] - e0)) + (cc1 + cc0));
        }
        Yout[c0] = y3;
      }
    }
    fz_in = fz_out;
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        T.t[p][q][0] = T.t[p][q][1];
        T.t[p][q][1] = T.t[p][q][2];
      }
  }
#undef RKL_LOAD_PLANE
}

// Final write of conduct_heat (conduct.f90:409-410): energy interior from
// the last stage, ghosts by energy_bcs (index map).  delta_mode != 0 applies
// lagrangian_step's energy0 + (energy - energy0) and MAX(., 0)
// (lagran.F90:168-181) over the whole cell box.
__global__ __launch_bounds__(BX * BY) void k_rkl2_final(Geo g, Box b, const real *__restrict__ Ys, real *__restrict__ energy,
                             int delta_mode) {
  BOX_INDEX(b, i, j, k);
  const double v = ghost_read(g, Ys, energy, i, j, k, TMODE_WRAP);
  const int c = g.idx(i, j, k);
  if (delta_mode) {
    const double e0 = energy[c];
    const double d = v - e0;
    energy[c] = max0(e0 + d);
  } else {
    energy[c] = v;
  }
}
