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

#define SIXTH (1.0 / 6.0)

template <int AX>
__device__ __forceinline__ int sstride(const Geo &g) {
  return AX == 0 ? 1 : (AX == 1 ? g.sy : g.sz);
}
template <int AX>
__device__ __forceinline__ int coord(int i, int j, int k) {
  return AX == 0 ? i : (AX == 1 ? j : k);
}

// Van Leer limited slope Di (xremap.f90:390-396) in the arithmetic type A.
// The gradients dfi and dfu come in FP64, and the limiter's sign tests and
// three-way minimum are made in FP64, so that near-zero gradients compare
// exactly; the slope estimate Da (two divisions) is formed in A.
template <class A>
__device__ __forceinline__ A vl_di(double dfi, double dfu, double sign_v, A dci, A dcu, A dbu, A phi) {
  A Da = ((A)2.0 - phi) * (A)fabs(dfi) / dci + ((A)1.0 + phi) * (A)fabs(dfu) / dcu;
  Da = Da * (A)SIXTH;
  const double ss = 0.5 * (copysign(1.0, dfi) + copysign(1.0, dfu));
  return (A)(sign_v * ss * dmin(dmin((double)fabs(Da) * (double)dbu, fabs(dfi)), fabs(dfu)));
}

// Inputs of the predictor energy and pressure (lagran.F90:234-249), which
// the predictor remap's last geometry pass (CV = 1) forms on cells 0..n+1
// from the cv1 it has just formed (p0 is the start-of-step pressure).
struct PredP {
  const real *energy, *rho, *visc_heat, *p0;
  real *pp;
  double gamma, dt2;
  EosArgs eq;  // partially ionised pressure when eq.xi is set
};

// Lagrangian geometry of a sweep (xremap.f90:57-105 and analogues).
// Velocities are the time-centred ones (v in predictor mode).  The node
// velocities of planes k and k-1 are kept in registers while marching in z.
// CV selects the outputs: 0 = d?b1 only (which needs only the sweep
// component of the velocity), 1 = cv1 and d?b1 (and the predictor pressure
// from pr), 2 = cv1, cv2 and d?b1.
template <int AX, int CV>
__global__ __launch_bounds__(BX * BY) void k_geom(Geo g, Box b, const real *__restrict__ ux, const real *__restrict__ uy,
                       const real *__restrict__ uz, real *__restrict__ cv1, real *__restrict__ cv2,
                       real *__restrict__ db1, double dt, double xpass, double ypass, double zpass, PredP pr) {
  ZCOLUMN(b, i, j, k0, k1);
  constexpr bool NX = CV > 0 || AX == 0, NY = CV > 0 || AX == 1, NZ = CV > 0 || AX == 2;
  const int Y = g.sy;
  const int cb = g.idx(i, j, k0 - 1);
  Quad ax = {}, ay = {}, az = {};  // plane k-1
  if (NX) ax = quad_back(ux, cb, Y);
  if (NY) ay = quad_back(uy, cb, Y);
  if (NZ) az = quad_back(uz, cb, Y);
  for (int k = k0; k <= k1; ++k) {
    const int c = g.idx(i, j, k);
    Quad bx_ = {}, by_ = {}, bz_ = {};  // plane k
    if (NX) bx_ = quad_back(ux, c, Y);
    if (NY) by_ = quad_back(uy, c, Y);
    if (NZ) bz_ = quad_back(uz, c, Y);
    // sums in the order of u(c) + u(c-Y) + u(c-Z) + u(c-Y-Z) etc.
    const acc q4 = 0.25, tdt = (acc)dt;
    const acc vxb = (bx_.p0 + bx_.py + ax.p0 + ax.py) * q4;
    const acc vxbm = (bx_.px + bx_.pxy + ax.px + ax.pxy) * q4;
    const acc vyb = (by_.p0 + by_.px + ay.p0 + ay.px) * q4;
    const acc vybm = (by_.py + by_.pxy + ay.py + ay.pxy) * q4;
    const acc vzb = (bz_.p0 + bz_.px + bz_.py + bz_.pxy) * q4;
    const acc vzbm = (az.p0 + az.px + az.py + az.pxy) * q4;
    const acc dxb = (acc)g.dxb[i], dyb = (acc)g.dyb[j], dzb = (acc)g.dzb[k];
    acc d1;
    if (AX == 0) d1 = dxb + (vxb - vxbm) * tdt;
    else if (AX == 1) d1 = dyb + (vyb - vybm) * tdt;
    else d1 = dzb + (vzb - vzbm) * tdt;
    if (CV > 0) {
      const acc cv = (acc)g.cv(i, j, k);
      acc dv1, dv2;
      if (AX == 0) {
        const acc tA = (acc)ypass * (vyb - vybm) / dyb;
        const acc tB = (acc)zpass * (vzb - vzbm) / dzb;
        dv1 = (tA + tB + (vxb - vxbm) / dxb) * tdt;
        dv2 = (tA + tB) * tdt;
      } else if (AX == 1) {
        const acc tA = (acc)xpass * (vxb - vxbm) / dxb;
        const acc tB = (acc)zpass * (vzb - vzbm) / dzb;
        dv1 = (tA + tB + (vyb - vybm) / dyb) * tdt;
        dv2 = (tA + tB) * tdt;
      } else {
        const acc tA = (acc)xpass * (vxb - vxbm) / dxb;
        const acc tB = (acc)ypass * (vyb - vybm) / dyb;
        dv1 = (tA + tB + (vzb - vzbm) / dzb) * tdt;
        dv2 = (tA + tB) * tdt;
      }
      const real c1 = cv * ((acc)1.0 + dv1);
      cv1[c] = c1;
      if (CV > 1) cv2[c] = cv * ((acc)1.0 + dv2);
      if (CV == 1 && i >= 0 && i <= g.nx + 1 && j >= 0 && j <= g.ny + 1 && k >= 0 && k <= g.nz + 1) {
        // predictor energy and pressure (lagran.F90:234-249)
        const acc r = pr.rho[c];
        const acc dv = c1 / cv - (acc)1.0;
        acc e1 = pr.energy[c] - pr.p0[c] * dv / r;
        e1 = e1 + pr.visc_heat[c] * (acc)pr.dt2 / r;
        const acc ei = pr.eq.xi ? eos_e<true>(e1, pr.eq, c) : eos_e<false>(e1, pr.eq, c);
        pr.pp[c] = ei * ((acc)pr.gamma - (acc)1.0) * r * cv / c1;
      }
    }
    db1[c] = d1;
    ax = bx_;
    ay = by_;
    az = bz_;
  }
}

// dxc1 at sweep coordinate s (index c), clamped as xremap.f90:110
template <int AX>
__device__ __forceinline__ acc dc1(const Geo &g, const real *__restrict__ db1, int c, int s) {
  const int S = sstride<AX>(g);
  const int n = g.n(AX);
  return (acc)0.5 * (db1[c] + db1[s + 1 <= n + 2 ? c + S : c]);
}

// Van Leer CT flux (xremap.f90:340-470) from the quotients f* = field /
// width at s-1 (clamped), s, s+1, s+2 (clamped), the widths db, dbp at s and
// s+1 (the field's widths averaged along its geometry axis), db1 at s and
// s+1, and dxc1 at s-1, s, s+1.  The quotients, the upwind value and the
// flux itself are FP64; the Courant number phi and the limited slope are
// formed in the arithmetic type.
__device__ __forceinline__ double ct_flux_q(double v_advect, double fm, double fi, double fp, double fp2, acc db,
                                            acc dbp, acc db1c, acc db1p, acc dcm, acc dci, acc dcp, double dt) {
  const double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
  const double sign_v = copysign(1.0, v_advect);
  const double vad_p = (sign_v + 1.0) * 0.5;
  const double vad_m = 1.0 - vad_p;
  const double fu = fi * vad_p + fp * vad_m;
  const double dfu = dfm * vad_p + dfp * vad_m;
  const acc ap = (acc)vad_p, am = (acc)vad_m;
  const acc dcu = dcm * ap + dcp * am;
  const acc dbu = db1c * ap + db1p * am;
  const acc du = db * ap + dbp * am;
  const acc phi = (acc)fabs(v_advect) * (acc)dt / du;
  const acc Di = vl_di<acc>(dfi, dfu, sign_v, dci, dcu, dbu, phi);
  return (fu + (double)(Di * ((acc)1.0 - phi))) * v_advect * dt;
}

// The index range of face B component Q (bx(-2:nx+2, -1:ny+2, -1:nz+2) and
// its analogues, remap.f90:44-67): the points that the scaling to face
// fluxes and back touches.
template <int Q>
__device__ __forceinline__ bool bface_ok(int i, int j, int k) {
  return Q == 0 ? (j >= -1 && k >= -1) : (Q == 1 ? (i >= -1 && k >= -1) : (i >= -1 && j >= -1));
}
// The widths dxb, dyb, dzb at one point, held in registers (the fused
// kernels load the two that are fixed along their march once).  Face B
// component Q becomes the face flux that the remap moves by multiplying by
// the other two widths, and goes back by dividing by them, in the
// reference's operation order (remap.f90:46-48 and 104-106).
struct Wid {
  double w[3];
};
template <int Q>
__device__ __forceinline__ double to_flux_w(double v, const Wid &m) {
  if (Q == 0) return v * m.w[1] * m.w[2];
  if (Q == 1) return v * m.w[0] * m.w[2];
  return v * m.w[0] * m.w[1];
}
template <int Q>
__device__ __forceinline__ double from_flux_w(double v, const Wid &m) {
  if (Q == 0) return v / m.w[1] / m.w[2];
  if (Q == 1) return v / m.w[0] / m.w[2];
  return v / m.w[0] / m.w[1];
}

// The transverse components (t1, t2) of a sweep along AX and its CT update
// pairs: x -> (by, bz), y -> (bx, bz), z -> (bx, by).
template <int AX> struct CtComp {
  static constexpr int T1 = AX == 0 ? 1 : 0;
  static constexpr int T2 = AX == 2 ? 1 : 2;
};
// ---------------------------------------------------------------------------
// Fused CT sweeps: the fluxes and the three face updates of a sweep in one
// pass, out of place (inputs t1i, t2i, ai; outputs t1o, t2o, ao), so that
// the fluxes never go to memory.  Each update takes a flux at the point and
// at one neighbour (xremap.f90:121-157 and analogues):
//   t1 = t1 - F1 + F1(s-1), t2 = t2 - F2 + F2(s-1) along the sweep;
//   a = a + F1 - F1(c - AS1) + F2 - F2(c - AS2) across the transverse axes.
// A block is a tile of the plane across the march axis, rows of lanes along
// x, covering a chunk of CTS_CHUNK points along the march.  It first forms
// the fluxes of the chunk and of the point before it into LDS, then, after
// one barrier, writes its points, taking each flux and its neighbours along
// the march (the previous point), along x (the lane below) and along the
// row axis (the row below) from LDS.  Lane 0 of each row segment and row 0
// of the block are halos: they form fluxes but write nothing.  Every point of the padded box -2..n+2 is
// written (copied where no update applies).  The fluxes are rounded to the
// storage type (FP32 in the FP32 mode) and each is added to one face and
// subtracted from its neighbour as the same value, while B accumulates in
// FP64, so div B stays at FP64 round-off.
#ifndef CTS_CHUNK
#define CTS_CHUNK 8
#endif

__device__ __forceinline__ bool in_box(const Box &b, int i, int j, int k) {
  return i >= b.x0 && i <= b.x1 && j >= b.y0 && j <= b.y1 && k >= b.z0 && k <= b.z1;
}

// Output of point (i, j, k) at index c from the input values r1, r2, ra of
// t1, t2 and a there: F1 and F2 are the fluxes at the point, F1s and F2s
// those at s-1 along the sweep, F1a and F2a those at c - AS1 and c - AS2
// (the fluxes as stored values: rounded to the storage type).  SC = +1
// converts the inputs to face fluxes and SC = -1 the results back, on each
// component's index range (remap.f90:44-67 and 102-125); points outside that
// range are copied as they are.
template <int AX, int SC>
__device__ __forceinline__ void ct_store(const Wid &wd, int c, int i, int j, int k, double r1, double r2, double ra,
                                         double *__restrict__ t1o, double *__restrict__ t2o,
                                         double *__restrict__ ao, real F1, real F1s, real F1a, real F2, real F2s,
                                         real F2a, const Box &bt1, const Box &ba1, const Box &bt2, const Box &ba2) {
  constexpr int T1 = CtComp<AX>::T1, T2 = CtComp<AX>::T2;
  {
    const bool ok = SC == 0 || bface_ok<T1>(i, j, k);
    double v = r1;
    if (SC > 0 && ok) v = to_flux_w<T1>(v, wd);
    if (in_box(bt1, i, j, k)) v = v - (double)F1 + (double)F1s;
    if (SC < 0 && ok) v = from_flux_w<T1>(v, wd);
    t1o[c] = v;
  }
  {
    const bool ok = SC == 0 || bface_ok<T2>(i, j, k);
    double v = r2;
    if (SC > 0 && ok) v = to_flux_w<T2>(v, wd);
    if (in_box(bt2, i, j, k)) v = v - (double)F2 + (double)F2s;
    if (SC < 0 && ok) v = from_flux_w<T2>(v, wd);
    t2o[c] = v;
  }
  {
    const bool ok = SC == 0 || bface_ok<AX>(i, j, k);
    double v = ra;
    if (SC > 0 && ok) v = to_flux_w<AX>(v, wd);
    if (in_box(ba1, i, j, k)) v = v + (double)F1 - (double)F1a;
    if (in_box(ba2, i, j, k)) v = v + (double)F2 - (double)F2a;
    if (SC < 0 && ok) v = from_flux_w<AX>(v, wd);
    ao[c] = v;
  }
}

// y and z sweeps (M = 1, 2).  Rows of CTM_XW lanes along x, CTM_ROWS rows
// along the other transverse axis P (which is t2's axis, so F2a comes from
// the row below; t1 is bx, so F1a comes from the lane below).  Two phases
// per chunk of CTS_CHUNK points along the sweep:
//  1. each thread marches its column and forms the fluxes of points m0-1 ..
//     m1 into LDS, holding the four-point stencils of db1, of both fields'
//     widths and of their quotients field / width in registers (a quotient
//     is formed once per point and serves as the s+2, s+1, s and s-1 value of
//     four successive fluxes, with the same operations as forming it at
//     each use);
//  2. after one barrier, each output thread writes its points m0 .. m1,
//     taking the fluxes at the point, at s-1, at the lane below and at the
//     row below from LDS.
#define CTM_XW 32
#ifndef CTM_THREADS
#define CTM_THREADS 256
#endif
#define CTM_ROWS (CTM_THREADS / CTM_XW)
template <int M, int SC>
__global__ __launch_bounds__(CTM_THREADS) void k_ct_sweep_m(Geo g, const double *__restrict__ t1i,
                                                            const double *__restrict__ t2i,
                                                            const double *__restrict__ ai, double *__restrict__ t1o,
                                                            double *__restrict__ t2o, double *__restrict__ ao,
                                                            const real *__restrict__ vel,
                                                            const real *__restrict__ db1, int VS1, int GS1, int VS2,
                                                            int GS2, double dt, int ntx, Box bt1, Box ba1, Box bt2,
                                                            Box ba2) {
  constexpr int T1 = CtComp<M>::T1, T2 = CtComp<M>::T2;
  constexpr int P = M == 2 ? 1 : 2;
  // fluxes F1, F2 of points m0-1 .. m1 (slot 0 = m0-1)
  __shared__ real fl[CTS_CHUNK + 1][2][CTM_ROWS][CTM_XW];
  const int xl = (int)threadIdx.x % CTM_XW, row = (int)threadIdx.x / CTM_XW;
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = (int)(tile % (unsigned)ntx), ty = (int)(tile / (unsigned)ntx);
  const int i = -3 + tx * (CTM_XW - 1) + xl;
  const int p = -3 + ty * (CTM_ROWS - 1) + row;
  const int n = g.n(M), nx = g.nx, nP = g.n(P);
  const int m0 = -2 + (int)blockIdx.z * CTS_CHUNK;
  const int m1 = m0 + CTS_CHUNK - 1 < n + 2 ? m0 + CTS_CHUNK - 1 : n + 2;
  // a flux column lies in the flux box across the sweep; an output column
  // in the padded box, off the halo lane and row
  const bool col = i >= -1 && i <= nx + 1 && p >= -1 && p <= nP + 1;
  const bool outc = xl >= 1 && row >= 1 && i >= -2 && i <= nx + 2 && p >= -2 && p <= nP + 2;
  auto at = [&](int m) { return M == 2 ? g.idx(i, p, m) : g.idx(i, m, p); };
  // widths at the point: x and P fixed along the march, the sweep axis's per point
  const double *__restrict__ dbm = M == 1 ? g.dyb : g.dzb;
  Wid wd;
  wd.w[0] = i >= -2 && i <= nx + 2 ? g.dxb[i] : 1.0;
  wd.w[P] = p >= -2 && p <= nP + 2 ? (P == 1 ? g.dyb[p] : g.dzb[p]) : 1.0;
  wd.w[M] = 1.0;
  {
    double q1[4], q2[4];
    acc w1[4], w2[4];
    real d[4];
    auto load = [&](int slot, int m) {
      const int c = at(m);
      d[slot] = db1[c];
      w1[slot] = (d[slot] + db1[c + GS1]) * (acc)0.5;
      w2[slot] = (d[slot] + db1[c + GS2]) * (acc)0.5;
      double v1 = t1i[c], v2 = t2i[c];
      if (SC > 0) {
        Wid wm = wd;
        wm.w[M] = dbm[m];
        v1 = to_flux_w<T1>(v1, wm);
        v2 = to_flux_w<T2>(v2, wm);
      }
      q1[slot] = v1 / (double)w1[slot];
      q2[slot] = v2 / (double)w2[slot];
    };
    // sweep velocity at the flux point and its two averaging neighbours
    real vc = 0, va = 0, vb = 0;
    auto loadv = [&](int m) {
      if (m >= -1 && m <= n + 1) {
        const int c = at(m);
        vc = vel[c];
        va = vel[c - VS1];
        vb = vel[c - VS2];
      }
    };
    auto clampm = [&](int m) { return m < -1 ? -1 : (m > n + 2 ? n + 2 : m); };
    // window over m-1, m, m+1, m+2 of the first flux point m0 - 1
    if (col) {
#pragma unroll
      for (int q = 0; q < 4; ++q) load(q, clampm(m0 - 2 + q));
      loadv(m0 - 1);
    }
    for (int m = m0 - 1; m <= m1; ++m) {
      real F1 = 0, F2 = 0;
      if (col && m >= -1 && m <= n + 1) {
        const acc dcm = (acc)0.5 * (d[0] + (m > -1 ? d[1] : d[2]));
        const acc dci = (acc)0.5 * (d[1] + d[2]);
        const acc dcp = (acc)0.5 * (d[2] + d[3]);
        F1 = (real)ct_flux_q((vc + va) * 0.5, q1[0], q1[1], q1[2], q1[3], w1[1], w1[2], d[1], d[2], dcm, dci, dcp,
                             dt);
        F2 = (real)ct_flux_q((vc + vb) * 0.5, q2[0], q2[1], q2[2], q2[3], w2[1], w2[2], d[1], d[2], dcm, dci, dcp,
                             dt);
      }
      fl[m - m0 + 1][0][row][xl] = F1;
      fl[m - m0 + 1][1][row][xl] = F2;
      if (col && m < m1) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          q1[q] = q1[q + 1];
          q2[q] = q2[q + 1];
          w1[q] = w1[q + 1];
          w2[q] = w2[q + 1];
          d[q] = d[q + 1];
        }
        load(3, clampm(m + 3));
        loadv(m + 1);
      }
    }
  }
  __syncthreads();
  if (!outc) return;
#pragma unroll 4
  for (int m = m0; m <= m1; ++m) {
    const int c = at(m);
    const int s = m - m0 + 1;
    const double r1 = t1i[c], r2 = t2i[c], ra = ai[c];
    if (SC != 0) wd.w[M] = dbm[m];
    const int jj = M == 1 ? m : p, kk = M == 2 ? m : p;
    ct_store<M, SC>(wd, c, i, jj, kk, r1, r2, ra, t1o, t2o, ao, fl[s][0][row][xl], fl[s - 1][0][row][xl],
                    fl[s][0][row][xl - 1], fl[s][1][row][xl], fl[s - 1][1][row][xl], fl[s][1][row - 1][xl], bt1, ba1,
                    bt2, ba2);
  }
}

// x sweep.  Rows of 64 lanes along x (the sweep axis: every lane forms the
// quotients of both fields at its own point, and the stencil values at s-1,
// s+1, s+2 come from other lanes by wavefront shuffles, so each quotient is
// divided once), CTX_ROWS rows along y (t1 is by, so F1a comes from the row
// below), marching in z (t2 is bz, so F2a is the previous plane's).  Lanes
// 1..61 form fluxes from the quotients of lanes 0..63; lanes 2..61 write,
// CTX_W = 60 points per row.  The same two phases as the y and z sweeps:
// the fluxes of planes k0-1 .. k1 into LDS, one barrier, the outputs.
#define CTX_W 60
#ifndef CTX_THREADS
#define CTX_THREADS 512
#endif
#define CTX_ROWS (CTX_THREADS / 64)
template <int SC>
__global__ __launch_bounds__(CTX_THREADS) void k_ct_sweep_x(Geo g, const double *__restrict__ t1i,
                                                            const double *__restrict__ t2i,
                                                            const double *__restrict__ ai, double *__restrict__ t1o,
                                                            double *__restrict__ t2o, double *__restrict__ ao,
                                                            const real *__restrict__ vel,
                                                            const real *__restrict__ db1, int GS1, int VS2, double dt,
                                                            int ntx, Box bt1, Box ba1, Box bt2, Box ba2) {
  constexpr int T1 = CtComp<0>::T1, T2 = CtComp<0>::T2;
  __shared__ real fl[CTS_CHUNK + 1][2][CTX_ROWS][64];
  const int lane = (int)threadIdx.x & 63, row = (int)threadIdx.x >> 6;
  const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = (int)(tile % (unsigned)ntx), ty = (int)(tile / (unsigned)ntx);
  const int i = -4 + tx * CTX_W + lane;
  const int j = -3 + ty * (CTX_ROWS - 1) + row;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int Z = g.sz;
  const int k0 = -2 + (int)blockIdx.z * CTS_CHUNK;
  const int k1 = k0 + CTS_CHUNK - 1 < nz + 2 ? k0 + CTS_CHUNK - 1 : nz + 2;
  // quotient points x = -1..n+2 and flux points x = -1..n+1 of flux rows
  // y = -1..n+1; output points of the padded box off the halo lanes and row
  const bool qrow = j >= -1 && j <= ny + 1 && i >= -1 && i <= nx + 2;
  const bool frow = qrow && lane >= 1 && lane <= CTX_W + 1 && i <= nx + 1;
  const bool outc = lane >= 2 && lane <= CTX_W + 1 && row >= 1 && i >= -2 && i <= nx + 2 && j >= -2 && j <= ny + 2;
  const int c0 = (qrow || outc) ? g.idx(i, j, 0) : 0;
  // widths at the point: x and y fixed along the march, z's per plane
  Wid wd;
  wd.w[0] = i >= -2 && i <= nx + 2 ? g.dxb[i] : 1.0;
  wd.w[1] = j >= -2 && j <= ny + 2 ? g.dyb[j] : 1.0;
  wd.w[2] = 1.0;
  {
    // inputs of plane k, fetched one iteration ahead
    real fd = 0, fg = 0, fz = 0, fv = 0, fv2 = 0, fvz = 0;
    double f1 = 0, f2 = 0;
    auto fetch = [&](int k) {
      const int c = c0 + k * Z;
      if (qrow && k >= -1 && k <= nz + 1) {
        fd = db1[c];
        fg = db1[c + GS1];
        fz = db1[c + Z];
        fv = vel[c];
        fv2 = vel[c - VS2];
        if (k == -1 || k == k0 - 1) fvz = vel[c - Z];
        f1 = t1i[c];
        f2 = t2i[c];
      }
    };
    fetch(k0 - 1);
    real vm = 0;
    for (int k = k0 - 1; k <= k1; ++k) {
      const bool qk = qrow && k >= -1 && k <= nz + 1;
      // this plane's inputs; then the next plane's loads go out
      const real d = fd, dg = fg, dz = fz, vc0 = fv, v2 = fv2, vz = fvz;
      const double r1 = f1, r2 = f2;
      if (k < k1) fetch(k + 1);
      if (SC > 0 && k >= -2) wd.w[2] = g.dzb[k];
      double q1 = 0, q2 = 0;
      acc w1 = 0, w2 = 0, dd = 0;
      real vc = 0;
      if (qk) {
        w1 = (d + dg) * (acc)0.5;
        w2 = (d + dz) * (acc)0.5;
        q1 = (SC > 0 ? to_flux_w<T1>(r1, wd) : r1) / (double)w1;
        q2 = (SC > 0 ? to_flux_w<T2>(r2, wd) : r2) / (double)w2;
        dd = d;
        vc = vc0;
        if (k == -1 || k == k0 - 1) vm = vz;
      }
      // neighbours along x; s-1 is clamped to -1 and s+2 to n+2
      double q1m = __shfl_up(q1, 1, 64), q2m = __shfl_up(q2, 1, 64);
      acc dm = __shfl_up(dd, 1, 64);
      const double q1p = __shfl_down(q1, 1, 64), q2p = __shfl_down(q2, 1, 64);
      const acc dp = __shfl_down(dd, 1, 64);
      double q1q = __shfl_down(q1, 2, 64), q2q = __shfl_down(q2, 2, 64);
      acc dq = __shfl_down(dd, 2, 64);
      const acc w1p = __shfl_down(w1, 1, 64), w2p = __shfl_down(w2, 1, 64);
      if (i - 1 < -1) {
        q1m = q1;
        q2m = q2;
        dm = dd;
      }
      if (i + 2 > nx + 2) {
        q1q = q1p;
        q2q = q2p;
        dq = dp;
      }
      real F1 = 0, F2 = 0;
      if (frow && qk) {
        const acc dcm = (acc)0.5 * (dm + (i > -1 ? dd : dp));
        const acc dci = (acc)0.5 * (dd + dp);
        const acc dcp = (acc)0.5 * (dp + dq);
        F1 = (real)ct_flux_q((vc + vm) * 0.5, q1m, q1, q1p, q1q, w1, w1p, dd, dp, dcm, dci, dcp, dt);
        F2 = (real)ct_flux_q((vc + v2) * 0.5, q2m, q2, q2p, q2q, w2, w2p, dd, dp, dcm, dci, dcp, dt);
      }
      vm = vc;
      fl[k - k0 + 1][0][row][lane] = F1;
      fl[k - k0 + 1][1][row][lane] = F2;
    }
  }
  __syncthreads();
  if (!outc) return;
#pragma unroll 4
  for (int k = k0; k <= k1; ++k) {
    const int c = c0 + k * Z;
    const int s = k - k0 + 1;
    const double r1 = t1i[c], r2 = t2i[c], ra = ai[c];
    if (SC != 0) wd.w[2] = g.dzb[k];
    ct_store<0, SC>(wd, c, i, j, k, r1, r2, ra, t1o, t2o, ao, fl[s][0][row][lane], fl[s][0][row][lane - 1],
                    fl[s][0][row - 1][lane], fl[s][1][row][lane], fl[s][1][row][lane - 1], fl[s - 1][1][row][lane],
                    bt1, ba1, bt2, ba2);
  }
}

// Face-averaged sweep velocity of the mass/energy fluxes (xremap.f90:492-493)
template <int AX>
__device__ __forceinline__ acc face_velocity(const Geo &g, const real *__restrict__ vel, int c) {
  const int X = 1, Y = g.sy, Z = g.sz;
  if (AX == 0) return (vel[c] + vel[c - Y] + vel[c - Z] + vel[c - Y - Z]) * (acc)0.25;
  if (AX == 1) return (vel[c] + vel[c - X] + vel[c - Z] + vel[c - X - Z]) * (acc)0.25;
  return (vel[c] + vel[c - X] + vel[c - Y] + vel[c - X - Y]) * (acc)0.25;
}
// face area, formed in FP64 from the metrics and rounded once
template <int AX>
__device__ __forceinline__ acc face_area(const Geo &g, int i, int j, int k) {
  if (AX == 0) return (acc)(g.dyb[j] * g.dzb[k]);
  if (AX == 1) return (acc)(g.dxb[i] * g.dzb[k]);
  return (acc)(g.dxb[i] * g.dyb[j]);
}

// Van Leer upwind value fu and limited slope Di of a cell-centred quantity
// along the sweep (xremap.f90:474-533), from its values at s-1..s+2, dxc1 at
// s-1, s, s+1 and db1 at s, s+1; also the upwind weights, dbu and phi.
// The field values come in FP64 (exact for FP32 storage), so the gradients
// that the limiter compares are FP64; the rest is formed in the arithmetic
// type.
__device__ __forceinline__ void vl_cell(acc v_advect, double fm, double fi, double fp, double fp2, acc dcm, acc dci,
                                        acc dcp, acc db1c, acc db1p, acc dt, acc &vad_p, acc &vad_m, acc &dbu,
                                        acc &fu, acc &Di, acc &phi) {
  const double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
  const double sign_v = copysign(1.0, (double)v_advect);
  vad_p = (acc)((sign_v + 1.0) * 0.5);
  vad_m = (acc)1.0 - vad_p;
  fu = (acc)fi * vad_p + (acc)fp * vad_m;
  const double dfu = dfm * (double)vad_p + dfp * (double)vad_m;
  const acc dcu = dcm * vad_p + dcp * vad_m;
  dbu = db1c * vad_p + db1p * vad_m;
  phi = fabs(v_advect) * dt / dbu;
  Di = vl_di<acc>(dfi, dfu, sign_v, dci, dcu, dbu, phi);
}

// Mass flux (x_mass_flux, xremap.f90:474-533), energy flux (x_energy_flux,
// :539-601) and the density and energy updates (:174-196) of the y and z
// sweeps (M = 1, 2) in one pass, marching along the sweep axis.  The launch
// box is the mass flux's: sweep 0..n, transverse 0..n+1.  Every column
// forms dm on its points from rho1 and db1 of m-1..m+2 in registers and
// stores it.  The columns with transverse 1..n also form the energy flux
// from the same window (energy of m-1..m+2) and update rho and energy at
// 1..n, the flux and dm at m-1 carried from the previous point (rounded to
// the storage type, as stored values are).  A chunk that starts at m0 > 0
// first forms the fluxes of m0-1, whose dm the chunk below stores.  On a
// periodic sweep axis the dm halo fill that follows overwrites dm(0) with
// dm(n), and the energy flux at 0 and the update at 1 take dm(0) after that
// fill, so there the first chunk forms dm(n) from its own window at n and
// uses it at 0.  The new density goes to rho_new and the new energy to
// e_new, since neighbouring chunks still read the old energy; only cells
// 1..n are written (the halo refresh after the sweep fills the rest).
template <int M>
__global__ __launch_bounds__(BX * BY) void k_mass_rho_energy_m(Geo g, Box b, const real *__restrict__ energy,
                                                               const real *__restrict__ rho1,
                                                               const real *__restrict__ vel,
                                                               const real *__restrict__ db1, real *__restrict__ dm,
                                                               const real *__restrict__ cv1,
                                                               const real *__restrict__ cv2,
                                                               real *__restrict__ rho_new, real *__restrict__ e_new,
                                                               double dt, int periodic) {
  MCOLUMN(M, b, i, p, m0, m1)
  const int MS = M == 2 ? g.sz : g.sy;
  const int nP = g.n(M == 2 ? 1 : 2);
  const bool inner = i >= 1 && i <= g.nx && p >= 1 && p <= nP;
  const int ms = m0 > 0 ? m0 - 1 : m0;  // first point whose fluxes are formed
  const acc tdt = (acc)dt;
  // dm(n) for the wrapped dm(0) of a periodic axis (first chunk only)
  real dm_wrap = 0;
  if (periodic && ms == 0 && inner) {
    const int nM = g.n(M);
    const int cn = g.idx(i, MJ(M, nM, p), MK(M, nM, p));
    real rw[4], dw[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      rw[q] = rho1[cn + (q - 1) * MS];
      dw[q] = db1[cn + (q - 1) * MS];
    }
    const acc area = face_area<M>(g, i, MJ(M, nM, p), MK(M, nM, p));
    const acc v_advect = face_velocity<M>(g, vel, cn);
    acc vad_p, vad_m, dbu, fu, Di, phi;
    vl_cell(v_advect, rw[0], rw[1], rw[2], rw[3], (acc)0.5 * (dw[0] + dw[1]), (acc)0.5 * (dw[1] + dw[2]),
            (acc)0.5 * (dw[2] + dw[3]), dw[1], dw[2], tdt, vad_p, vad_m, dbu, fu, Di, phi);
    dm_wrap = (fu + Di * ((acc)1.0 - phi)) * v_advect * tdt * area;
  }
  // window at ms: rho1, db1 (and energy in inner columns) at ms-1..ms+2
  real r[4], d[4], e[4] = {0, 0, 0, 0};
  int c = g.idx(i, MJ(M, ms, p), MK(M, ms, p));
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    r[q] = rho1[c + (q - 1) * MS];
    d[q] = db1[c + (q - 1) * MS];
    if (inner) e[q] = energy[c + (q - 1) * MS];
  }
  real Fp = 0, dp = 0;
  for (int m = ms; m <= m1; ++m) {
    if (m > ms) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        r[q] = r[q + 1];
        d[q] = d[q + 1];
        e[q] = e[q + 1];
      }
      r[3] = rho1[c + 3 * MS];
      d[3] = db1[c + 3 * MS];
      if (inner) e[3] = energy[c + 3 * MS];
      c += MS;
    }
    const int j = MJ(M, m, p), k = MK(M, m, p);
    const acc area = face_area<M>(g, i, j, k);
    const acc v_advect = face_velocity<M>(g, vel, c);
    const acc dcm = (acc)0.5 * (d[0] + d[1]), dci = (acc)0.5 * (d[1] + d[2]), dcp = (acc)0.5 * (d[2] + d[3]);
    acc vad_p, vad_m, dbu, fu, Di, phi;
    vl_cell(v_advect, r[0], r[1], r[2], r[3], dcm, dci, dcp, d[1], d[2], tdt, vad_p, vad_m, dbu, fu, Di, phi);
    const real dm_own = (fu + Di * ((acc)1.0 - phi)) * v_advect * tdt * area;
    if (m >= m0) dm[c] = dm_own;
    if (!inner) continue;
    const real dmc = (periodic && m == 0) ? dm_wrap : dm_own;
    vl_cell(v_advect, e[0], e[1], e[2], e[3], dcm, dci, dcp, d[1], d[2], tdt, vad_p, vad_m, dbu, fu, Di, phi);
    const acc rhou = (acc)r[1] * vad_p + (acc)r[2] * vad_m;
    const acc dd = dmc;
    const acc dmu = fabs(dd) / area / dbu / rhou;
    const real F = (fu + Di * ((acc)1.0 - dmu)) * dd;
    if (m >= m0 && m >= 1) {
      const acc rr = r[1];
      const acc rn = (rr * cv1[c] + dp - dmc) / cv2[c];
      rho_new[c] = rn;
      e_new[c] = (e[1] * cv1[c] * rr + Fp - F) / (cv2[c] * rn);
    }
    Fp = F;
    dp = dmc;
  }
}

// The same for the x sweep (mass flux, energy flux and the density and
// energy updates in one pass), the flux and dm at s-1 taken from the
// neighbouring lane: waves cover runs of 64 points of a plane of the mass
// flux's box (x 0..n, transverse 0..n+1) flattened over x, overlapping by
// one; lanes 1..REX_OUT own their points.  Every point forms dm and its
// owner stores it; points with transverse 1..n also form the energy flux,
// and those at x >= 1 update rho and energy.  On a periodic x axis the
// lanes at x = 0 form dm(n) as well and use it, as the dm halo fill would
// have made them.
#define REX_OUT 63
__device__ __forceinline__ acc mass_flux_x(const Geo &g, const real *__restrict__ rho, const real *__restrict__ vel,
                                          const real *__restrict__ db1, int c, int i, int j, int k, acc dt) {
  const acc area = face_area<0>(g, i, j, k);
  const acc v_advect = face_velocity<0>(g, vel, c);
  acc vad_p, vad_m, dbu, fu, Di, phi;
  vl_cell(v_advect, rho[c - 1], rho[c], rho[c + 1], rho[c + 2], dc1<0>(g, db1, c - 1, i - 1), dc1<0>(g, db1, c, i),
          dc1<0>(g, db1, c + 1, i + 1), db1[c], db1[c + 1], dt, vad_p, vad_m, dbu, fu, Di, phi);
  return (fu + Di * ((acc)1.0 - phi)) * v_advect * dt * area;
}
__global__ __launch_bounds__(BX * BY) void k_mass_rho_energy_x(Geo g, Box b, const real *__restrict__ energy,
                                                               const real *__restrict__ rho1,
                                                               const real *__restrict__ vel,
                                                               const real *__restrict__ db1, real *__restrict__ dm,
                                                               const real *__restrict__ cv1,
                                                               const real *__restrict__ cv2,
                                                               real *__restrict__ rho_new, real *__restrict__ e_new,
                                                               double dt, int periodic) {
  const int lane = (int)threadIdx.x;  // BX = 64: threadIdx.y is the wave
  const int wq = b.x1 - b.x0 + 1;
  const int npts = wq * (b.y1 - b.y0 + 1);
  const int f = (int)(xcd_tile(blockIdx.x, gridDim.x) * BY + threadIdx.y) * REX_OUT + lane - 1;
  const bool valid = f >= 0 && f < npts;
  const int ff = valid ? f : 0;  // lanes past the plane read point 0, unused
  const int jr = ff / wq;
  const int i = b.x0 + (ff - jr * wq);
  const int j = b.y0 + jr;
  const int k = b.z0 + (int)blockIdx.z;
  const int c = g.idx(i, j, k);
  const acc tdt = (acc)dt;
  const real dm_own = mass_flux_x(g, rho1, vel, db1, c, i, j, k, tdt);
  if (valid && lane >= 1) dm[c] = dm_own;
  const bool inner = j >= 1 && j <= g.ny && k >= 1 && k <= g.nz;
  real dmc = dm_own, F = 0;
  if (inner) {
    if (periodic && i == 0) {
      const int cn = g.idx(g.nx, j, k);
      dmc = mass_flux_x(g, rho1, vel, db1, cn, g.nx, j, k, tdt);
    }
    // the energy flux at this point
    const acc area = face_area<0>(g, i, j, k);
    const acc v_advect = face_velocity<0>(g, vel, c);
    acc vad_p, vad_m, dbu, fu, Di, phi;
    vl_cell(v_advect, energy[c - 1], energy[c], energy[c + 1], energy[c + 2], dc1<0>(g, db1, c - 1, i - 1),
            dc1<0>(g, db1, c, i), dc1<0>(g, db1, c + 1, i + 1), db1[c], db1[c + 1], tdt, vad_p, vad_m, dbu, fu, Di,
            phi);
    const acc rhou = rho1[c] * vad_p + rho1[c + 1] * vad_m;
    const acc d = dmc;
    const acc dmu = fabs(d) / area / dbu / rhou;
    F = (fu + Di * ((acc)1.0 - dmu)) * d;
  }
  // lane l-1 holds point i-1 of the same row whenever i > x0
  const real Fm = __shfl_up(F, 1, 64);
  const real dmm = __shfl_up(dmc, 1, 64);
  if (!valid || lane < 1 || !inner || i < 1) return;
  const acc r1 = rho1[c];
  const acc rn = (r1 * cv1[c] + dmm - dmc) / cv2[c];
  rho_new[c] = rn;
  e_new[c] = (energy[c] * cv1[c] * r1 + Fm - F) / (cv2[c] * rn);
}



template <int AX>
__device__ __forceinline__ acc mom_area(const Geo &g, int i, int j, int k) {
  if (AX == 0) return (acc)(g.dyc[j] * g.dzc[k]);
  if (AX == 1) return (acc)(g.dxc[i] * g.dzc[k]);
  return (acc)(g.dxc[i] * g.dyc[j]);
}

// Vertex quantities of xremap.f90:201-280 at a node: rho_v, cv1_v, cv2_v
// and dm_v over the eight cells (i..i+1, j..j+1, k..k+1), summed in the
// order (i,j,k), (i+1,j,k), (i,j+1,k), (i+1,j+1,k), then the same four at
// k+1, and rounded to the storage type as stored values are.  A kernel that
// marches along z or y forms them node after node: VSum holds the four
// running sums, and VMarch<M> what the next node along M can take over
// from this one, so that only the new plane (z) or row (y) of cells is
// loaded there.  The sums are formed in the same order either way.
struct VertexV {
  real rv, c1, c2, dm;
};
struct VSum {
  acc s, s1, s2, sd;  // sum of rho1*cv1, of cv1, of cv2, of dm
};
// cells at c, c+1 (and c+D, c+1+D for N = 4): one x pair, or a quad
template <int N>
struct VCellN {
  real r[N], a[N], b[N], d[N];
};
template <int N>
__device__ __forceinline__ VCellN<N> vcells_at(const real *__restrict__ rho1, const real *__restrict__ cv1,
                                               const real *__restrict__ cv2, const real *__restrict__ dm, int c,
                                               int D) {
  VCellN<N> v;
#pragma unroll
  for (int t = 0; t < N; ++t) {
    const int o = c + (t & 1) + (t >> 1) * D;
    v.r[t] = rho1[o];
    v.a[t] = cv1[o];
    v.b[t] = cv2[o];
    v.d[t] = dm[o];
  }
  return v;
}
// sums over the cells of v, starting with v's first cell
template <int N>
__device__ __forceinline__ VSum vsum_first(const VCellN<N> &v) {
  VSum p;
  p.s = (acc)v.r[0] * v.a[0];
  p.s1 = v.a[0];
  p.s2 = v.b[0];
  p.sd = v.d[0];
#pragma unroll
  for (int t = 1; t < N; ++t) {
    p.s = p.s + (acc)v.r[t] * v.a[t];
    p.s1 = p.s1 + v.a[t];
    p.s2 = p.s2 + v.b[t];
    p.sd = p.sd + v.d[t];
  }
  return p;
}
// p continued over the cells of v
template <int N>
__device__ __forceinline__ VSum vsum_add(VSum p, const VCellN<N> &v) {
#pragma unroll
  for (int t = 0; t < N; ++t) {
    p.s = p.s + (acc)v.r[t] * v.a[t];
    p.s1 = p.s1 + v.a[t];
    p.s2 = p.s2 + v.b[t];
    p.sd = p.sd + v.d[t];
  }
  return p;
}
__device__ __forceinline__ VertexV vertex_of(const VSum &p) {
  VertexV v;
  v.rv = p.s / p.s1;
  v.c1 = p.s1 * (acc)0.125;
  v.c2 = p.s2 * (acc)0.125;
  v.dm = p.sd * (acc)0.125;
  return v;
}
// What a node passes to the next one along M.  Along z (M = 2) the next
// node's first four cells are this node's last four, so their sums carry
// over whole.  Along y (M = 1) the next node's cells 0, 1 are this node's
// 2, 3 (their sums carry over) and its cells 4, 5 are this node's 6, 7, which
// come between new cells in the order and so carry over as values.
template <int M>
struct VMarch {
  VSum head;
  VCellN<M == 1 ? 2 : 1> tail;  // unused for M = 2
};
// Vertex quantities of the node at c.  With have set, m holds what the node
// one step below c along M passed on; on return it holds what this node
// passes on.
template <int M>
__device__ __forceinline__ VertexV vertex_march(const Geo &g, const real *__restrict__ rho1,
                                                const real *__restrict__ cv1, const real *__restrict__ cv2,
                                                const real *__restrict__ dm, int c, VMarch<M> &m, bool have) {
  const int Y = g.sy, Z = g.sz;
  if constexpr (M == 2) {
    const VSum lo = have ? m.head : vsum_first(vcells_at<4>(rho1, cv1, cv2, dm, c, Y));
    const VCellN<4> hi = vcells_at<4>(rho1, cv1, cv2, dm, c + Z, Y);
    m.head = vsum_first(hi);
    return vertex_of(vsum_add(lo, hi));
  } else {
    VSum p;
    VCellN<2> c45;
    if (have) {
      p = m.head;
      c45 = m.tail;
    } else {
      p = vsum_first(vcells_at<2>(rho1, cv1, cv2, dm, c, Y));
      c45 = vcells_at<2>(rho1, cv1, cv2, dm, c + Z, Y);
    }
    const VCellN<2> c23 = vcells_at<2>(rho1, cv1, cv2, dm, c + Y, Y);
    const VCellN<2> c67 = vcells_at<2>(rho1, cv1, cv2, dm, c + Y + Z, Y);
    p = vsum_add(p, c23);
    p = vsum_add(p, c45);
    p = vsum_add(p, c67);
    m.head = vsum_first(c23);
    m.tail = c67;
    return vertex_of(p);
  }
}

// Momentum remap of the x sweep: the vertex quantities (xremap.f90:201-280),
// the fluxes x_momx/y/z_flux (xremap.f90:605-895, rke off; sweep -1..n,
// transverse 0..n, the three components sharing the upwind switch and
// geometry) and the updates of xremap.f90:282-329 (nodes 0..n) in one pass.
// Waves cover runs of 64 points of a plane of the launch box (the whole
// padded box) flattened along x, successive waves overlapping by two, and
// march in z over ZCHUNK planes, keeping the upper plane of vertex cells for
// the next node: each lane forms the vertex quantities of its node, takes
// rho_v of node i+1 from the lane to its right and the flux and dm_v of node
// i-1 from the lane to its left; lanes 1..MOMX_OUT write.  Every point is
// written to vxo..vzo: nodes 0..n updated, the rest copied.
//
// RKE: also the kinetic energy the sweep loses at each edge between nodes
// s and s+1 (xremap.f90:664-695, s = 0..n-1, transverse 0..n), one array
// per velocity component, indexed by the edge's lower node; rke_gather
// adds them to delta_ke in the reference's order.  The edge term of node i
// needs the flux velocity of node i+1, whose upwind density is node i+2's,
// so with RKE the waves overlap by three and lanes 1..61 write.
struct RkeOut {
  real *dk[3];
};
// The edge term dk of xremap.f90:673-686 from the old velocities vs, vp of
// the edge's nodes, the flux velocities (before the dm factor) at the edges
// s-1, s, s+1, the vertex dm there, and the node masses ms, mp after the
// sweep (rho_v1 cv2).
__device__ __forceinline__ acc rke_edge(acc vs, acc vp, acc flm, acc fl, acc flp, acc dmm, acc dms, acc dmp, acc ms,
                                        acc mp) {
  const acc ai = (vs - flm) * dmm / ms - (vs - fl) * dms / ms;
  const acc aip = (vp - fl) * dms / mp - (vp - flp) * dmp / mp;
  acc dk = (vp - vs) * (fl - (acc)0.5 * (vp + vs)) - (acc)0.5 * ai * (vs - fl) + (acc)0.5 * aip * (vp - fl);
  dk = dk * dms * (acc)0.5;
  return dk;
}
#define MOMX_OUT 62
#define MOMX_OUT_RKE 61
template <bool RKE>
__global__ __launch_bounds__(BX * BY) void k_mom_remap_x(Geo g, Box b, const real *__restrict__ vx,
                                                         const real *__restrict__ vy, const real *__restrict__ vz,
                                                         const real *__restrict__ vel1,
                                                         const real *__restrict__ db1,
                                                         const real *__restrict__ rho1,
                                                         const real *__restrict__ cv1,
                                                         const real *__restrict__ cv2,
                                                         const real *__restrict__ dm, real *__restrict__ vxo,
                                                         real *__restrict__ vyo, real *__restrict__ vzo, double dt,
                                                         RkeOut rk) {
  const int lane = (int)threadIdx.x;  // BX = 64: threadIdx.y is the wave
  const int wq = b.x1 - b.x0 + 1;
  const int npts = wq * (b.y1 - b.y0 + 1);
  constexpr int OUT = RKE ? MOMX_OUT_RKE : MOMX_OUT;
  const int f = (int)(xcd_tile(blockIdx.x, gridDim.x) * BY + threadIdx.y) * OUT + lane - 1;
  const bool valid = f >= 0 && f < npts;
  const int ff = valid ? f : 0;  // lanes past the plane read point 0, unused
  const int jr = ff / wq;
  const int i = b.x0 + (ff - jr * wq);
  const int j = b.y0 + jr;
  ZRANGE(b, k0, k1)
  const real *V[3] = {vx, vy, vz};
  real *VO[3] = {vxo, vyo, vzo};
  const int n = g.nx;
  const bool writer = valid && lane >= 1 && lane <= OUT;
  // vertex box: x -1..n+1, transverse 0..n
  const bool vin = j >= 0 && j <= g.ny && i >= -1 && i <= n + 1;
  VMarch<2> w;
  bool have = false;  // w holds what node k-1 passed on
  for (int k = k0; k <= k1; ++k) {
    const int c = g.idx(i, j, k);
    const bool iny = j >= 0 && j <= g.ny && k >= 0 && k <= g.nz;
    VertexV vq = {0, 0, 0, 0};
    if (vin && k >= 0 && k <= g.nz) {
      vq = vertex_march<2>(g, rho1, cv1, cv2, dm, c, w, have);
      have = true;
    }
    const real rvp = __shfl_down(vq.rv, 1, 64);  // rho_v of node i+1 (lanes 0..62)
    // fluxes on x = -1..n (transverse 0..n); FL: the flux velocity before
    // the dm factor (stored as the reference's flux array is)
    real F[3] = {0, 0, 0}, FL[3] = {0, 0, 0};
    if (iny && i >= -1 && i <= n) {
      const acc area = mom_area<0>(g, i, j, k);
      const acc v_advect = (vel1[c] + vel1[c + 1]) * (acc)0.5;
      const double sign_v = copysign(1.0, (double)v_advect);
      const acc vad_p = (acc)((sign_v + 1.0) * 0.5);
      const acc vad_m = (acc)1.0 - vad_p;
      const acc dci = db1[c + 1];
      const acc dcu = db1[c] * vad_p + db1[c + 2] * vad_m;
      const acc dbu = dc1<0>(g, db1, c, i) * vad_p + dc1<0>(g, db1, c + 1, i + 1) * vad_m;
      const acc phi = fabs(v_advect) * (acc)dt / dbu;
      const acc rhou = (acc)vq.rv * vad_p + (acc)rvp * vad_m;
      const acc d = vq.dm;
      const acc dmu = fabs(d) / area / dbu / rhou;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const real *u = V[q];
        const double fm = u[c - 1], fi = u[c], fp = u[c + 1], fp2 = u[c + 2];
        const double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
        const acc fu = (acc)fi * vad_p + (acc)fp * vad_m;
        const double dfu = dfm * (double)vad_p + dfp * (double)vad_m;
        const acc Di = vl_di<acc>(dfi, dfu, sign_v, dci, dcu, dbu, phi);
        const acc fl = fu + Di * ((acc)1.0 - dmu);
        FL[q] = fl;
        F[q] = fl * d;
      }
    }
    // lane l-1 holds node i-1 of the same row whenever i > x0
    real Fm[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) Fm[q] = __shfl_up(F[q], 1, 64);
    const real dmm = __shfl_up(vq.dm, 1, 64);
    // rho_v1 of node i (xremap.f90:282-290)
    const bool upd = iny && i >= 0 && i <= n;
    const acc rv = vq.rv, c1 = vq.c1, c2 = vq.c2;
    acc rvn = 0;
    if (upd) rvn = (rv * c1 + dmm - vq.dm) / c2;
    if constexpr (RKE) {
      // lane l+1 holds node i+1 of the same row for the writers
      const acc mass = (acc)(real)rvn * c2;
      const acc mp = __shfl_down(mass, 1, 64);
      const real dmp = __shfl_down(vq.dm, 1, 64);
      real FLm[3], FLp[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        FLm[q] = __shfl_up(FL[q], 1, 64);
        FLp[q] = __shfl_down(FL[q], 1, 64);
      }
      if (writer && iny && i >= 0 && i < n) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
          rk.dk[q][c] = rke_edge(V[q][c], V[q][c + 1], FLm[q], FL[q], FLp[q], dmm, vq.dm, dmp, mass, mp);
      }
    }
    if (!writer) continue;
    if (upd) {
      // the update (xremap.f90:282-329)
#pragma unroll
      for (int q = 0; q < 3; ++q) VO[q][c] = (rv * V[q][c] * c1 + Fm[q] - F[q]) / (c2 * rvn);
    } else {
#pragma unroll
      for (int q = 0; q < 3; ++q) VO[q][c] = V[q][c];
    }
  }
}

// Momentum remap of the y and z sweeps (M = 1, 2): the vertex quantities,
// fluxes and update of the x-sweep kernel above, in one pass marching along
// the sweep axis with the velocity stencil, db1 and the sweep velocity in
// registers; the vertex quantities of node m+1 are formed once, from the
// cells it shares with node m and the next row or plane of cells, and reused
// as node m's in the next iteration, and the flux at s-1 that the update at s
// needs is the previous iteration's (rounded to the storage type, as a
// stored flux is).  The new velocities go to vxo..vzo, because neighbouring
// chunks still read the old ones; every point of the launch box (the whole
// padded box) is written: nodes 0..n updated, the rest copied.
//
// RKE: the edge terms of yremap.f90:665-696 / zremap.f90:664-695 (edges s =
// 0..n-1 along the sweep, transverse 0..n), formed one point behind the
// march: the term of the edge at m-1 needs the flux velocity at m, so a
// chunk runs one point past its last node when that node's edge exists.
template <int M, bool RKE>
__global__ __launch_bounds__(BX * BY) void k_mom_remap_m(Geo g, Box b, const real *__restrict__ vx,
                                                         const real *__restrict__ vy, const real *__restrict__ vz,
                                                         const real *__restrict__ vel1,
                                                         const real *__restrict__ db1,
                                                         const real *__restrict__ rho1,
                                                         const real *__restrict__ cv1,
                                                         const real *__restrict__ cv2,
                                                         const real *__restrict__ dm, real *__restrict__ vxo,
                                                         real *__restrict__ vyo, real *__restrict__ vzo, double dt,
                                                         RkeOut rk) {
  MCOLUMN(M, b, i, p, m0, m1)
  const int MS = M == 2 ? g.sz : g.sy;
  const int nM = g.n(M), nP = g.n(M == 2 ? 1 : 2);
  const real *V[3] = {vx, vy, vz};
  real *VO[3] = {vxo, vyo, vzo};
  // update range of this column's chunk: nodes 0..n on every axis
  const bool inT = i >= 0 && i <= g.nx && p >= 0 && p <= nP;
  const int ms = inT ? (m0 > 0 ? m0 : 0) : m1 + 1;
  const int me = inT ? (m1 < nM ? m1 : nM) : m0 - 1;
  for (int m = m0; m <= m1; ++m) {
    if (m >= ms && m <= me) continue;
    const int c = g.idx(i, MJ(M, m, p), MK(M, m, p));
#pragma unroll
    for (int q = 0; q < 3; ++q) VO[q][c] = V[q][c];
  }
  if (ms > me) return;
  // window at node ms-1 (>= -1): v at ms-2..ms+1, db1 at ms-1..ms+1, the
  // sweep velocity at ms-1, ms; vertex quantities of nodes ms-1 and ms
  real u[3][4], d[3], v0, v1;
  int c = g.idx(i, MJ(M, ms - 1, p), MK(M, ms - 1, p));
#pragma unroll
  for (int q = 0; q < 3; ++q)
#pragma unroll
    for (int t = 0; t < 4; ++t) u[q][t] = V[q][c + (t - 1) * MS];
#pragma unroll
  for (int t = 0; t < 3; ++t) d[t] = db1[c + t * MS];
  v0 = vel1[c];
  v1 = vel1[c + MS];
  VMarch<M> w;
  VertexV vc = vertex_march<M>(g, rho1, cv1, cv2, dm, c, w, false);
  VertexV vn = vertex_march<M>(g, rho1, cv1, cv2, dm, c + MS, w, true);
  real Fp[3], dp = 0;
  // rke: flux velocities, dm and node mass carried from m-1 and m-2
  real FLp[3] = {0, 0, 0}, FLpp[3] = {0, 0, 0}, dpp = 0;
  acc massp = 0;
  const int mend = (RKE && me < nM) ? me + 1 : me;
  for (int m = ms - 1; m <= mend; ++m) {
    if (m > ms - 1) {
      // advance the window from m-1 (index c) to m
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        u[q][0] = u[q][1];
        u[q][1] = u[q][2];
        u[q][2] = u[q][3];
        u[q][3] = V[q][c + 3 * MS];
      }
      d[0] = d[1];
      d[1] = d[2];
      d[2] = db1[c + 3 * MS];
      v0 = v1;
      v1 = vel1[c + 2 * MS];
      vc = vn;
      vn = vertex_march<M>(g, rho1, cv1, cv2, dm, c + 2 * MS, w, true);
      c += MS;
    }
    const int j = MJ(M, m, p), k = MK(M, m, p);
    // fluxes at m
    const acc area = mom_area<M>(g, i, j, k);
    const acc v_advect = (v0 + v1) * (acc)0.5;
    const double sign_v = copysign(1.0, (double)v_advect);
    const acc vad_p = (acc)((sign_v + 1.0) * 0.5);
    const acc vad_m = (acc)1.0 - vad_p;
    const acc dci = d[1];
    const acc dcu = d[0] * vad_p + d[2] * vad_m;
    const acc dbu = (acc)0.5 * (d[0] + d[1]) * vad_p + (acc)0.5 * (d[1] + d[2]) * vad_m;
    const acc phi = fabs(v_advect) * (acc)dt / dbu;
    const acc rhou = (acc)vc.rv * vad_p + (acc)vn.rv * vad_m;
    const real dm_ = vc.dm;
    const acc dd = dm_;
    const acc dmu = fabs(dd) / area / dbu / rhou;
    real F[3], FL[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double fm = u[q][0], fi = u[q][1], fp = u[q][2], fp2 = u[q][3];
      const double dfm = fi - fm, dfi = fp - fi, dfp = fp2 - fp;
      const acc fu = (acc)fi * vad_p + (acc)fp * vad_m;
      const double dfu = dfm * (double)vad_p + dfp * (double)vad_m;
      const acc Di = vl_di<acc>(dfi, dfu, sign_v, dci, dcu, dbu, phi);
      const acc f = fu + Di * ((acc)1.0 - dmu);
      FL[q] = f;
      F[q] = f * dd;
    }
    if (m >= ms) {
      // rho_v1 at m, and the update at m (nodes ms..me)
      const acc rv = vc.rv, c1 = vc.c1, c2 = vc.c2;
      const acc rvn = (rv * c1 + dp - dm_) / c2;
      if (m <= me) {
#pragma unroll
        for (int q = 0; q < 3; ++q) VO[q][c] = (rv * u[q][1] * c1 + Fp[q] - F[q]) / (c2 * rvn);
      }
      if constexpr (RKE) {
        const acc mass = (acc)(real)rvn * c2;
        // the edge between m-1 and m (its lower node m-1 in ms..me, below n)
        if (m > ms && m <= nM) {
#pragma unroll
          for (int q = 0; q < 3; ++q)
            rk.dk[q][c - MS] = rke_edge(u[q][0], u[q][1], FLpp[q], FLp[q], FL[q], dpp, dp, dm_, massp, mass);
        }
        massp = mass;
      }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) Fp[q] = F[q];
    if constexpr (RKE) {
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        FLpp[q] = FLp[q];
        FLp[q] = FL[q];
      }
      dpp = dp;
    }
    dp = dm_;
  }
}

// delta_ke of the cells 1..n gathers the edge terms of one sweep along AX
// (xremap.f90:687-690, yremap.f90:688-691, zremap.f90:687-690): cell
// (.., s, ..) takes the edges whose lower node has sweep coordinate s-1 and
// transverse coordinates t-1 or t.  For each velocity component in turn
// (x, y, z) the four terms are added in the order the reference's loops
// reach them: the outer transverse axis T2 lower first, then the inner T1.
template <int AX>
__global__ __launch_bounds__(BX * BY) void k_rke_gather(Geo g, Box b, RkeOut rk, real *__restrict__ dke) {
  BOX_INDEX(b, i, j, k);
  constexpr int T1 = CtComp<AX>::T1, T2 = CtComp<AX>::T2;
  const int c = g.idx(i, j, k);
  const int S1 = g.stride(T1), S2 = g.stride(T2);
  const int e = c - g.stride(AX);
  acc v = dke[c];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const real *d = rk.dk[q];
    v = v + d[e - S1 - S2];
    v = v + d[e - S2];
    v = v + d[e - S1];
    v = v + d[e];
  }
  dke[c] = v;
}

// energy_correction (diagnostics.F90:613-629) on cells 1..n: the kinetic
// energy lost by the remap (delta_ke < 0) added to the specific internal
// energy.  Only the interior of delta_ke is used, so only it is formed.
__global__ __launch_bounds__(BX * BY) void k_energy_correction(Geo g, Box b, real *__restrict__ dke,
                                                               const real *__restrict__ rho,
                                                               real *__restrict__ energy) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  acc d = -(acc)dke[c];
  if (d < (acc)0.0) d = 0.0;
  d = d / ((acc)rho[c] * (acc)g.cv(i, j, k));
  dke[c] = d;
  energy[c] = energy[c] + d;
}
