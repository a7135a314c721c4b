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

// Face B component Q at (i, j, k) as the face flux that the remap moves,
// and back (remap.f90:46-48 and 104-106, same operation order), with the
// index range each component is defined on (bx(-2:nx+2, -1:ny+2, -1:nz+2)
// and its analogues, remap.f90:44-67).
template <int Q>
__device__ __forceinline__ bool bface_ok(int i, int j, int k) {
  return Q == 0 ? (j >= -1 && k >= -1) : (Q == 1 ? (i >= -1 && k >= -1) : (i >= -1 && j >= -1));
}
template <int Q>
__device__ __forceinline__ double to_flux(const Geo &g, double v, int i, int j, int k) {
  if (Q == 0) return v * g.dyb[j] * g.dzb[k];
  if (Q == 1) return v * g.dxb[i] * g.dzb[k];
  return v * g.dxb[i] * g.dyb[j];
}
template <int Q>
__device__ __forceinline__ double from_flux(const Geo &g, double v, int i, int j, int k) {
  if (Q == 0) return v / g.dyb[j] / g.dzb[k];
  if (Q == 1) return v / g.dxb[i] / g.dzb[k];
  return v / g.dxb[i] / g.dyb[j];
}
// The transverse components (t1, t2) of a sweep along AX and its CT update
// pairs: x -> (by, bz), y -> (bx, bz), z -> (bx, by).
template <int AX> struct CtComp {
  static constexpr int T1 = AX == 0 ? 1 : 0;
  static constexpr int T2 = AX == 2 ? 1 : 2;
};
// Stencil value of component Q at the point whose sweep coordinate is s
// (other coordinates those of (i, j, k)); SC = 1 reads the field as B and
// converts it to the face flux on the fly (first sweep of a remap).
template <int AX, int Q, int SC>
__device__ __forceinline__ double ct_load(const Geo &g, const double *__restrict__ f, int c, int s, int i, int j,
                                          int k) {
  const double v = f[c];
  if (!SC) return v;
  return to_flux<Q>(g, v, AX == 0 ? s : i, AX == 1 ? s : j, AX == 2 ? s : k);
}

// CT fluxes of the two transverse face fields of the x sweep (vx_by_flux and
// vx_bz_flux, xremap.f90:340-470) in one pass, the stencil running along the
// lanes.
// Each wave covers 64 consecutive points of the (x, y) plane flattened over
// x = x0..n+2 (one column beyond the flux box), successive waves overlapping
// by three points.  Every lane forms the quotients field / width of both
// fields at its own point, and lanes 1..CTX_OUT take the values at s-1, s+1
// and s+2 from their neighbours by wavefront shuffles, so each quotient is
// divided once.  The kernel marches in z with v of plane k-1 and db1 of
// plane k+1 in registers.  GS1 and VS2 are the y stride (by's widths, bz's
// velocity average); by's velocity and bz's widths are averaged along z.
#define CTX_OUT 61
template <int SC>
__global__ __launch_bounds__(BX * BY) void k_ct_flux_x(Geo g, Box b, const double *__restrict__ fld1,
                                                       const double *__restrict__ fld2,
                                                       const real *__restrict__ vel, const real *__restrict__ db1,
                                                       real *__restrict__ F1, real *__restrict__ F2, int GS1,
                                                       int VS2, double dt) {
  constexpr int T1 = CtComp<0>::T1, T2 = CtComp<0>::T2;
  const int lane = (int)threadIdx.x;  // BX = 64: threadIdx.y is the wave
  const int wq = g.nx + 2 - b.x0 + 1;  // row of quotient points x0..n+2
  const int npts = wq * (b.y1 - b.y0 + 1);
  const int f = (int)(xcd_tile(blockIdx.x, gridDim.x) * BY + threadIdx.y) * CTX_OUT + lane - 1;
  const bool valid = f >= 0 && f < npts;
  const int ff = valid ? f : 0;  // lanes past the plane read point 0, unused
  const int jr = ff / wq;
  const int i = b.x0 + (ff - jr * wq);
  const int j = b.y0 + jr;
  const bool out = valid && lane >= 1 && lane <= CTX_OUT && i <= b.x1;
  ZRANGE(b, k0, k1)
  const int Z = g.sz;
  const int nx = g.nx;
  int c = g.idx(i, j, k0);
  real vm = vel[c - Z];
  real d = db1[c];
  for (int k = k0; k <= k1; ++k, c += Z) {
    const real dzo = db1[c + Z];
    const acc w1 = (d + db1[c + GS1]) * (acc)0.5;
    const acc w2 = (d + dzo) * (acc)0.5;
    const double q1 = ct_load<0, T1, SC>(g, fld1, c, i, i, j, k) / (double)w1;
    const double q2 = ct_load<0, T2, SC>(g, fld2, c, i, i, j, k) / (double)w2;
    const acc dd = d;
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
    const real vc = vel[c];
    if (out) {
      // dxc1 at s-1 (upper point s, or s+1 when s = -1), s and s+1
      const acc dcm = (acc)0.5 * (dm + (i > -1 ? dd : dp));
      const acc dci = (acc)0.5 * (dd + dp);
      const acc dcp = (acc)0.5 * (dp + dq);
      F1[c] = ct_flux_q((vc + vm) * 0.5, q1m, q1, q1p, q1q, w1, w1p, dd, dp, dcm, dci, dcp, dt);
      F2[c] = ct_flux_q((vc + vel[c - VS2]) * 0.5, q2m, q2, q2p, q2q, w2, w2p, dd, dp, dcm, dci, dcp, dt);
    }
    vm = vc;
    d = dzo;
  }
}

// The same fluxes for the y and z sweeps (M = 1, 2).  Each thread keeps its
// (x, other transverse) column and marches along the sweep axis for ZCHUNK
// points, holding the four-point stencils of db1, of both fields' widths
// and of their quotients field / width in registers.  A quotient is formed
// once per point and serves as the s+2, s+1, s and s-1 value of four
// successive fluxes, with the same operations as forming it at each use.
template <int M, int SC>
__global__ __launch_bounds__(BX * BY) void k_ct_flux_m(Geo g, Box b, const double *__restrict__ fld1,
                                                       const double *__restrict__ fld2,
                                                       const real *__restrict__ vel, const real *__restrict__ db1,
                                                       real *__restrict__ F1, real *__restrict__ F2, int VS1,
                                                       int GS1, int VS2, int GS2, double dt) {
  constexpr int T1 = CtComp<M>::T1, T2 = CtComp<M>::T2;
  // the launch box with the march axis last: (x, P, M), P the other axis
  const Box pb = M == 2 ? b : Box{b.x0, b.x1, b.z0, b.z1, b.y0, b.y1};
  PLANE_POS(pb, i, p)
  ZRANGE(pb, m0, m1)
  if (p > pb.y1) return;
  const int n = g.n(M);
  // index, and quotients of both fields, at march coordinate m
  auto at = [&](int m) { return M == 2 ? g.idx(i, p, m) : g.idx(i, m, p); };
  double q1[4], q2[4];
  acc w1[4], w2[4];
  real d[4];
  auto load = [&](int slot, int m) {
    const int c = at(m);
    const int jj = M == 1 ? m : p, kk = M == 2 ? m : p;
    d[slot] = db1[c];
    w1[slot] = (d[slot] + db1[c + GS1]) * (acc)0.5;
    w2[slot] = (d[slot] + db1[c + GS2]) * (acc)0.5;
    q1[slot] = ct_load<M, T1, SC>(g, fld1, c, m, i, jj, kk) / (double)w1[slot];
    q2[slot] = ct_load<M, T2, SC>(g, fld2, c, m, i, jj, kk) / (double)w2[slot];
  };
  // window over m-1, m, m+1, m+2, each clamped to [-1, n+2]
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int mm = m0 - 1 + q;
    mm = mm < -1 ? -1 : (mm > n + 2 ? n + 2 : mm);
    load(q, mm);
  }
  for (int m = m0; m <= m1; ++m) {
    const int c = at(m);
    // dxc1 at the clamped s-1 (its upper point is s, or s+1 when s = -1),
    // at s, and at s+1 (upper point clamped to n+2, as d[3] already is)
    const acc dcm = (acc)0.5 * (d[0] + (m > -1 ? d[1] : d[2]));
    const acc dci = (acc)0.5 * (d[1] + d[2]);
    const acc dcp = (acc)0.5 * (d[2] + d[3]);
    const real vc = vel[c];
    F1[c] = ct_flux_q((vc + vel[c - VS1]) * 0.5, q1[0], q1[1], q1[2], q1[3], w1[1], w1[2], d[1], d[2], dcm, dci, dcp,
                      dt);
    F2[c] = ct_flux_q((vc + vel[c - VS2]) * 0.5, q2[0], q2[1], q2[2], q2[3], w2[1], w2[2], d[1], d[2], dcm, dci, dcp,
                      dt);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      q1[q] = q1[q + 1];
      q2[q] = q2[q + 1];
      w1[q] = w1[q + 1];
      w2[q] = w2[q + 1];
      d[q] = d[q + 1];
    }
    if (m < m1) load(3, m + 3 > n + 2 ? n + 2 : m + 3);
  }
}

// CT updates of a sweep along AX (xremap.f90:121-157 and analogues), both
// pairs in the reference's order:
//   t1 = t1 - F1 + F1(c - TS) on bt1;  a = a + F1 - F1(c - AS1) on ba1;
//   t2 = t2 - F2 + F2(c - TS) on bt2;  a = a + F2 - F2(c - AS2) on ba2.
// SC = 0: in place on the face fluxes (t1i == t1 etc.), over the box b.
// SC = +1 (first sweep of a remap): the inputs are B; every point of each
// component's range in b is converted to the face flux (remap.f90:44-67),
// updated where its box says so, and stored to the outputs.
// SC = -1 (last sweep of the predictor remap): the updated face flux of
// every such point is converted back to B (remap.f90:102-125).
// The fluxes F1, F2 are stored in the storage type (FP32 in the FP32 mode);
// each is added to one face and subtracted from its neighbour as the same
// stored value, and B accumulates in FP64, so div B stays at FP64 round-off.
template <int AX, int SC>
__global__ __launch_bounds__(BX * BY) void k_ct_update(Geo g, Box b, const real *__restrict__ F1,
                                                       const real *__restrict__ F2, const double *t1i,
                                                       const double *t2i, const double *ai, double *t1, double *t2,
                                                       double *a, int TS, int AS1, int AS2, Box bt1, Box ba1,
                                                       Box bt2, Box ba2) {
  BOX_INDEX(b, i, j, k);
  constexpr int T1 = CtComp<AX>::T1, T2 = CtComp<AX>::T2;
  const int c = g.idx(i, j, k);
  const bool u1 = i >= bt1.x0 && i <= bt1.x1 && j >= bt1.y0 && j <= bt1.y1 && k >= bt1.z0 && k <= bt1.z1;
  const bool w1 = i >= ba1.x0 && i <= ba1.x1 && j >= ba1.y0 && j <= ba1.y1 && k >= ba1.z0 && k <= ba1.z1;
  const bool u2 = i >= bt2.x0 && i <= bt2.x1 && j >= bt2.y0 && j <= bt2.y1 && k >= bt2.z0 && k <= bt2.z1;
  const bool w2 = i >= ba2.x0 && i <= ba2.x1 && j >= ba2.y0 && j <= ba2.y1 && k >= ba2.z0 && k <= ba2.z1;
  if (SC == 0) {
    if (u1) t1[c] = t1[c] - F1[c] + F1[c - TS];
    if (u2) t2[c] = t2[c] - F2[c] + F2[c - TS];
    if (w1 || w2) {
      double va = a[c];
      if (w1) va = va + F1[c] - F1[c - AS1];
      if (w2) va = va + F2[c] - F2[c - AS2];
      a[c] = va;
    }
    return;
  }
  if (bface_ok<T1>(i, j, k)) {
    double v = t1i[c];
    if (SC > 0) v = to_flux<T1>(g, v, i, j, k);
    if (u1) v = v - F1[c] + F1[c - TS];
    if (SC < 0) v = from_flux<T1>(g, v, i, j, k);
    t1[c] = v;
  }
  if (bface_ok<T2>(i, j, k)) {
    double v = t2i[c];
    if (SC > 0) v = to_flux<T2>(g, v, i, j, k);
    if (u2) v = v - F2[c] + F2[c - TS];
    if (SC < 0) v = from_flux<T2>(g, v, i, j, k);
    t2[c] = v;
  }
  if (bface_ok<AX>(i, j, k)) {
    double va = ai[c];
    if (SC > 0) va = to_flux<AX>(g, va, i, j, k);
    if (w1) va = va + F1[c] - F1[c - AS1];
    if (w2) va = va + F2[c] - F2[c - AS2];
    if (SC < 0) va = from_flux<AX>(g, va, i, j, k);
    a[c] = va;
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
