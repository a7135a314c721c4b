// openbc.hpp -- far-field characteristic (open) boundaries, openboundary.f90.
//
// open_bcs runs at the end of the corrector, after the velocity update and
// velocity_bcs and before the final density and energy update
// (lagran.F90:408-412).  Here the corrector's fused kernel updates density
// and energy together with the velocity, so k_open_save keeps the
// pre-update energy and density of the first interior layer of each open
// face beforehand.  The faces then run one launch each, in the reference's
// order (x_min, x_max, y_min, y_max, z_min, z_max): a face reads ghost
// cells that an earlier face may just have set (its transverse range
// 0..n+1 overlaps the neighbouring faces' first ghost layer).

// Face geometry: axis a, side hi; transverse axes u (contiguous where
// possible) and w; the reference's transverse slots ty, tz ("y" and "z" of
// the face: for x (y, z), for y (x, z), for z (y, x)); layers along a.
struct OpenFace {
  int a, hi, u, w, ty, tz;
  int lb, lc, lfc, lfn, lset_n, lset_c;
  double sg;  // -1 on the low side: the normal components point outwards
};

// One face point of open_bcs_fast / open_bcs_alfven (openboundary.f90:327-455).
// Index 1: values carried out of the domain; 0: the boundary values set.
struct ObcState {
  double vxbc1, vybc1, vzbc1, bxbc1, bybc1, bzbc1, rbc1, ebc1;
  double pfar, rhofar, uxfar, uyfar, uzfar, bxfar, byfar, bzfar, pbc, vnorm;
  double bxbc0, bybc0, bzbc0, rbc0, ebc0;
};

__device__ __forceinline__ void obc_fast(ObcState &b, double gamma) {
  const double c0far = sqrt(gamma * b.pfar / b.rhofar);
  const double ctfar = sqrt((b.byfar * b.byfar + b.bzfar * b.bzfar) / b.rhofar);
  const double cffar = sqrt(c0far * c0far + ctfar * ctfar);
  const double c0 = sqrt(gamma * (gamma - 1.0) * b.ebc1);
  const double ct = sqrt((b.bybc1 * b.bybc1 + b.bzbc1 * b.bzbc1) / b.rbc1);
  const double cf = sqrt(c0 * c0 + ct * ct);
  const double vtest[3] = {b.vnorm + cf, b.vnorm - cf, b.vnorm};
  // each characteristic takes the interior or the far-field state
  const double pin = b.pbc + 0.5 * (b.bybc1 * b.bybc1 + b.bzbc1 * b.bzbc1);
  const double pmin = 0.5 * (b.bybc1 * b.bybc1 + b.bzbc1 * b.bzbc1);
  const double pout = b.pfar + 0.5 * (b.byfar * b.byfar + b.bzfar * b.bzfar);
  const double pmout = 0.5 * (b.byfar * b.byfar + b.bzfar * b.bzfar);
  double pstar[3], pmagstar[3], vstar[3], rhostar[3], bystar[3], bzstar[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const bool in = vtest[i] >= 0.0;
    pstar[i] = in ? pin : pout;
    pmagstar[i] = in ? pmin : pmout;
    vstar[i] = in ? b.vxbc1 : b.uxfar;
    rhostar[i] = in ? b.rbc1 : b.rhofar;
    bystar[i] = in ? b.bybc1 : b.byfar;
    bzstar[i] = in ? b.bzbc1 : b.bzfar;
  }
  b.bxbc0 = 0.5 * (b.bxbc1 + b.bxfar);
  b.bybc0 = 0.5 * (bystar[0] + bystar[1]);
  b.bzbc0 = 0.5 * (bzstar[0] + bzstar[1]);
  double pg = 0.5 * (pstar[0] + pstar[1] + b.rhofar * cffar * (vstar[0] - vstar[1])) - 0.5 * (pmagstar[0] + pmagstar[1]);
  pg = dmax(pg, 0.0);
  const double rhog = rhostar[2] + (pg - (pstar[2] - pmagstar[2])) / (c0far * c0far);
  b.rbc0 = dmax(rhog, NONE_ZERO);
  b.ebc0 = pg / (gamma - 1.0) / b.rbc0;
}

__device__ __forceinline__ void obc_alfven(ObcState &b, double gamma) {
  const double lambdayfar = -b.bxfar * b.byfar, lambdazfar = -b.bxfar * b.bzfar;
  const double c0far = sqrt(gamma * b.pfar / b.rhofar);
  const double cxfar = sqrt(b.bxfar * b.bxfar / b.rhofar);
  const double c0 = sqrt(gamma * (gamma - 1.0) * b.ebc1);
  const double cx = sqrt(b.bxbc1 * b.bxbc1 / b.rbc1);
  const double vtest[5] = {b.vnorm + c0, b.vnorm - c0, b.vnorm + cx, b.vnorm - cx, b.vnorm};
  const double pin = b.pbc + 0.5 * (b.bybc1 * b.bybc1 + b.bzbc1 * b.bzbc1 - b.bxbc1 * b.bxbc1);
  const double pmin = 0.5 * (b.bybc1 * b.bybc1 + b.bzbc1 * b.bzbc1 - b.bxbc1 * b.bxbc1);
  const double pout = b.pfar + 0.5 * (b.byfar * b.byfar + b.bzfar * b.bzfar - b.bxfar * b.bxfar);
  const double pmout = 0.5 * (b.byfar * b.byfar + b.bzfar * b.bzfar - b.bxfar * b.bxfar);
  const double lyin = -b.bybc1 * b.bxbc1, lzin = -b.bzbc1 * b.bxbc1;
  double pstar[5], pmagstar[5], lys[5], lzs[5], uxs[5], uys[5], uzs[5], rhostar[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const bool in = vtest[i] >= 0.0;
    pstar[i] = in ? pin : pout;
    pmagstar[i] = in ? pmin : pmout;
    lys[i] = in ? lyin : lambdayfar;
    lzs[i] = in ? lzin : lambdazfar;
    uxs[i] = in ? b.vxbc1 : b.uxfar;
    uys[i] = in ? b.vybc1 : b.uyfar;
    uzs[i] = in ? b.vzbc1 : b.uzfar;
    rhostar[i] = in ? b.rbc1 : b.rhofar;
  }
  b.bxbc0 = 0.5 * (b.bxbc1 + b.bxfar);
  if (fabs(b.bxbc0) <= NONE_ZERO) b.bxbc0 = NONE_ZERO;
  double lambdag = 0.5 * (lys[2] + lys[3] + b.rhofar * cxfar * (uys[2] - uys[3]));
  b.bybc0 = -lambdag / b.bxbc0;
  lambdag = 0.5 * (lzs[2] + lzs[3] + b.rhofar * cxfar * (uzs[2] - uzs[3]));
  b.bzbc0 = -lambdag / b.bxbc0;
  double pg = 0.5 * (pstar[0] + pstar[1] + b.rhofar * c0far * (uxs[0] - uxs[1])) -
              0.5 * (b.bybc0 * b.bybc0 + b.bzbc0 * b.bzbc0 - b.bxbc0 * b.bxbc0);
  pg = dmax(pg, NONE_ZERO);
  const double rhog = rhostar[4] + (pg - (pstar[4] - pmagstar[4])) / (c0far * c0far);
  b.rbc0 = dmax(rhog, NONE_ZERO);
  b.ebc0 = pg / (gamma - 1.0) / b.rbc0;
}

// index of transverse point (ju, jw) at layer l along the face's axis
__device__ __forceinline__ int obc_idx(const Geo &g, const OpenFace &f, int ju, int jw, int l) {
  int q[3];
  q[f.a] = l;
  q[f.u] = ju;
  q[f.w] = jw;
  return g.idx(q[0], q[1], q[2]);
}

// the pre-corrector energy and density of the face's first interior layer,
// transverse points 0..n+1, into save[0 .. 2 np)
__global__ __launch_bounds__(256) void k_open_save(Geo g, OpenFace f, const real *__restrict__ energy,
                                                   const real *__restrict__ rho, double *__restrict__ save) {
  const int nu = g.n(f.u), nw = g.n(f.w);
  const int np = (nu + 2) * (nw + 2);
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= np) return;
  const int ju = t % (nu + 2), jw = t / (nu + 2);
  const int c = obc_idx(g, f, ju, jw, f.lc);
  save[t] = energy[c];
  save[np + t] = rho[c];
}

// open_bcs of one face (openboundary.f90:30-323) over transverse 0..n+1.
// The first interior layer's energy and density come from the saved
// pre-corrector values where the cell is interior, from the arrays where it
// is a transverse ghost (never touched by the corrector; possibly set by an
// earlier face).
__global__ __launch_bounds__(256) void k_open_bcs(Geo g, OpenFace f, const real *__restrict__ vx,
                                                  const real *__restrict__ vy, const real *__restrict__ vz,
                                                  double *__restrict__ bx, double *__restrict__ by,
                                                  double *__restrict__ bz, real *__restrict__ rho,
                                                  real *__restrict__ energy, const double *__restrict__ save,
                                                  double gamma) {
  const int nu = g.n(f.u), nw = g.n(f.w);
  const int np = (nu + 2) * (nw + 2);
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= np) return;
  const int ju = t % (nu + 2), jw = t / (nu + 2);
  const real *v[3] = {vx, vy, vz};
  double *bf[3] = {bx, by, bz};
  const int cb = obc_idx(g, f, ju, jw, f.lb), cc = obc_idx(g, f, ju, jw, f.lc);
  const int cfc = obc_idx(g, f, ju, jw, f.lfc), cfn = obc_idx(g, f, ju, jw, f.lfn);
  const bool inner = ju >= 1 && ju <= nu && jw >= 1 && jw <= nw;
  const double e1 = inner ? save[t] : (double)energy[cc];
  const double r1 = inner ? save[np + t] : (double)rho[cc];
  ObcState b;
  b.vxbc1 = f.sg * (double)v[f.a][cb];
  b.vybc1 = v[f.ty][cb];
  b.vzbc1 = v[f.tz][cb];
  b.bxbc1 = f.sg * bf[f.a][cb];
  b.bybc1 = bf[f.ty][cc];
  b.bzbc1 = bf[f.tz][cc];
  b.rbc1 = r1;
  b.ebc1 = e1;
  b.pbc = (gamma - 1.0) * e1 * r1;
  b.pfar = (gamma - 1.0) * (double)energy[cfc] * (double)rho[cfc];
  b.uxfar = f.sg * (double)v[f.a][cfn];
  b.uyfar = v[f.ty][cfn];
  b.uzfar = v[f.tz][cfn];
  b.bxfar = f.sg * bf[f.a][cfn];
  b.byfar = bf[f.ty][cfc];
  b.bzfar = bf[f.tz][cfc];
  b.rhofar = rho[cfc];
  b.vnorm = f.sg * (double)v[f.a][cb];
  const double bperp = sqrt(b.byfar * b.byfar + b.bzfar * b.bzfar);
  if (fabs(bperp) <= NONE_ZERO) obc_alfven(b, gamma);
  else obc_fast(b, gamma);
  bf[f.a][obc_idx(g, f, ju, jw, f.lset_n)] = f.sg * b.bxbc0;
  const int cs = obc_idx(g, f, ju, jw, f.lset_c);
  bf[f.ty][cs] = b.bybc0;
  bf[f.tz][cs] = b.bzbc0;
  rho[cs] = b.rbc0;
  energy[cs] = b.ebc0;
}

// Carry the ghost layers of BC_OPEN sides from src to dst over the padded
// box: cells at or below 0 / at or above n+1, nodes at or below -1 / at or
// above n+1 (node = 1).  Used where a kernel writes its new values to a
// second buffer that then swaps roles with the field: a BC_OPEN side is
// never filled, so its ghosts keep the field's previous values.
__global__ __launch_bounds__(BX * BY) void k_copy_open(Geo g, Box b, const real *__restrict__ src,
                                                       real *__restrict__ dst, int node) {
  BOX_INDEX(b, i, j, k);
  const int c[3] = {i, j, k};
  bool ghost = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int n = g.n(a);
    if (g.open_lo[a] && c[a] <= (node ? -1 : 0)) ghost = true;
    if (g.open_hi[a] && c[a] >= n + 1) ghost = true;
  }
  if (ghost) dst[g.idx(i, j, k)] = src[g.idx(i, j, k)];
}
