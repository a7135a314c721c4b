// neutral.hpp -- partially ionised plasma (neutral.f90): the neutral
// fraction of every cell and the Cowling resistivity at the nodes.
//
// Both are per-point and independent, so each thread owns one point; the
// bisection of neutral_fraction runs at most 100 halvings with a per-lane
// exit, and everything is FP64 whatever the storage type (the bisection
// stops at a 1e-8 step, below FP32 resolution).

// exp(x): Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2 (ln2 split in a
// 32-bit head, so k*LN2_HI is exact, and a tail), the Taylor polynomial of
// degree 13 in r by Horner's rule with fused multiply-adds (truncation below
// 0.05 ulp) and the exact scaling by 2^k: within an ulp of a libm exp, and
// the same operations, hence the same bits, as the oracle's.
#define LN2_HI 0x1.62e42feep-1
#define LN2_LO 0x1.a39ef35793c76p-33
#define INV_LN2 0x1.71547652b82fep+0
__device__ __forceinline__ double lare_exp(double x) {
  if (x != x) return x;
  if (x > 710.0) return __builtin_huge_val();
  if (x < -746.0) return 0.0;
  const double k = rint(x * INV_LN2);
  double r = fma(-k, LN2_HI, x);
  r = fma(-k, LN2_LO, r);
  double p = 0x1.6124613a86d09p-33;       // 1/13!
  p = fma(p, r, 0x1.1eed8eff8d898p-29);   // 1/12!
  p = fma(p, r, 0x1.ae64567f544e4p-26);   // 1/11!
  p = fma(p, r, 0x1.27e4fb7789f5cp-22);   // 1/10!
  p = fma(p, r, 0x1.71de3a556c734p-19);   // 1/9!
  p = fma(p, r, 0x1.a01a01a01a01ap-16);   // 1/8!
  p = fma(p, r, 0x1.a01a01a01a01ap-13);   // 1/7!
  p = fma(p, r, 0x1.6c16c16c16c17p-10);   // 1/6!
  p = fma(p, r, 0x1.1111111111111p-7);    // 1/5!
  p = fma(p, r, 0x1.5555555555555p-5);    // 1/4!
  p = fma(p, r, 0x1.5555555555555p-3);    // 1/3!
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

// Normalised constants of setup_neutral / normalise_transport
// (neutral.f90:46-61, normalise.f90:40-49)
struct NeutralArgs {
  double gm1, ip, f_bar, t_bar, r_bar, eta_bar, t_rad;
};

// get_neutral (neutral.f90:140-154): the Saha-type neutral fraction at
// temperature t_v and density rho_v, in the dilute radiation field of
// temperature t_rad (dilution 1/2)
__device__ __forceinline__ double saha_neutral(const NeutralArgs &a, double t_v, double rho_v) {
  const double dilution = 0.5;
  const double bof = 1.0 / (dilution * a.f_bar * a.t_rad * sqrt(t_v)) *
                     lare_exp((0.25 * (t_v / a.t_rad - 1.0) + 1.0) * a.t_bar / t_v);
  const double r = 0.5 * (-1.0 + sqrt(1.0 + a.r_bar * rho_v * bof));
  return r / (1.0 + r);
}

// neutral_fraction (neutral.f90:158-195) over box b (the cell box): the
// temperature by bisection of T = gm1 (e - (1 - xi(T)) ip) / (2 - xi(T))
// between gm1 max((e - ip)/2, tiny) and gm1 e, at most 100 halvings,
// stopping when the step or the residual falls below 1e-8; xi_n from the
// last probe.  MODE TMODE_DIRECT reads energy as stored; TMODE_WRAP reads
// the ghosts of Y through the boundary map (an RKL2 stage value whose
// ghosts were never filled, conduct.f90:400-404; E0 on a BC_OPEN side),
// rho as stored.  An inverted bracket (the reference's STOP, :172-175)
// sets *flag.
template <int MODE>
__global__ __launch_bounds__(BX * BY) void k_neutral(Geo g, Box b, const real *__restrict__ Y,
                                                     const real *__restrict__ E0, const real *__restrict__ rho,
                                                     real *__restrict__ xi, int *__restrict__ flag, NeutralArgs a) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const double rho0 = rho[c];
  const double e0 = ghost_read(g, Y, E0, i, j, k, MODE);
  const double ta1 = a.gm1 * dmax((e0 - a.ip) / 2.0, NONE_ZERO), ta2 = a.gm1 * e0;
  if (ta1 > ta2) *flag = 1;
  double dx = ta2 - ta1, t = ta1, x = t;
  for (int loop = 1; loop <= 100; ++loop) {
    dx = dx / 2.0;
    x = t + dx;
    const double xi_a = saha_neutral(a, x, rho0);
    const double fa = x - a.gm1 * (e0 - (1.0 - xi_a) * a.ip) / (2.0 - xi_a);
    if (fa <= 0.0) t = x;
    if (fabs(dx) < 1.0e-8 || fabs(fa) < 1.0e-8) break;
  }
  xi[c] = saha_neutral(a, x, rho0);
}

// perpendicular_resistivity (neutral.f90:68-136) at nodes 0..n: the volume
// averaged density, the temperature averaged over the eight cells, the
// neutral fraction there, and eta_perp = eta_bar xi/(1-xi) |B|^2 / rho^2 /
// sqrt(T), capped at 100.  The vertex field is the reference's: bxv and byv
// take bz at the upper faces (:97-98).
__global__ __launch_bounds__(BX * BY) void k_eta_perp(Geo g, Box b, const real *__restrict__ rho,
                                                      const real *__restrict__ energy, const real *__restrict__ xi,
                                                      const double *__restrict__ bx, const double *__restrict__ by,
                                                      const double *__restrict__ bz, real *__restrict__ eta_perp,
                                                      NeutralArgs a) {
  BOX_INDEX(b, i, j, k);
  const int X = 1, Y = g.sy, Z = g.sz;
  const int c = g.idx(i, j, k);
  const int off[8] = {0, X, Y, X + Y, Z, X + Z, Y + Z, X + Y + Z};
  const int di[8] = {0, 1, 0, 1, 0, 1, 0, 1}, dj[8] = {0, 0, 1, 1, 0, 0, 1, 1}, dk[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  double cvs[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) cvs[q] = g.cv(i + di[q], j + dj[q], k + dk[q]);
  double rho_v = (double)rho[c] * cvs[0];
#pragma unroll
  for (int q = 1; q < 8; ++q) rho_v = rho_v + (double)rho[c + off[q]] * cvs[q];
  double sc = cvs[0];
#pragma unroll
  for (int q = 1; q < 8; ++q) sc = sc + cvs[q];
  rho_v = rho_v / sc;
  double bxv = bx[c] + bx[c + Y] + bz[c + Z] + bz[c + Y + Z];
  double byv = by[c] + by[c + X] + bz[c + Z] + bz[c + X + Z];
  double bzv = bz[c] + bz[c + X] + bz[c + Y] + bz[c + X + Y];
  bxv = 0.25 * bxv;
  byv = 0.25 * byv;
  bzv = 0.25 * bzv;
  const double bfieldsq = bxv * bxv + byv * byv + bzv * bzv;
  double t_v = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double x = xi[c + off[q]];
    t_v = t_v + ((double)energy[c + off[q]] - (1.0 - x) * a.ip) / (2.0 - x);
  }
  t_v = t_v * a.gm1 / 8.0;
  const double xi_v = saha_neutral(a, t_v, rho_v);
  const double f = dmax(1.0 - xi_v, NONE_ZERO);
  double ep = 0.0;
  if (f > 0) ep = a.eta_bar * xi_v / f * bfieldsq / (rho_v * rho_v) / sqrt(t_v);
  eta_perp[c] = dmin(ep, 100.0);
}
