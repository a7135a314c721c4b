// lare_dev.hip -- C-ABI runtime of the MI355X Lare3d hot path.
//
// One translation unit: the kernels live in the headers included below.
// The runtime owns every device array from lare_dev_init to
// lare_dev_finalize, queues all work on one HIP stream, and synchronises
// only where the reference's control flow needs a device scalar on the
// host (dt after set_dt, the RKL2 stage count, energy diagnostics).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/lare3d_dev.h"
#include "bc.hpp"
#include "common.hpp"
#include "conduct.hpp"
#include "lagran.hpp"
#include "remap.hpp"
#include "surrogate.hpp"

namespace {

char g_err[512] = "no error";

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(-(int)e_, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

// ----------------------------------------------------------------------------
// kernel timing with HIP events on the library stream
// ----------------------------------------------------------------------------
struct ProfRec {
  const char *name;
  hipEvent_t a, b;
};
struct ProfAcc {
  std::string name;
  int count;
  double ms;
};

struct Dev {
  bool ready = false, loaded = false;
  int n[3];
  int sy, sz;
  size_t ntot;
  int bc[6];
  int per[3], user[3];
  lare_params p;
  double tfac, visc2_norm;
  hipStream_t stream = nullptr;
  double *metric = nullptr;  // 7 arrays of length nmax+5, index i+2
  Geo g;
  // state
  double *rho, *rho_alt, *energy, *vx, *vy, *vz, *vx1, *vy1, *vz1, *bx, *by, *bz;
  double *eta, *visc3, *visc_dep, *ohmic_dep;
  // lagrangian scratch
  double *rho_v, *cs_v, *al1, *al2, *al3, *visc_heat, *p0, *pp, *bpx, *bpy, *bpz;
  double *fxr, *fyr, *fzr, *curlb, *qx, *qy, *qz;
  // remap scratch
  double *cv1, *cv2, *db1, *F, *dm, *Fx, *Fy, *Fz, *rho_vr, *cv1v, *cv2v, *dmv;
  // conduction scratch
  double *larsen, *Lc, *Yb[3];
  // reductions
  double *partial = nullptr;
  size_t partial_cap = 0;
  double *dscal = nullptr;   // device scalars: [0..3] last reduce, [8] tvh acc, [9] toh acc
  double *hscal = nullptr;   // pinned host mirror
  std::vector<double *> allocs;
  // host scalars
  double dt = 0.0, dt2 = 0.0, dtr = 0.0, time = 0.0;
  int n_s_stages = 0, peak_substeps = 0, last_substeps = 0, first_set_dt = 1, step = 0;
  // profiling
  bool prof = false;
  std::vector<ProfRec> pending;
  std::vector<hipEvent_t> pool;
  std::vector<ProfAcc> acc;
};

Dev S;

hipEvent_t ev_get() {
  if (!S.pool.empty()) {
    hipEvent_t e = S.pool.back();
    S.pool.pop_back();
    return e;
  }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}

void prof_flush() {
  if (S.pending.empty()) return;
  (void)hipStreamSynchronize(S.stream);
  for (auto &r : S.pending) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    bool found = false;
    for (auto &a : S.acc)
      if (a.name == r.name) {
        a.count++;
        a.ms += ms;
        found = true;
        break;
      }
    if (!found) S.acc.push_back({r.name, 1, (double)ms});
    S.pool.push_back(r.a);
    S.pool.push_back(r.b);
  }
  S.pending.clear();
}

struct ProfScope {
  const char *name;
  hipEvent_t a = nullptr;
  explicit ProfScope(const char *n) : name(n) {
    if (S.prof) {
      a = ev_get();
      (void)hipEventRecord(a, S.stream);
    }
  }
  ~ProfScope() {
    if (S.prof) {
      hipEvent_t b = ev_get();
      (void)hipEventRecord(b, S.stream);
      S.pending.push_back({name, a, b});
      if (S.pending.size() > 4096) prof_flush();
    }
  }
};

#define LAUNCH(name, kern, grid, block, ...)                                   \
  do {                                                                         \
    ProfScope ps_(name);                                                       \
    hipLaunchKernelGGL(kern, grid, block, 0, S.stream, __VA_ARGS__);           \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) return fail(-(int)e_, "launch %s: %s", name, hipGetErrorString(e_)); \
  } while (0)

#define LAUNCH_BOX(name, kern, box, ...) LAUNCH(name, kern, box_grid(box), box_block(), S.g, box, __VA_ARGS__)

int nx() { return S.n[0]; }
int ny() { return S.n[1]; }
int nz() { return S.n[2]; }
Box full_box() { return mkbox(-2, nx() + 2, -2, ny() + 2, -2, nz() + 2); }
Box cell_box() { return mkbox(-1, nx() + 2, -1, ny() + 2, -1, nz() + 2); }
Box interior() { return mkbox(1, nx(), 1, ny(), 1, nz()); }
Box nodes0n() { return mkbox(0, nx(), 0, ny(), 0, nz()); }

// reduction over partials of a grid launched from index 0 on every axis
dim3 grid_from0(int x1, int y1, int z1) { return box_grid(mkbox(0, x1, 0, y1, 0, z1)); }

// Two-level reduction of the per-block partials (rows of nv values; bit q of
// minmask selects min, else sum).  Level 1: RED_ROWS blocks each fold one
// contiguous chunk of rows; level 2: one block folds the RED_ROWS results and
// stores (or adds) the nv scalars.
#define RED_ROWS 512
#define RED_MAXV 4

__device__ __forceinline__ void fold_rows(const double *__restrict__ rows, int r0, int r1, int nv, int minmask,
                                          double *__restrict__ out, int accumulate) {
  __shared__ double red[4][RED_MAXV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int q = 0; q < nv; ++q) {
    const bool isMin = (minmask >> q) & 1;
    double v = isMin ? 1.7976931348623157e+308 : 0.0;
    for (int r = r0 + (int)threadIdx.x; r < r1; r += 256) {
      const double x = rows[(size_t)r * nv + q];
      v = isMin ? dmin(v, x) : v + x;
    }
    v = isMin ? wave_min(v) : wave_sum(v);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    const int q = threadIdx.x;
    const bool isMin = (minmask >> q) & 1;
    double r = red[0][q];
    for (int w = 1; w < 4; ++w) r = isMin ? dmin(r, red[w][q]) : r + red[w][q];
    if (accumulate) r = out[q] + r;
    out[q] = r;
  }
}

__global__ __launch_bounds__(256) void k_reduce_rows(const double *__restrict__ partial, int nrows, int nv, int minmask,
                                                     int chunk, double *__restrict__ level1) {
  const int r0 = (int)blockIdx.x * chunk;
  const int r1 = r0 + chunk < nrows ? r0 + chunk : nrows;
  fold_rows(partial, r0, r1, nv, minmask, level1 + (size_t)blockIdx.x * nv, 0);
}

__global__ __launch_bounds__(256) void k_final_reduce(const double *__restrict__ level1, int nrows, int nv, int minmask,
                                                      int accumulate, double *__restrict__ out) {
  fold_rows(level1, 0, nrows, nv, minmask, out, accumulate);
}

int final_reduce(dim3 grid, int nv, int minmask, int slot, int accumulate) {
  const int nb = (int)(grid.x * grid.y * grid.z);
  if (nv > RED_MAXV || (size_t)nb * nv > S.partial_cap)
    return fail(LARE_E_ARG, "reduction of %d x %d values exceeds the partial buffer", nb, nv);
  const int chunk = (nb + RED_ROWS - 1) / RED_ROWS;
  const int rows = (nb + chunk - 1) / chunk;  // every level-1 block has >= 1 row
  double *level1 = S.partial + S.partial_cap;
  LAUNCH("reduce_rows", k_reduce_rows, dim3(rows), dim3(256), S.partial, nb, nv, minmask, chunk, level1);
  LAUNCH("final_reduce", k_final_reduce, dim3(1), dim3(256), level1, rows, nv, minmask, accumulate, S.dscal + slot);
  return 0;
}

int read_scalars(int slot, int count, double *out) {
  HIPCHK(hipMemcpyAsync(S.hscal, S.dscal + slot, sizeof(double) * count, hipMemcpyDeviceToHost, S.stream));
  HIPCHK(hipStreamSynchronize(S.stream));
  for (int q = 0; q < count; ++q) out[q] = S.hscal[q];
  return 0;
}

// ----------------------------------------------------------------------------
// boundary groups
// ----------------------------------------------------------------------------
enum FieldKind { FK_CELL, FK_NODE, FK_BX, FK_BY, FK_BZ };

struct BcField {
  double *a;
  FieldKind kind;
};

int op_for(FieldKind k, int axis, bool periodic) {
  switch (k) {
    case FK_CELL: return periodic ? BC_PER_CELL : BC_MIR_CELL;
    case FK_NODE: return periodic ? BC_PER_NODE : BC_ZERO_NODE;
    case FK_BX: return axis == 0 ? (periodic ? BC_PER_FACE : BC_MIR_FACE) : (periodic ? BC_PER_CELL : BC_MIR_CELL);
    case FK_BY: return axis == 1 ? (periodic ? BC_PER_FACE : BC_MIR_FACE) : (periodic ? BC_PER_CELL : BC_MIR_CELL);
    case FK_BZ: return axis == 2 ? (periodic ? BC_PER_FACE : BC_MIR_FACE) : (periodic ? BC_PER_CELL : BC_MIR_CELL);
  }
  return BC_NONE;
}

int launch_bc_axis(const BcSet &s, int axis, const char *name) {
  const int u = axis == 0 ? 1 : 0, w = axis == 2 ? 1 : 2;
  dim3 grid((unsigned)((S.n[u] + 5 + 63) / 64), (unsigned)(S.n[w] + 5));
  LAUNCH(name, k_bc_axis, grid, dim3(64), S.g, s, axis);
  return 0;
}

// exchanges on periodic axes (x, y, z), then user fills (x, y, z)
int apply_bcs(const BcField *f, int nf, const char *name) {
  for (int pass = 0; pass < 2; ++pass)
    for (int ax = 0; ax < 3; ++ax) {
      const bool periodic = S.per[ax];
      if ((pass == 0) != periodic) continue;
      BcSet s;
      s.nf = nf;
      for (int q = 0; q < nf; ++q) {
        s.f[q] = f[q].a;
        s.op[q] = op_for(f[q].kind, ax, periodic);
      }
      int rc = launch_bc_axis(s, ax, name);
      if (rc) return rc;
    }
  return 0;
}

int bfield_bcs(double *bx, double *by, double *bz) {
  BcField f[3] = {{bx, FK_BX}, {by, FK_BY}, {bz, FK_BZ}};
  return apply_bcs(f, 3, "bc_bfield");
}
int energy_bcs() {
  BcField f[1] = {{S.energy, FK_CELL}};
  return apply_bcs(f, 1, "bc_energy");
}
int boundary_conditions() {
  BcField f[8] = {{S.bx, FK_BX},     {S.by, FK_BY},   {S.bz, FK_BZ},   {S.energy, FK_CELL},
                  {S.rho, FK_CELL},  {S.vx, FK_NODE}, {S.vy, FK_NODE}, {S.vz, FK_NODE}};
  return apply_bcs(f, 8, "bc_all");
}

// ----------------------------------------------------------------------------
// conduction
// ----------------------------------------------------------------------------
int calc_s_stages() {
  dim3 gr = grid_from0(nx(), ny(), nz());
  LAUNCH("larsen", k_larsen, gr, box_block(), S.g, S.energy, S.rho, S.larsen, S.partial, S.p.flux_limiter,
         S.p.kappa_0, S.tfac, S.p.heat_flux_limiter);
  int rc = final_reduce(gr, 1, 1, 0, 0);
  if (rc) return rc;
  double dtp;
  rc = read_scalars(0, 1, &dtp);
  if (rc) return rc;
  // the reference starts its minimum at 1e10 (conduct.f90:110)
  double dt_parab = dtp < 1.0e10 ? dtp : 1.0e10;
  dt_parab = 0.9 * dt_parab / sqrt(3.0);
  const double stages = 0.5 * (sqrt(9.0 + 16.0 * (S.dt / dt_parab)) - 1.0);
  int ns = (int)ceil(stages);
  if (ns % 2 == 0) ns = ns + 1;
  S.n_s_stages = ns;
  return 0;
}

// RKL2 super-step; delta_mode = 1 inside lagrangian_step
int heat_conduct_sts2(int delta_mode) {
  const int s = S.n_s_stages;
  Box in = interior();
  if (s <= 1) {  // the reference's Y(3) is never written: energy unchanged
    if (delta_mode) LAUNCH_BOX("energy_max", k_energy_max, cell_box(), S.energy);
    return 0;
  }
  std::vector<double> b(s + 1), a(s + 1), mu_t(s + 1), mu(s + 1), nu(s + 1), gt(s + 1);
  const double ds = (double)s;
  const double omega_1 = 4.0 / (ds * ds + ds - 2.0);
  for (int j = 0; j <= 2; ++j) b[j] = 1.0 / 3.0;
  for (int j = 3; j <= s; ++j) {
    const double dj = (double)j;
    b[j] = (dj * dj + dj - 2.0) / (2.0 * dj * (dj + 1.0));
  }
  for (int j = 0; j <= s; ++j) a[j] = 1.0 - b[j];
  mu_t[1] = omega_1 / 3.0;
  for (int j = 2; j <= s; ++j) {
    const double dj = (double)j;
    mu_t[j] = ((2.0 * dj - 1.0) / dj) * omega_1 * (b[j] / b[j - 1]);
    mu[j] = ((2.0 * dj - 1.0) / dj) * (b[j] / b[j - 1]);
    nu[j] = -1.0 * ((dj - 1.0) / dj) * (b[j] / b[j - 2]);
    gt[j] = -1.0 * a[j - 1] * mu_t[j];
  }
  CondArgs ca = {S.p.kappa_0, S.tfac, S.dt};
  double *E0 = S.energy;
  double *outp = S.Yb[0];
  LAUNCH_BOX("rkl2_stage1", k_rkl2_stage1, in, E0, S.rho, S.bx, S.by, S.bz, S.larsen, S.Lc, outp, ca, mu_t[1]);
  double *y1 = E0, *y2 = outp;
  int next = 1;
  for (int j = 2; j <= s; ++j) {
    double *yo = S.Yb[next];
    next = (next + 1) % 3;
    const int mode = j == 2 ? TMODE_STALE : TMODE_WRAP;
    LAUNCH_BOX("rkl2_stage", k_rkl2_stage, in, y2, y1, E0, S.rho, S.bx, S.by, S.bz, S.larsen, S.Lc, yo, ca, mode,
               mu[j], nu[j], (1.0 - mu[j]) - nu[j], mu_t[j], gt[j]);
    y1 = y2;
    y2 = yo;
  }
  LAUNCH_BOX("rkl2_final", k_rkl2_final, cell_box(), y2, S.energy, delta_mode);
  return 0;
}

// ----------------------------------------------------------------------------
// remap
// ----------------------------------------------------------------------------
int strides(int ax) { return ax == 0 ? 1 : (ax == 1 ? S.sy : S.sz); }

template <int AX>
int ct_pair(double *bxa, double *bya, double *bza, const double *vel, double dt) {
  const int N0 = nx(), N1 = ny(), N2 = nz();
  Box fb = mkbox(-1, N0 + 1, -1, N1 + 1, -1, N2 + 1);
  Box ub = mkbox(0, N0 + 1, 0, N1 + 1, 0, N2 + 1);
  double *B[3] = {bxa, bya, bza};
  // (transverse field t, velocity-average axis, geometry-average axis, box_t, box_a) x 2
  int t1, va1, ga1, t2, va2, ga2;
  Box bt1 = ub, ba1 = ub, bt2 = ub, ba2 = ub;
  if (AX == 0) {
    t1 = 1; va1 = 2; ga1 = 1;  // vx_by_flux
    t2 = 2; va2 = 1; ga2 = 2;  // vx_bz_flux
    bt2 = mkbox(1, N0, 1, N1, 0, N2);
    ba2 = mkbox(0, N0, 0, N1 + 1, 0, N2 + 1);
  } else if (AX == 1) {
    t1 = 0; va1 = 2; ga1 = 0;  // vy_bx_flux
    t2 = 2; va2 = 0; ga2 = 2;  // vy_bz_flux
    ba2 = mkbox(1, N0, 0, N1, 1, N2);
  } else {
    t1 = 0; va1 = 1; ga1 = 0;  // vz_bx_flux
    t2 = 1; va2 = 0; ga2 = 1;  // vz_by_flux
  }
  LAUNCH_BOX("ct_flux", k_ct_flux<AX>, fb, B[t1], vel, S.db1, S.F, strides(va1), strides(ga1), dt);
  LAUNCH_BOX("ct_update", k_ct_update, ub, S.F, B[t1], B[AX], strides(AX), strides(t1), bt1, ba1);
  LAUNCH_BOX("ct_flux", k_ct_flux<AX>, fb, B[t2], vel, S.db1, S.F, strides(va2), strides(ga2), dt);
  LAUNCH_BOX("ct_update", k_ct_update, ub, S.F, B[t2], B[AX], strides(AX), strides(t2), bt2, ba2);
  return 0;
}

// one predictor-mode sweep (B only, no flag reset, no halo refresh)
template <int AX>
int sweep_predictor(double dt, bool last) {
  LAUNCH_BOX("geom", k_geom<AX>, cell_box(), S.vx, S.vy, S.vz, S.cv1, S.cv2, S.db1, dt, 1.0, 1.0, 1.0,
             last ? 1 : 0, 0);
  const double *vel = AX == 0 ? S.vx : (AX == 1 ? S.vy : S.vz);
  return ct_pair<AX>(S.bpx, S.bpy, S.bpz, vel, dt);
}

template <int AX>
int sweep_full(double dt, double pass[3]) {
  const int N[3] = {nx(), ny(), nz()};
  const double *vel = AX == 0 ? S.vx1 : (AX == 1 ? S.vy1 : S.vz1);
  LAUNCH_BOX("geom", k_geom<AX>, cell_box(), S.vx1, S.vy1, S.vz1, S.cv1, S.cv2, S.db1, dt, pass[0], pass[1],
             pass[2], 1, 1);
  int rc = ct_pair<AX>(S.bx, S.by, S.bz, vel, dt);
  if (rc) return rc;
  int hi[3] = {N[0] + 1, N[1] + 1, N[2] + 1};
  hi[AX] = N[AX];
  Box mb = mkbox(0, hi[0], 0, hi[1], 0, hi[2]);
  LAUNCH_BOX("mass_flux", k_mass_flux<AX>, mb, S.rho, vel, S.db1, S.dm, dt);
  {
    BcSet s;
    s.nf = 1;
    s.f[0] = S.dm;
    s.op[0] = S.per[AX] ? BC_PER_CELL : BC_DM_CLAMP;
    rc = launch_bc_axis(s, AX, "bc_dm");
    if (rc) return rc;
  }
  LAUNCH_BOX("energy_flux", k_energy_flux<AX>, nodes0n(), S.energy, S.rho, vel, S.db1, S.dm, S.F, dt);
  LAUNCH_BOX("rho_energy_update", k_rho_energy_update<AX>, interior(), S.rho, S.rho_alt, S.energy, S.cv1, S.cv2,
             S.dm, S.F);
  {
    int lo[3] = {0, 0, 0}, h2[3] = {N[0], N[1], N[2]};
    lo[AX] = -1;
    h2[AX] = N[AX] + 1;
    Box vb = mkbox(lo[0], h2[0], lo[1], h2[1], lo[2], h2[2]);
    LAUNCH_BOX("vertex", k_vertex<AX>, vb, S.rho, S.cv1, S.cv2, S.dm, S.rho_vr, S.cv1v, S.cv2v, S.dmv);
    h2[AX] = N[AX];
    Box fb = mkbox(lo[0], h2[0], lo[1], h2[1], lo[2], h2[2]);
    LAUNCH_BOX("mom_flux", k_mom_flux<AX>, fb, S.vx, S.vy, S.vz, vel, S.db1, S.rho_vr, S.dmv, S.Fx, S.Fy, S.Fz, dt);
  }
  LAUNCH_BOX("mom_update", k_mom_update<AX>, nodes0n(), S.vx, S.vy, S.vz, S.rho_vr, S.cv1v, S.cv2v, S.dmv, S.Fx,
             S.Fy, S.Fz);
  double *t = S.rho;
  S.rho = S.rho_alt;
  S.rho_alt = t;
  rc = boundary_conditions();
  if (rc) return rc;
  pass[AX] = 0.0;
  return 0;
}

const int kStrang[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}};

int eulerian_remap(int step, bool predictor) {
  const double dt = S.dt;
  const int c = ((step % 6) + 6) % 6;
  if (predictor) {
    LAUNCH_BOX("bscale", k_bscale, full_box(), S.bx, S.by, S.bz, S.bpx, S.bpy, S.bpz, 1);
    for (int q = 0; q < 3; ++q) {
      const int ax = kStrang[c][q];
      const bool last = q == 2;
      int rc = ax == 0 ? sweep_predictor<0>(dt, last) : (ax == 1 ? sweep_predictor<1>(dt, last) : sweep_predictor<2>(dt, last));
      if (rc) return rc;
    }
    LAUNCH_BOX("bscale", k_bscale, full_box(), S.bpx, S.bpy, S.bpz, S.bpx, S.bpy, S.bpz, -1);
    return bfield_bcs(S.bpx, S.bpy, S.bpz);
  }
  double pass[3] = {1.0, 1.0, 1.0};
  LAUNCH_BOX("bscale", k_bscale, full_box(), S.bx, S.by, S.bz, S.bx, S.by, S.bz, 1);
  for (int q = 0; q < 3; ++q) {
    const int ax = kStrang[c][q];
    int rc = ax == 0 ? sweep_full<0>(dt, pass) : (ax == 1 ? sweep_full<1>(dt, pass) : sweep_full<2>(dt, pass));
    if (rc) return rc;
  }
  LAUNCH_BOX("bscale", k_bscale, full_box(), S.bx, S.by, S.bz, S.bx, S.by, S.bz, -1);
  return bfield_bcs(S.bx, S.by, S.bz);
}

// ----------------------------------------------------------------------------
// lagrangian step
// ----------------------------------------------------------------------------
int eta_calc() {
  if (S.p.resistive_mhd) {
    LAUNCH_BOX("eta", k_eta, mkbox(-1, nx() + 1, -1, ny() + 1, -1, nz() + 1), S.bx, S.by, S.bz, S.eta,
               S.p.eta_background, S.p.eta0, S.p.j_max);
  } else {
    LAUNCH_BOX("fill", k_fill, cell_box(), S.eta, 0.0);
  }
  return 0;
}

LagArgs lag_args() {
  LagArgs a;
  a.gamma = S.p.gamma;
  a.visc1 = S.p.visc1;
  a.visc2_norm = S.visc2_norm;
  a.dt = S.dt;
  a.dt2 = S.dt2;
  a.eta_bg = S.p.eta_background;
  a.eta0 = S.p.eta0;
  a.j_max = S.p.j_max;
  a.resistive = S.p.resistive_mhd;
  return a;
}

int set_dt_and_stages(double dt_local, double dtr_local) {
  if (S.first_set_dt) {
    S.first_set_dt = 0;
    if (S.p.restart) {
      S.dt = S.p.dt_from_restart;
      return 0;
    }
  }
  S.dt = 0.9 * dt_local;
  S.dtr = 0.9 * dtr_local;
  if (S.p.conduction) {
    int rc = calc_s_stages();
    if (rc) return rc;
    const int ss_limit = 60;
    if (S.n_s_stages >= ss_limit) {
      const double f = (double)(2 * ss_limit * ss_limit - 9) / (double)(2 * S.n_s_stages * S.n_s_stages - 9);
      S.dt = S.dt * f;
      S.dtr = S.dtr * f;
    }
  }
  S.time = S.time + S.dt;
  return 0;
}

// learned conduction surrogate (surrogate_rt.hpp)
int sur_enabled();
int sur_conduct_step(int delta_mode, double *res_out, int *acc_out);

int lagrangian_step() {
  const int N0 = nx(), N1 = ny(), N2 = nz();
  LagArgs a = lag_args();  // dt = previous step's dt (Q1)
  LAUNCH_BOX("vertex_prep", k_vertex_prep, mkbox(-1, N0 + 1, -1, N1 + 1, -1, N2 + 1), S.rho, S.energy, S.bx, S.by,
             S.bz, S.rho_v, S.cs_v, S.p.gamma);
  if (S.p.use_viscous_damping) {
    LAUNCH_BOX("damp_q", k_damp_q, nodes0n(), S.vx, S.vy, S.vz, S.visc3, S.qx, S.qy, S.qz);
    LAUNCH_BOX("damp_apply", k_damp_apply, nodes0n(), S.vx, S.vy, S.vz, S.qx, S.qy, S.qz, S.dt);
  }
  LAUNCH_BOX("alpha", k_alpha, mkbox(-1, N0 + 1, -1, N1 + 2, 0, N2 + 2), S.vx, S.vy, S.vz, S.rho_v, S.cs_v, S.al1,
             S.al2, S.al3, a);
  dim3 gr = grid_from0(N0 + 1, N1 + 1, N2 + 1);
  LAUNCH("visc_dt", k_visc_dt, gr, box_block(), S.g, S.vx, S.vy, S.vz, S.rho, S.energy, S.bx, S.by, S.bz, S.eta,
         S.al1, S.al2, S.al3, S.visc_heat, S.p0, S.partial, a);
  int rc = final_reduce(gr, 2, 3, 0, 0);
  if (rc) return rc;
  double mins[2];
  rc = read_scalars(0, 2, mins);
  if (rc) return rc;
  rc = set_dt_and_stages(mins[0], mins[1]);
  if (rc) return rc;
  S.dt2 = 0.5 * S.dt;
  S.last_substeps = 0;

  if (S.p.resistive_mhd) {
    const double dt_sub = S.dtr;
    const int substeps = (int)(S.dt / dt_sub) + 1;
    if (substeps > S.peak_substeps) S.peak_substeps = substeps;
    S.last_substeps = substeps;
    const double actual_dt = S.dt;
    S.dt = S.dt / (double)substeps;
    for (int sc = 1; sc <= substeps; ++sc) {
      rc = eta_calc();
      if (rc) return rc;
      dim3 g2 = grid_from0(N0, N1, N2);
      LAUNCH("rkstep", k_rkstep, g2, box_block(), S.g, S.bx, S.by, S.bz, S.eta, S.fxr, S.fyr, S.fzr, S.curlb,
             S.partial, S.dt);
      rc = final_reduce(g2, 1, 0, 9, 1);
      if (rc) return rc;
      LAUNCH_BOX("bstep", k_bstep, nodes0n(), S.bx, S.by, S.bz, S.fxr, S.fyr, S.fzr, S.dt);
      rc = bfield_bcs(S.bx, S.by, S.bz);
      if (rc) return rc;
      LAUNCH_BOX("ohmic", k_ohmic, cell_box(), S.energy, S.ohmic_dep, S.rho, S.curlb, S.dt);
      rc = energy_bcs();
      if (rc) return rc;
    }
    S.dt = actual_dt;
  }

  if (S.p.conduction) {
    if (sur_enabled()) {
      rc = sur_conduct_step(1, nullptr, nullptr);
      if (rc) return rc;
    } else {
      rc = calc_s_stages();
      if (rc) return rc;
      rc = heat_conduct_sts2(1);
      if (rc) return rc;
    }
  } else {
    LAUNCH_BOX("energy_max", k_energy_max, cell_box(), S.energy);
  }

  // predictor-corrector (lagran.F90:208-478)
  const double dt_full = S.dt;
  S.dt = 0.5 * dt_full;
  rc = eulerian_remap(S.step, true);
  S.dt = dt_full;
  if (rc) return rc;
  LAUNCH_BOX("pred_pressure", k_pred_pressure, mkbox(0, N0 + 1, 0, N1 + 1, 0, N2 + 1), S.energy, S.rho, S.cv1,
             S.visc_heat, S.p0, S.pp, S.p.gamma, S.dt2);
  LAUNCH_BOX("force", k_force, nodes0n(), S.vx, S.vy, S.vz, S.pp, S.bpx, S.bpy, S.bpz, S.rho_v, S.al1, S.al2, S.al3,
             S.vx1, S.vy1, S.vz1, S.dt2);
  {
    BcField f[3] = {{S.vx1, FK_NODE}, {S.vy1, FK_NODE}, {S.vz1, FK_NODE}};
    rc = apply_bcs(f, 3, "bc_v1");
    if (rc) return rc;
  }
  dim3 g3 = box_grid(interior());
  LAUNCH("shock_corr", k_shock_corr, g3, box_block(), S.g, S.vx, S.vy, S.vz, S.vx1, S.vy1, S.vz1, S.al1, S.al2,
         S.al3, S.pp, S.energy, S.rho, S.visc_dep, S.partial, S.dt);
  rc = final_reduce(g3, 1, 0, 8, 1);
  if (rc) return rc;
  LAUNCH_BOX("vupdate", k_vupdate, nodes0n(), S.vx, S.vy, S.vz, S.vx1, S.vy1, S.vz1);
  {
    BcField f[5] = {{S.energy, FK_CELL}, {S.rho, FK_CELL}, {S.vx, FK_NODE}, {S.vy, FK_NODE}, {S.vz, FK_NODE}};
    rc = apply_bcs(f, 5, "bc_lagrangian");
    if (rc) return rc;
  }
  return 0;
}

// ----------------------------------------------------------------------------
// memory
// ----------------------------------------------------------------------------
void sur_free();
void free_all() {
  sur_free();
  if (S.stream) (void)hipStreamSynchronize(S.stream);
  for (double *p : S.allocs) (void)hipFree(p);
  S.allocs.clear();
  if (S.metric) (void)hipFree(S.metric);
  S.metric = nullptr;
  if (S.partial) (void)hipFree(S.partial);
  S.partial = nullptr;
  if (S.dscal) (void)hipFree(S.dscal);
  S.dscal = nullptr;
  if (S.hscal) (void)hipHostFree(S.hscal);
  S.hscal = nullptr;
  for (auto &r : S.pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  S.pending.clear();
  for (auto e : S.pool) (void)hipEventDestroy(e);
  S.pool.clear();
  S.acc.clear();
  S.ready = S.loaded = false;
}

int dalloc(double **p) {
  hipError_t e = hipMalloc((void **)p, sizeof(double) * S.ntot);
  if (e != hipSuccess) return fail(LARE_E_NOMEM, "hipMalloc of %zu doubles failed: %s", S.ntot, hipGetErrorString(e));
  S.allocs.push_back(*p);
  e = hipMemsetAsync(*p, 0, sizeof(double) * S.ntot, S.stream);
  if (e != hipSuccess) return fail(-(int)e, "hipMemset failed");
  return 0;
}

// Fortran array bounds (lower) per field kind
void lower_of(FieldKind k, int lo[3]) {
  lo[0] = lo[1] = lo[2] = (k == FK_NODE) ? -2 : -1;
  if (k == FK_BX) lo[0] = -2;
  if (k == FK_BY) lo[1] = -2;
  if (k == FK_BZ) lo[2] = -2;
}

std::vector<double> g_stage;

int xfer(double *dev, double *host, FieldKind k, bool to_dev) {
  int lo[3];
  lower_of(k, lo);
  const long ex = nx() + 2 - lo[0] + 1, ey = ny() + 2 - lo[1] + 1;
  if (g_stage.size() != S.ntot) g_stage.assign(S.ntot, 0.0);
  if (to_dev) {
    HIPCHK(hipMemcpy(g_stage.data(), dev, sizeof(double) * S.ntot, hipMemcpyDeviceToHost));
    for (int kk = lo[2]; kk <= nz() + 2; ++kk)
      for (int jj = lo[1]; jj <= ny() + 2; ++jj)
        for (int ii = lo[0]; ii <= nx() + 2; ++ii)
          g_stage[S.g.idx(ii, jj, kk)] = host[((long)(kk - lo[2]) * ey + (jj - lo[1])) * ex + (ii - lo[0])];
    HIPCHK(hipMemcpy(dev, g_stage.data(), sizeof(double) * S.ntot, hipMemcpyHostToDevice));
  } else {
    HIPCHK(hipMemcpy(g_stage.data(), dev, sizeof(double) * S.ntot, hipMemcpyDeviceToHost));
    for (int kk = lo[2]; kk <= nz() + 2; ++kk)
      for (int jj = lo[1]; jj <= ny() + 2; ++jj)
        for (int ii = lo[0]; ii <= nx() + 2; ++ii)
          host[((long)(kk - lo[2]) * ey + (jj - lo[1])) * ex + (ii - lo[0])] = g_stage[S.g.idx(ii, jj, kk)];
  }
  return 0;
}

}  // namespace

// ============================================================================
// C-ABI
// ============================================================================
extern "C" {

const char *lare_dev_last_error(void) { return g_err; }

int lare_dev_init(const lare_grid *grid, const lare_params *params) {
  if (!grid || !params) return fail(LARE_E_ARG, "lare_dev_init: NULL argument");
  free_all();
  const int n[3] = {grid->nx, grid->ny, grid->nz};
  const int bc[6] = {grid->xbc_min, grid->xbc_max, grid->ybc_min, grid->ybc_max, grid->zbc_min, grid->zbc_max};
  for (int a = 0; a < 3; ++a) {
    if (n[a] < 4) return fail(LARE_E_ARG, "grid size %d on axis %d is below 4", n[a], a);
    const int lo = bc[2 * a], hi = bc[2 * a + 1];
    if (lo != hi || (lo != LARE_BC_PERIODIC && lo != LARE_BC_USER))
      return fail(LARE_E_ARG, "axis %d: boundary codes (%d, %d) unsupported (need both periodic or both BC_USER)", a, lo, hi);
  }
  if (params->precision != LARE_PREC_FP64) return fail(LARE_E_ARG, "precision mode %d not available", params->precision);
  if (!grid->dxb || !grid->dxc || !grid->dyb || !grid->dyc || !grid->dzb || !grid->dzc)
    return fail(LARE_E_ARG, "lare_dev_init: metric arrays missing");
  memcpy(S.n, n, sizeof(n));
  memcpy(S.bc, bc, sizeof(bc));
  for (int a = 0; a < 3; ++a) {
    S.per[a] = bc[2 * a] == LARE_BC_PERIODIC;
    S.user[a] = bc[2 * a] == LARE_BC_USER;
  }
  S.p = *params;
  S.tfac = (params->gamma - 1.0) / (2.0 - params->xi_n);
  S.visc2_norm = 0.25 * (params->gamma + 1.0) * params->visc2;
  S.sy = ((n[0] + 19 + 15) / 16) * 16;
  S.sz = S.sy * (n[1] + 5);
  S.ntot = (size_t)S.sz * (size_t)(n[2] + 5);
  if (S.ntot > (size_t)2147483647) return fail(LARE_E_ARG, "grid too large for 32-bit indexing");
  if (!S.stream) HIPCHK(hipStreamCreate(&S.stream));  // default priority
  // metrics: 7 arrays, each with index offset 2, length nmax + 6
  const int nmax = n[0] > n[1] ? (n[0] > n[2] ? n[0] : n[2]) : (n[1] > n[2] ? n[1] : n[2]);
  const int L = nmax + 6;
  std::vector<double> hm(7 * (size_t)L, 0.0);
  const double *src[7] = {grid->dxb, grid->dxc, grid->dyb, grid->dyc, grid->dzb, grid->dzc, grid->grav};
  for (int q = 0; q < 7; ++q) {
    const int ax = q < 6 ? q / 2 : 2;
    const int hi = (q < 6 && (q % 2) == 1) ? n[ax] + 1 : n[ax] + 2;
    for (int i = -1; i <= hi; ++i) hm[(size_t)q * L + (i + 2)] = src[q] ? src[q][i + 1] : 0.0;
  }
  HIPCHK(hipMalloc((void **)&S.metric, sizeof(double) * hm.size()));
  HIPCHK(hipMemcpy(S.metric, hm.data(), sizeof(double) * hm.size(), hipMemcpyHostToDevice));
  S.g.nx = n[0];
  S.g.ny = n[1];
  S.g.nz = n[2];
  S.g.sy = S.sy;
  S.g.sz = S.sz;
  const double *m[7];
  for (int q = 0; q < 7; ++q) m[q] = S.metric + (size_t)q * L + 2;
  S.g.dxb = m[0]; S.g.dxc = m[1]; S.g.dyb = m[2]; S.g.dyc = m[3]; S.g.dzb = m[4]; S.g.dzc = m[5]; S.g.grav = m[6];
  for (int a = 0; a < 3; ++a) {
    S.g.per[a] = S.per[a];
    S.g.user[a] = S.user[a];
  }
  double **arrays[] = {&S.rho, &S.rho_alt, &S.energy, &S.vx, &S.vy, &S.vz, &S.vx1, &S.vy1, &S.vz1, &S.bx, &S.by,
                       &S.bz, &S.eta, &S.visc3, &S.visc_dep, &S.ohmic_dep, &S.rho_v, &S.cs_v, &S.al1, &S.al2,
                       &S.al3, &S.visc_heat, &S.p0, &S.pp, &S.bpx, &S.bpy, &S.bpz, &S.fxr, &S.fyr, &S.fzr,
                       &S.curlb, &S.cv1, &S.cv2, &S.db1, &S.F, &S.dm, &S.Fx, &S.Fy, &S.Fz, &S.rho_vr,
                       &S.cv1v, &S.cv2v, &S.dmv, &S.larsen, &S.Lc, &S.Yb[0], &S.Yb[1], &S.Yb[2]};
  for (double **pp : arrays) {
    int rc = dalloc(pp);
    if (rc) return rc;
  }
  if (params->use_viscous_damping) {
    int rc = dalloc(&S.qx);
    if (!rc) rc = dalloc(&S.qy);
    if (!rc) rc = dalloc(&S.qz);
    if (rc) return rc;
  } else {
    S.qx = S.qy = S.qz = nullptr;
  }
  // partial sums: largest reduction grid (0..n+1 per axis) x 4 values
  dim3 gr = grid_from0(n[0] + 2, n[1] + 2, n[2] + 2);
  S.partial_cap = (size_t)gr.x * gr.y * gr.z * 4;
  // + RED_ROWS x RED_MAXV level-1 results after the per-block partials
  HIPCHK(hipMalloc((void **)&S.partial, sizeof(double) * (S.partial_cap + RED_ROWS * RED_MAXV)));
  HIPCHK(hipMalloc((void **)&S.dscal, sizeof(double) * 16));
  HIPCHK(hipMemset(S.dscal, 0, sizeof(double) * 16));
  HIPCHK(hipHostMalloc((void **)&S.hscal, sizeof(double) * 16));
  HIPCHK(hipStreamSynchronize(S.stream));
  S.dt = S.dt2 = S.dtr = S.time = 0.0;
  S.n_s_stages = S.peak_substeps = S.last_substeps = 0;
  S.first_set_dt = 1;
  S.step = 0;
  S.ready = true;
  return 0;
}

int lare_dev_upload(const lare_fields *f) {
  if (!S.ready) return fail(LARE_E_STATE, "lare_dev_upload before lare_dev_init");
  if (!f || !f->rho || !f->energy || !f->vx || !f->vy || !f->vz || !f->bx || !f->by || !f->bz)
    return fail(LARE_E_ARG, "lare_dev_upload: missing field");
  if (S.p.use_viscous_damping && !f->visc3) return fail(LARE_E_ARG, "visc3 required with viscous damping");
  HIPCHK(hipStreamSynchronize(S.stream));
  int rc = 0;
  rc |= xfer(S.rho, f->rho, FK_CELL, true);
  rc |= xfer(S.energy, f->energy, FK_CELL, true);
  rc |= xfer(S.vx, f->vx, FK_NODE, true);
  rc |= xfer(S.vy, f->vy, FK_NODE, true);
  rc |= xfer(S.vz, f->vz, FK_NODE, true);
  rc |= xfer(S.bx, f->bx, FK_BX, true);
  rc |= xfer(S.by, f->by, FK_BY, true);
  rc |= xfer(S.bz, f->bz, FK_BZ, true);
  if (f->visc3) rc |= xfer(S.visc3, f->visc3, FK_NODE, true);
  if (f->eta) rc |= xfer(S.eta, f->eta, FK_CELL, true);
  if (f->visc_dep) rc |= xfer(S.visc_dep, f->visc_dep, FK_CELL, true);
  if (f->ohmic_dep) rc |= xfer(S.ohmic_dep, f->ohmic_dep, FK_CELL, true);
  if (rc) return fail(LARE_E_STATE, "upload failed: %s", g_err);
  S.loaded = true;
  return 0;
}

int lare_dev_download(lare_fields *f) {
  if (!S.loaded) return fail(LARE_E_STATE, "lare_dev_download before lare_dev_upload");
  if (!f) return fail(LARE_E_ARG, "lare_dev_download: NULL");
  HIPCHK(hipStreamSynchronize(S.stream));
  int rc = 0;
  if (f->rho) rc |= xfer(S.rho, f->rho, FK_CELL, false);
  if (f->energy) rc |= xfer(S.energy, f->energy, FK_CELL, false);
  if (f->vx) rc |= xfer(S.vx, f->vx, FK_NODE, false);
  if (f->vy) rc |= xfer(S.vy, f->vy, FK_NODE, false);
  if (f->vz) rc |= xfer(S.vz, f->vz, FK_NODE, false);
  if (f->bx) rc |= xfer(S.bx, f->bx, FK_BX, false);
  if (f->by) rc |= xfer(S.by, f->by, FK_BY, false);
  if (f->bz) rc |= xfer(S.bz, f->bz, FK_BZ, false);
  if (f->eta) rc |= xfer(S.eta, f->eta, FK_CELL, false);
  if (f->visc_dep) rc |= xfer(S.visc_dep, f->visc_dep, FK_CELL, false);
  if (f->ohmic_dep) rc |= xfer(S.ohmic_dep, f->ohmic_dep, FK_CELL, false);
  if (rc) return fail(LARE_E_STATE, "download failed: %s", g_err);
  return 0;
}

int lare_dev_get_scalars(lare_scalars *o) {
  if (!S.ready) return fail(LARE_E_STATE, "not initialised");
  double acc[2];
  int rc = read_scalars(8, 2, acc);
  if (rc) return rc;
  o->dt = S.dt;
  o->dtr = S.dtr;
  o->time = S.time;
  o->total_visc_heating = acc[0];
  o->total_ohmic_heating = acc[1];
  o->n_s_stages = S.n_s_stages;
  o->peak_substeps = S.peak_substeps;
  o->last_substeps = S.last_substeps;
  o->first_set_dt = S.first_set_dt;
  o->step = S.step;
  return 0;
}

int lare_dev_set_scalars(const lare_scalars *in) {
  if (!S.ready) return fail(LARE_E_STATE, "not initialised");
  S.dt = in->dt;
  S.dtr = in->dtr;
  S.time = in->time;
  S.n_s_stages = in->n_s_stages;
  S.peak_substeps = in->peak_substeps;
  S.last_substeps = in->last_substeps;
  S.first_set_dt = in->first_set_dt;
  S.step = in->step;
  double acc[2] = {in->total_visc_heating, in->total_ohmic_heating};
  HIPCHK(hipMemcpy(S.dscal + 8, acc, sizeof(acc), hipMemcpyHostToDevice));
  return 0;
}

#define REQUIRE_LOADED() \
  if (!S.loaded) return fail(LARE_E_STATE, "%s before lare_dev_upload", __func__)

int lare_dev_lagrangian_step(void) {
  REQUIRE_LOADED();
  return lagrangian_step();
}

int lare_dev_eulerian_remap(int32_t step) {
  REQUIRE_LOADED();
  return eulerian_remap(step, false);
}

int lare_dev_eta_calc(void) {
  REQUIRE_LOADED();
  return eta_calc();
}

int lare_dev_conduct_heat(void) {
  REQUIRE_LOADED();
  if (S.p.kappa_0 <= 0.0) return fail(LARE_E_ARG, "conduct_heat needs kappa_0 > 0");
  int rc = calc_s_stages();
  if (rc) return rc;
  return heat_conduct_sts2(0);
}

int lare_dev_boundary_conditions(void) {
  REQUIRE_LOADED();
  return boundary_conditions();
}

int lare_dev_energy_account(double en[3], double *divb) {
  REQUIRE_LOADED();
  if (!en) return fail(LARE_E_ARG, "energy_account: NULL output");
  dim3 gr = grid_from0(nx(), ny(), nz());
  LAUNCH("energy", k_energy, gr, box_block(), S.g, S.rho, S.energy, S.vx, S.vy, S.vz, S.bx, S.by, S.bz, S.partial);
  int rc = final_reduce(gr, 4, 8, 0, 0);
  if (rc) return rc;
  double v[4];
  rc = read_scalars(0, 4, v);
  if (rc) return rc;
  en[0] = v[0];
  en[1] = v[1];
  en[2] = v[2];
  if (divb) *divb = -v[3];
  return 0;
}

int lare_dev_step(int32_t step, double *dt, double *time) {
  REQUIRE_LOADED();
  S.step = step;
  int rc = lagrangian_step();
  if (rc) return rc;
  rc = eulerian_remap(step, false);
  if (rc) return rc;
  rc = eta_calc();
  if (rc) return rc;
  if (!(S.dt == S.dt) || (S.dt <= 0.0 && !S.p.restart))
    return fail(LARE_F_NONFINITE, "step %d: invalid dt %g", step, S.dt);
  if (dt) *dt = S.dt;
  if (time) *time = S.time;
  return 0;
}

int lare_dev_profile(int32_t enable) {
  if (!S.ready) return fail(LARE_E_STATE, "not initialised");
  prof_flush();
  if (enable && !S.prof) S.acc.clear();
  S.prof = enable != 0;
  return 0;
}

int lare_dev_kernel_stats(char *names, int32_t name_len, int32_t *counts, double *total_ms, int32_t cap) {
  if (!S.ready) return fail(LARE_E_STATE, "not initialised");
  prof_flush();
  int q = 0;
  for (auto &a : S.acc) {
    if (q >= cap) break;
    if (names && name_len > 0) {
      strncpy(names + (size_t)q * name_len, a.name.c_str(), (size_t)name_len - 1);
      names[(size_t)q * name_len + name_len - 1] = 0;
    }
    if (counts) counts[q] = a.count;
    if (total_ms) total_ms[q] = a.ms;
    ++q;
  }
  return q;
}

int lare_dev_synchronize(void) {
  if (!S.stream) return 0;
  HIPCHK(hipStreamSynchronize(S.stream));
  return 0;
}

int lare_dev_finalize(void) {
  free_all();
  if (S.stream) {
    (void)hipStreamDestroy(S.stream);
    S.stream = nullptr;
  }
  return 0;
}

}  // extern "C"

#include "surrogate_rt.hpp"
