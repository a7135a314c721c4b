// runtime.hpp -- device state, launch sequencing and the entry points of one
// precision mode.  lare_dev.hip includes this file twice: in namespace fp64
// with LARE_REAL = double and in namespace fp32 with LARE_REAL = float.
// `real` is the storage type of the bulk fields (density, energy, velocity,
// their Lagrangian and remap work arrays, conduction stages); the face
// magnetic field, its constrained-transport fluxes and resistive EMFs, the
// metrics and every reduction stay FP64, and kernels do their arithmetic in
// FP64 registers.  No include guard: one inclusion per namespace.
typedef LARE_REAL real;
// Arithmetic type of the bulk kernels that have been given one: double in
// the FP64 mode (the reference's arithmetic), float in the FP32 mode.  The
// face field's CT updates, the limiter comparisons and every reduction
// stay FP64 in both.
typedef LARE_ACC acc;

#include "common.hpp"
#include "bc.hpp"
#include "conduct.hpp"
#include "lagran.hpp"
#include "remap.hpp"
#include "surrogate.hpp"
#include "initial.hpp"
#include "neutral.hpp"
#include "openbc.hpp"

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
  bool any_open = false;  // some side is BC_OPEN
  lare_params p;
  double tfac, visc2_norm;
  hipStream_t stream = nullptr;
  double *metric = nullptr;  // 7 arrays of length nmax+5, index i+2
  Geo g;
  // state (bulk fields in the storage precision, face B in FP64)
  real *rho, *rho_alt, *energy, *vx, *vy, *vz, *vx1, *vy1, *vz1;
  double *bx, *by, *bz;
  real *eta, *visc3, *visc_dep, *ohmic_dep;
  // lagrangian scratch
  real *rho_v, *cs_v, *al1, *al2, *al3, *visc_heat, *p0, *pp;
  double *bpx, *bpy, *bpz;       // predictor B
  double *fxr, *fyr, *fzr;       // resistive EMFs (CT update of B)
  real *qx, *qy, *qz;
  // remap scratch: F, F2 the CT face fluxes (FP64-sized arrays, holding
  // the fluxes in the storage type; the potential field's FP64 scratch),
  // Fe the energy flux
  real *cv1, *cv2, *db1, *Fe, *dm, *Fx, *Fy, *Fz;
  double *F, *F2;
  // conduction scratch
  real *larsen, *Lc, *Yb[3];
  // options: gamma_boris (boris), cool_term_v / cool_term_b and the stored
  // curlb of the ohmic cooling (cooling_term); NULL when off
  real *gamma_boris = nullptr, *cool_v = nullptr, *cool_b = nullptr, *curlb = nullptr;
  // rke: delta_ke (cells) and one sweep's edge terms per velocity component
  real *delta_ke = nullptr;
  RkeOut rk = {};
  // FOURTHORDER: one stage's EMFs and curlb, and their weighted sum
  RkK rk4k = {}, rk4s = {};
  // partially ionised EOS: xi_n (cells), eta_perp (nodes, with Cowling), the
  // bracket fault flag of neutral_fraction; ip = ionise_pot (EOS_ION only)
  real *xi = nullptr, *eta_perp = nullptr;
  int *nf_flag = nullptr;
  // BC_OPEN: the pre-corrector energy and density of each open face's first
  // interior layer (6 faces x 2 fields x (nmax+2)^2)
  double *obc_save = nullptr;
  size_t obc_face = 0;
  bool pi_eos = false;
  double ip = 0.0;
  // reductions
  double *partial = nullptr;
  size_t partial_cap = 0;
  double *dscal = nullptr;   // device scalars: [0..3] last reduce, [8] tvh acc, [9] toh acc
  double *hscal = nullptr;   // pinned host mirror
  std::vector<void *> allocs;
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

// grid of k_resist over the padded box: tiles of RS_OUT_X x RS_OUT_Y
// written nodes, ZCHUNK planes per block
dim3 resist_grid(int *tiles_x) {
  const Box fb = full_box();
  *tiles_x = (fb.x1 - fb.x0 + RS_OUT_X) / RS_OUT_X;
  const int tiles_y = (fb.y1 - fb.y0 + RS_OUT_Y) / RS_OUT_Y;
  return dim3((unsigned)(*tiles_x * tiles_y), 1, (unsigned)((fb.z1 - fb.z0 + ZCHUNK) / ZCHUNK));
}

// grid of a z-marching launch: BX*BY columns, ZCHUNK planes per block
dim3 zchunk_grid(const Box &b) {
  return dim3(plane_blocks(b), 1, (unsigned)((b.z1 - b.z0 + 1 + ZCHUNK - 1) / ZCHUNK));
}

// the box whose zchunk grid a launch marching along axis M (1, 2) uses:
// y and z ranges swapped for M = 1 (MCOLUMN)
Box march_box(const Box &b, int M) { return M == 2 ? b : mkbox(b.x0, b.x1, b.z0, b.z1, b.y0, b.y1); }

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
  void *a;
  FieldKind kind;
};

// FP64 storage: face B (and predictor B); everything else in `real`
bool wide_kind(FieldKind k) { return k == FK_BX || k == FK_BY || k == FK_BZ || sizeof(real) == sizeof(double); }

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
  if (axis == 0) {
    const int lines = (S.n[1] + 5) * (S.n[2] + 5);
    LAUNCH(name, k_bc_x, dim3((unsigned)((lines * 8 + 255) / 256)), dim3(256), S.g, s);
    return 0;
  }
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
      BcSet s = {};
      s.nf = nf;
      for (int q = 0; q < nf; ++q) {
        s.f[q] = f[q].a;
        s.op[q] = op_for(f[q].kind, ax, periodic);
        s.wide[q] = wide_kind(f[q].kind) ? 1 : 0;
      }
      s.skip_lo = S.bc[2 * ax] == LARE_BC_OPEN;
      s.skip_hi = S.bc[2 * ax + 1] == LARE_BC_OPEN;
      if (s.skip_lo && s.skip_hi) continue;
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
CondArgs cond_args() {
  CondArgs ca;
  ca.kappa_0 = S.p.kappa_0;
  ca.tfac = S.tfac;
  ca.dt = S.dt;
  ca.xi = S.pi_eos ? S.xi : nullptr;
  ca.gm1 = S.p.gamma - 1.0;
  ca.ip = S.ip;
  return ca;
}

int calc_s_stages() {
  dim3 gr = grid_from0(nx(), ny(), nz());
  if (S.pi_eos)
    LAUNCH("larsen", k_larsen<true>, gr, box_block(), S.g, S.energy, S.rho, S.larsen, S.partial, S.p.flux_limiter,
           S.p.kappa_0, cond_args(), S.p.heat_flux_limiter);
  else
    LAUNCH("larsen", k_larsen<false>, gr, box_block(), S.g, S.energy, S.rho, S.larsen, S.partial, S.p.flux_limiter,
           S.p.kappa_0, cond_args(), S.p.heat_flux_limiter);
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

// grid of the z-marching stage kernels over the interior box: waves of
// RKL_OUT cells over rows of x0-1..x1, ZCHUNK planes per block
dim3 rkl2_grid(const Box &in) {
  const long npts = (long)(in.x1 - in.x0 + 2) * (in.y1 - in.y0 + 1);
  const long waves = (npts + RKL_OUT - 1) / RKL_OUT;
  return dim3((unsigned)((waves + BY - 1) / BY), 1, (unsigned)((in.z1 - in.z0 + 1 + ZCHUNK - 1) / ZCHUNK));
}
// stage 1 from E0 (ghosts as stored): Lc and Y1 = E0 + mu_t1 dt Lc
int rkl2_stage1(const real *E0, real *Yout, CondArgs ca, double mu_t1) {
  const Box in = interior();
  // the stage-1 read is direct: no ghost mapping, BC_OPEN or not
  if (ca.xi)
    LAUNCH("rkl2_stage1", (k_rkl2_z<TMODE_DIRECT, 1, true, false>), rkl2_grid(in), box_block(), S.g, in, E0, E0, E0,
           S.rho, S.bx, S.by, S.bz, S.larsen, S.Lc, Yout, ca, 0.0, 0.0, 0.0, mu_t1, 0.0);
  else
    LAUNCH("rkl2_stage1", (k_rkl2_z<TMODE_DIRECT, 1, false, false>), rkl2_grid(in), box_block(), S.g, in, E0, E0, E0,
           S.rho, S.bx, S.by, S.bz, S.larsen, S.Lc, Yout, ca, 0.0, 0.0, 0.0, mu_t1, 0.0);
  return 0;
}
// stage j >= 2 from Y2 (and Y1, E0, Lc) under a ghost-read mode
int rkl2_stage(int mode, const real *Y2, const real *Y1, const real *E0, real *Yout, CondArgs ca, double mu, double nu,
               double one_m_mu_nu, double mu_t, double gam_t) {
  const Box in = interior();
#define RKL2_STAGE(M, P, O)                                                                                       \
  LAUNCH("rkl2_stage", (k_rkl2_z<M, 0, P, O>), rkl2_grid(in), box_block(), S.g, in, Y2, Y1, E0, S.rho, S.bx, S.by, \
         S.bz, S.larsen, S.Lc, Yout, ca, mu, nu, one_m_mu_nu, mu_t, gam_t)
  if (S.any_open) {
    if (mode == TMODE_STALE) {
      if (ca.xi) RKL2_STAGE(TMODE_STALE, true, true);
      else RKL2_STAGE(TMODE_STALE, false, true);
    } else {
      if (ca.xi) RKL2_STAGE(TMODE_WRAP, true, true);
      else RKL2_STAGE(TMODE_WRAP, false, true);
    }
  } else if (mode == TMODE_STALE) {
    if (ca.xi) RKL2_STAGE(TMODE_STALE, true, false);
    else RKL2_STAGE(TMODE_STALE, false, false);
  } else {
    if (ca.xi) RKL2_STAGE(TMODE_WRAP, true, false);
    else RKL2_STAGE(TMODE_WRAP, false, false);
  }
#undef RKL2_STAGE
  return 0;
}

int neutral_fraction(const real *Y, int mode);
EosArgs eos_args();

// RKL2 super-step; delta_mode = 1 inside lagrangian_step
int heat_conduct_sts2(int delta_mode) {
  const int s = S.n_s_stages;
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
  CondArgs ca = cond_args();
  real *E0 = S.energy;
  real *outp = S.Yb[0];
  int rc = rkl2_stage1(E0, S.Yb[0], ca, mu_t[1]);
  if (rc) return rc;
  real *y1 = E0, *y2 = outp;
  int next = 1;
  for (int j = 2; j <= s; ++j) {
    real *yo = S.Yb[next];
    next = (next + 1) % 3;
    rc = rkl2_stage(j == 2 ? TMODE_STALE : TMODE_WRAP, y2, y1, E0, yo, ca, mu[j], nu[j], (1.0 - mu[j]) - nu[j],
                    mu_t[j], gt[j]);
    if (rc) return rc;
    // the neutral fraction of the stage value with its boundary fill
    // (conduct.f90:400-405), read by the next stage's temperatures
    if (S.pi_eos && j < s) {
      rc = neutral_fraction(yo, TMODE_WRAP);
      if (rc) return rc;
    }
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

// CT fluxes and updates of one sweep in one pass, from the face arrays in
// to the face arrays out (never the same: the stencil reads neighbours that
// other blocks write).  sc = 0: both hold face fluxes.  sc = +1 (first sweep
// of a remap): in holds B, converted on the fly (the scaling pass of
// remap.f90:44-67 folded in).  sc = -1 (last predictor sweep): out gets B
// back (remap.f90:102-125 folded in).
template <int AX>
int ct_sweep(const double *const in[3], double *const out[3], const real *vel, double dt, int sc) {
  const int N[3] = {nx(), ny(), nz()};
  const Box ub = mkbox(0, N[0] + 1, 0, N[1] + 1, 0, N[2] + 1);
  // (transverse field t, velocity-average axis, geometry-average axis, box_t, box_a) x 2
  int t1, va1, ga1, t2, va2, ga2;
  Box bt1 = ub, ba1 = ub, bt2 = ub, ba2 = ub;
  if (AX == 0) {
    t1 = 1; va1 = 2; ga1 = 1;  // vx_by_flux
    t2 = 2; va2 = 1; ga2 = 2;  // vx_bz_flux
    bt2 = mkbox(1, N[0], 1, N[1], 0, N[2]);
    ba2 = mkbox(0, N[0], 0, N[1] + 1, 0, N[2] + 1);
  } else if (AX == 1) {
    t1 = 0; va1 = 2; ga1 = 0;  // vy_bx_flux
    t2 = 2; va2 = 0; ga2 = 2;  // vy_bz_flux
    ba2 = mkbox(1, N[0], 0, N[1], 1, N[2]);
  } else {
    t1 = 0; va1 = 1; ga1 = 0;  // vz_bx_flux
    t2 = 1; va2 = 0; ga2 = 1;  // vz_by_flux
  }
  const dim3 blk(CTX_THREADS, 1, 1);
  if constexpr (AX == 0) {
    // x sweep: va1 = z and ga2 = z (see the flux table above); tiles of
    // CTX_W points in x by CTX_ROWS - 1 rows in y, marching in z
    const int ntx = (N[0] + 5 + CTX_W - 1) / CTX_W, nty = (N[1] + 5 + CTX_ROWS - 2) / (CTX_ROWS - 1);
    const dim3 grid((unsigned)(ntx * nty), 1, (unsigned)((N[2] + 5 + CTS_CHUNK - 1) / CTS_CHUNK));
    if (sc > 0)
      LAUNCH("ct_sweep", k_ct_sweep_x<1>, grid, blk, S.g, in[t1], in[t2], in[AX], out[t1], out[t2], out[AX], vel,
             S.db1, strides(ga1), strides(va2), dt, ntx, bt1, ba1, bt2, ba2);
    else if (sc < 0)
      LAUNCH("ct_sweep", k_ct_sweep_x<-1>, grid, blk, S.g, in[t1], in[t2], in[AX], out[t1], out[t2], out[AX], vel,
             S.db1, strides(ga1), strides(va2), dt, ntx, bt1, ba1, bt2, ba2);
    else
      LAUNCH("ct_sweep", k_ct_sweep_x<0>, grid, blk, S.g, in[t1], in[t2], in[AX], out[t1], out[t2], out[AX], vel,
             S.db1, strides(ga1), strides(va2), dt, ntx, bt1, ba1, bt2, ba2);
  } else {
    // y and z sweeps march along the sweep axis; tiles of CTM_XW - 1 points
    // in x by CTM_ROWS - 1 rows along the other transverse axis
    const int P = AX == 2 ? 1 : 2;
    const int ntx = (N[0] + 5 + CTM_XW - 2) / (CTM_XW - 1), nty = (N[P] + 5 + CTM_ROWS - 2) / (CTM_ROWS - 1);
    const dim3 grid((unsigned)(ntx * nty), 1, (unsigned)((N[AX] + 5 + CTS_CHUNK - 1) / CTS_CHUNK));
    const dim3 blkm(CTM_THREADS, 1, 1);
    if (sc > 0)
      LAUNCH("ct_sweep", (k_ct_sweep_m<AX, 1>), grid, blkm, S.g, in[t1], in[t2], in[AX], out[t1], out[t2], out[AX],
             vel, S.db1, strides(va1), strides(ga1), strides(va2), strides(ga2), dt, ntx, bt1, ba1, bt2, ba2);
    else if (sc < 0)
      LAUNCH("ct_sweep", (k_ct_sweep_m<AX, -1>), grid, blkm, S.g, in[t1], in[t2], in[AX], out[t1], out[t2], out[AX],
             vel, S.db1, strides(va1), strides(ga1), strides(va2), strides(ga2), dt, ntx, bt1, ba1, bt2, ba2);
    else
      LAUNCH("ct_sweep", (k_ct_sweep_m<AX, 0>), grid, blkm, S.g, in[t1], in[t2], in[AX], out[t1], out[t2], out[AX],
             vel, S.db1, strides(va1), strides(ga1), strides(va2), strides(ga2), dt, ntx, bt1, ba1, bt2, ba2);
  }
  return 0;
}

// the face arrays a and the spare EMF arrays swap roles (the CT output went
// to the spares)
void swap_spare(double *&ax, double *&ay, double *&az) {
  double *t;
  t = ax; ax = S.fxr; S.fxr = t;
  t = ay; ay = S.fyr; S.fyr = t;
  t = az; az = S.fzr; S.fzr = t;
}

// one predictor-mode sweep (B only, no flag reset, no halo refresh) of the
// copy bp of B; the first sweep reads B itself, the last converts back
template <int AX>
int sweep_predictor(double dt, bool first, bool last) {
  // the last pass also forms the predictor energy and pressure
  const PredP pr = {S.energy, S.rho, S.visc_heat, S.p0, S.pp, S.p.gamma, S.dt2, eos_args()};
  if (last)
    LAUNCH("geom", (k_geom<AX, 1>), zchunk_grid(cell_box()), box_block(), S.g, cell_box(), S.vx, S.vy, S.vz, S.cv1,
           S.cv2, S.db1, dt, 1.0, 1.0, 1.0, pr);
  else
    LAUNCH("geom", (k_geom<AX, 0>), zchunk_grid(cell_box()), box_block(), S.g, cell_box(), S.vx, S.vy, S.vz, S.cv1,
           S.cv2, S.db1, dt, 1.0, 1.0, 1.0, pr);
  const real *vel = AX == 0 ? S.vx : (AX == 1 ? S.vy : S.vz);
  if (first) {
    // B itself stays untouched: the first sweep writes the copy
    const double *const Bin[3] = {S.bx, S.by, S.bz};
    double *const Bp[3] = {S.bpx, S.bpy, S.bpz};
    return ct_sweep<AX>(Bin, Bp, vel, dt, 1);
  }
  const double *const Bp[3] = {S.bpx, S.bpy, S.bpz};
  double *const spare[3] = {S.fxr, S.fyr, S.fzr};
  int rc = ct_sweep<AX>(Bp, spare, vel, dt, last ? -1 : 0);
  if (rc) return rc;
  swap_spare(S.bpx, S.bpy, S.bpz);
  return 0;
}

template <int AX>
int sweep_full(double dt, double pass[3], bool first, bool last) {
  const int N[3] = {nx(), ny(), nz()};
  const real *vel = AX == 0 ? S.vx1 : (AX == 1 ? S.vy1 : S.vz1);
  LAUNCH("geom", (k_geom<AX, 2>), zchunk_grid(cell_box()), box_block(), S.g, cell_box(), S.vx1, S.vy1, S.vz1, S.cv1,
         S.cv2, S.db1, dt, pass[0], pass[1], pass[2], PredP{});
  {
    const double *const B[3] = {S.bx, S.by, S.bz};
    double *const spare[3] = {S.fxr, S.fyr, S.fzr};
    int rc = ct_sweep<AX>(B, spare, vel, dt, first ? 1 : (last ? -1 : 0));
    if (rc) return rc;
    swap_spare(S.bx, S.by, S.bz);
  }
  int rc = 0;
  int hi[3] = {N[0] + 1, N[1] + 1, N[2] + 1};
  hi[AX] = N[AX];
  Box mb = mkbox(0, hi[0], 0, hi[1], 0, hi[2]);
  // mass flux, energy flux and the density and energy updates in one pass
  // over the mass flux's box (updates on cells 1..n); the new energy lands
  // in the energy flux array, which then swaps roles with the energy array
  // (the new density goes to rho_alt as before)
  if constexpr (AX == 0) {
    const long npts = (long)(mb.x1 - mb.x0 + 1) * (mb.y1 - mb.y0 + 1);
    const long waves = (npts + REX_OUT - 1) / REX_OUT;
    const dim3 gx((unsigned)((waves + BY - 1) / BY), 1, (unsigned)(mb.z1 - mb.z0 + 1));
    LAUNCH("rho_energy", k_mass_rho_energy_x, gx, box_block(), S.g, mb, S.energy, S.rho, vel, S.db1, S.dm, S.cv1,
           S.cv2, S.rho_alt, S.Fe, dt, S.per[AX] ? 1 : 0);
  } else {
    LAUNCH("rho_energy", k_mass_rho_energy_m<AX>, zchunk_grid(march_box(mb, AX)), box_block(), S.g, mb, S.energy,
           S.rho, vel, S.db1, S.dm, S.cv1, S.cv2, S.rho_alt, S.Fe, dt, S.per[AX] ? 1 : 0);
  }
  {
    BcSet s = {};
    s.nf = 1;
    s.f[0] = S.dm;
    s.op[0] = S.per[AX] ? BC_PER_CELL : BC_DM_CLAMP;
    s.wide[0] = wide_kind(FK_CELL) ? 1 : 0;
    rc = launch_bc_axis(s, AX, "bc_dm");
    if (rc) return rc;
  }
  if (S.any_open) {
    // the new energy and density cover cells 1..n; the ghosts of a BC_OPEN
    // side keep their values (no fill there)
    LAUNCH_BOX("copy_open", k_copy_open, full_box(), S.energy, S.Fe, 0);
    LAUNCH_BOX("copy_open", k_copy_open, full_box(), S.rho, S.rho_alt, 0);
  }
  {
    real *te = S.energy;
    S.energy = S.Fe;
    S.Fe = te;
  }
  {
    // vertex quantities, momentum fluxes and update in one pass over the
    // padded box (rho1 = rho before this sweep's update, which went to
    // rho_alt); the new velocities land in the flux arrays, which then swap
    // roles with the velocity arrays
    const Box xb = full_box();
    if constexpr (AX == 0) {
      const long npts = (long)(xb.x1 - xb.x0 + 1) * (xb.y1 - xb.y0 + 1);
      const long out = S.delta_ke ? MOMX_OUT_RKE : MOMX_OUT;
      const long waves = (npts + out - 1) / out;
      const dim3 gx((unsigned)((waves + BY - 1) / BY), 1, (unsigned)((xb.z1 - xb.z0 + ZCHUNK) / ZCHUNK));
      if (S.delta_ke)
        LAUNCH("mom_remap", k_mom_remap_x<true>, gx, box_block(), S.g, xb, S.vx, S.vy, S.vz, vel, S.db1, S.rho,
               S.cv1, S.cv2, S.dm, S.Fx, S.Fy, S.Fz, dt, S.rk);
      else
        LAUNCH("mom_remap", k_mom_remap_x<false>, gx, box_block(), S.g, xb, S.vx, S.vy, S.vz, vel, S.db1, S.rho,
               S.cv1, S.cv2, S.dm, S.Fx, S.Fy, S.Fz, dt, S.rk);
    } else {
      if (S.delta_ke)
        LAUNCH("mom_remap", (k_mom_remap_m<AX, true>), zchunk_grid(march_box(xb, AX)), box_block(), S.g, xb, S.vx,
               S.vy, S.vz, vel, S.db1, S.rho, S.cv1, S.cv2, S.dm, S.Fx, S.Fy, S.Fz, dt, S.rk);
      else
        LAUNCH("mom_remap", (k_mom_remap_m<AX, false>), zchunk_grid(march_box(xb, AX)), box_block(), S.g, xb, S.vx,
               S.vy, S.vz, vel, S.db1, S.rho, S.cv1, S.cv2, S.dm, S.Fx, S.Fy, S.Fz, dt, S.rk);
    }
    // rke: this sweep's edge terms into delta_ke
    if (S.delta_ke) LAUNCH_BOX("rke_gather", k_rke_gather<AX>, interior(), S.rk, S.delta_ke);
    real *tv;
    tv = S.vx; S.vx = S.Fx; S.Fx = tv;
    tv = S.vy; S.vy = S.Fy; S.Fy = tv;
    tv = S.vz; S.vz = S.Fz; S.Fz = tv;
  }
  real *t = S.rho;
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
  // The B -> face flux scaling (remap.f90:44-67) happens inside the first
  // sweep's CT kernels and the scaling back (remap.f90:102-125) inside the
  // last sweep's update.  In the full remap the last sweep's halo refresh
  // then runs on B instead of on face fluxes; the B halos it writes are
  // written again from the same interior by the bfield_bcs that follows the
  // scaling back in the reference, and nothing reads them in between.
  if (predictor) {
    for (int q = 0; q < 3; ++q) {
      const int ax = kStrang[c][q];
      const bool first = q == 0, last = q == 2;
      int rc = ax == 0 ? sweep_predictor<0>(dt, first, last)
                       : (ax == 1 ? sweep_predictor<1>(dt, first, last) : sweep_predictor<2>(dt, first, last));
      if (rc) return rc;
    }
    return bfield_bcs(S.bpx, S.bpy, S.bpz);
  }
  // rke: delta_ke = 0 (remap.f90:38); only the interior is ever read
  if (S.delta_ke) LAUNCH_BOX("fill", k_fill, interior(), S.delta_ke, 0.0);
  double pass[3] = {1.0, 1.0, 1.0};
  for (int q = 0; q < 3; ++q) {
    const int ax = kStrang[c][q];
    const bool first = q == 0, last = q == 2;
    int rc = ax == 0 ? sweep_full<0>(dt, pass, first, last)
                     : (ax == 1 ? sweep_full<1>(dt, pass, first, last) : sweep_full<2>(dt, pass, first, last));
    if (rc) return rc;
  }
  return bfield_bcs(S.bx, S.by, S.bz);
}

// ----------------------------------------------------------------------------
// lagrangian step
// ----------------------------------------------------------------------------
int eta_calc() {
  if (S.p.resistive_mhd) {
    const Box eb = mkbox(-1, nx() + 1, -1, ny() + 1, -1, nz() + 1);
    LAUNCH("eta", k_eta, zchunk_grid(eb), box_block(), S.g, eb, S.bx, S.by, S.bz, S.eta,
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
  a.va_max2 = S.p.va_max * S.p.va_max;
  a.alpha_av = S.p.alpha_av;
  a.resistive = S.p.resistive_mhd;
  a.boris = S.p.boris;
  a.write_gb = !(S.first_set_dt && S.p.restart);
  a.limiter = S.p.shock_limiter;
  a.compression = S.p.shock_compression;
  a.cowling = S.p.cowling_resistivity;
  a.eta_perp = S.eta_perp;
  return a;
}

// ----------------------------------------------------------------------------
// partially ionised plasma (neutral.f90)
// ----------------------------------------------------------------------------
EosArgs eos_args() { return EosArgs{S.pi_eos ? S.xi : nullptr, S.ip}; }

NeutralArgs neutral_args() {
  NeutralArgs a;
  a.gm1 = S.p.gamma - 1.0;
  a.ip = S.ip;
  a.f_bar = S.p.f_bar;
  a.t_bar = S.p.t_bar;
  a.r_bar = S.p.r_bar;
  a.eta_bar = S.p.eta_bar;
  a.t_rad = S.p.t_rad;
  return a;
}

// neutral_fraction over the cell box from Y (energy, or an RKL2 stage value
// read through the boundary map when mode is TMODE_WRAP)
int neutral_fraction(const real *Y, int mode) {
  if (mode == TMODE_WRAP)
    LAUNCH_BOX("neutral", k_neutral<TMODE_WRAP>, cell_box(), Y, S.energy, S.rho, S.xi, S.nf_flag, neutral_args());
  else
    LAUNCH_BOX("neutral", k_neutral<TMODE_DIRECT>, cell_box(), Y, S.energy, S.rho, S.xi, S.nf_flag, neutral_args());
  return 0;
}

int perpendicular_resistivity() {
  LAUNCH_BOX("eta_perp", k_eta_perp, nodes0n(), S.rho, S.energy, S.xi, S.bx, S.by, S.bz, S.eta_perp,
             neutral_args());
  return 0;
}

// ----------------------------------------------------------------------------
// open boundaries (openboundary.f90)
// ----------------------------------------------------------------------------
OpenFace open_face(int a, int hi) {
  OpenFace f;
  const int n = S.n[a];
  f.a = a;
  f.hi = hi;
  f.u = a == 0 ? 1 : 0;
  f.w = a == 2 ? 1 : 2;
  f.ty = a == 1 ? 0 : 1;
  f.tz = a == 2 ? 0 : 2;
  f.lb = hi ? n : 0;
  f.lc = hi ? n : 1;
  f.lfc = hi ? n + 2 : -1;
  f.lfn = hi ? n + 2 : -2;
  f.lset_n = hi ? n + 1 : -1;
  f.lset_c = hi ? n + 1 : 0;
  f.sg = hi ? 1.0 : -1.0;
  return f;
}

unsigned face_points(const OpenFace &f) { return (unsigned)((S.n[f.u] + 2) * (S.n[f.w] + 2)); }

// the pre-corrector energy and density of each open face's first interior layer
int open_save() {
  for (int a = 0; a < 3; ++a)
    for (int hi = 0; hi <= 1; ++hi) {
      if (S.bc[2 * a + hi] != LARE_BC_OPEN) continue;
      const OpenFace f = open_face(a, hi);
      LAUNCH("open_save", k_open_save, dim3((face_points(f) + 255) / 256), dim3(256), S.g, f, S.energy, S.rho,
             S.obc_save + (size_t)(2 * a + hi) * S.obc_face);
    }
  return 0;
}

// open_bcs: the faces one launch each, in the reference's order
int open_bcs() {
  for (int a = 0; a < 3; ++a)
    for (int hi = 0; hi <= 1; ++hi) {
      if (S.bc[2 * a + hi] != LARE_BC_OPEN) continue;
      const OpenFace f = open_face(a, hi);
      LAUNCH("open_bcs", k_open_bcs, dim3((face_points(f) + 255) / 256), dim3(256), S.g, f, S.vx, S.vy, S.vz, S.bx,
             S.by, S.bz, S.rho, S.energy, S.obc_save + (size_t)(2 * a + hi) * S.obc_face, S.p.gamma);
    }
  return 0;
}

// the reference STOPs on an inverted temperature bracket (neutral.f90:172-175):
// report it once the work queued so far has run, and clear the flag
int check_neutral_fault(const char *where) {
  if (!S.nf_flag) return 0;
  int h = 0;
  HIPCHK(hipMemcpyAsync(&h, S.nf_flag, sizeof(int), hipMemcpyDeviceToHost, S.stream));
  HIPCHK(hipStreamSynchronize(S.stream));
  if (!h) return 0;
  HIPCHK(hipMemsetAsync(S.nf_flag, 0, sizeof(int), S.stream));
  return fail(LARE_F_TBOUNDS, "%s: neutral_fraction temperature bracket inverted (energy <= 0 in some cell)", where);
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

// One FOURTHORDER resistive substep (lagran.F90:146-149 with :1329-1415):
// eta_calc; B1 = B (in the spare fxr..fzr arrays); rkstep and bstep four
// times; the final bstep with the weighted EMFs; the heating; energy_bcs.
int resist_rk4_substep() {
  int rc = eta_calc();
  if (rc) return rc;
  const size_t nb = sizeof(double) * S.ntot;
  HIPCHK(hipMemcpyAsync(S.fxr, S.bx, nb, hipMemcpyDeviceToDevice, S.stream));
  HIPCHK(hipMemcpyAsync(S.fyr, S.by, nb, hipMemcpyDeviceToDevice, S.stream));
  HIPCHK(hipMemcpyAsync(S.fzr, S.bz, nb, hipMemcpyDeviceToDevice, S.stream));
  const Box n0 = nodes0n();
  const double dt = S.dt, half_dt = dt * 0.5, dt6 = dt * (1.0 / 6.0);
  const double h[3] = {half_dt, half_dt, dt};
  const int w[4] = {0, 2, 2, 1};
  for (int q = 0; q < 4; ++q) {
    if (q > 0) {
      LAUNCH_BOX("rk4_bstep", k_rk4_bstep, n0, S.fxr, S.fyr, S.fzr, S.rk4k, S.bx, S.by, S.bz, h[q - 1]);
      rc = bfield_bcs(S.bx, S.by, S.bz);
      if (rc) return rc;
    }
    if (S.eta_perp)
      LAUNCH_BOX("rk4_emf", k_rk4_emf<true>, n0, S.bx, S.by, S.bz, S.eta, S.rk4k, S.rk4s, w[q], S.eta_perp);
    else
      LAUNCH_BOX("rk4_emf", k_rk4_emf<false>, n0, S.bx, S.by, S.bz, S.eta, S.rk4k, S.rk4s, w[q], S.eta_perp);
  }
  LAUNCH_BOX("rk4_bstep", k_rk4_bstep, n0, S.fxr, S.fyr, S.fzr, S.rk4s, S.bx, S.by, S.bz, dt6);
  rc = bfield_bcs(S.bx, S.by, S.bz);
  if (rc) return rc;
  LAUNCH_BOX("rk4_heat", k_rk4_heat, interior(), S.rk4s.k[3], S.rho, S.energy, dt6);
  rc = energy_bcs();
  if (rc) return rc;
  const dim3 gr = box_grid(n0);
  LAUNCH("ohmic_total", k_ohmic_total, gr, box_block(), S.g, n0, S.rk4s.k[3], S.partial, dt6);
  return final_reduce(gr, 1, 0, 9, 1);
}

int lagrangian_step() {
  const int N0 = nx(), N1 = ny(), N2 = nz();
  LagArgs a = lag_args();  // dt = previous step's dt (Q1)
  {
    const Box vb = mkbox(-1, N0 + 1, -1, N1 + 1, -1, N2 + 1);
    const EosArgs eq = eos_args();
#define VERTEX_PREP(GB, PI)                                                                                        \
  LAUNCH("vertex_prep", (k_vertex_prep<GB, PI>), zchunk_grid(vb), box_block(), S.g, vb, S.rho, S.energy, S.bx, S.by, \
         S.bz, S.gamma_boris, S.rho_v, S.cs_v, S.p.gamma, eq)
    if (S.gamma_boris) {
      if (S.pi_eos) VERTEX_PREP(true, true);
      else VERTEX_PREP(true, false);
    } else {
      if (S.pi_eos) VERTEX_PREP(false, true);
      else VERTEX_PREP(false, false);
    }
#undef VERTEX_PREP
  }
  if (S.p.use_viscous_damping) {
    LAUNCH_BOX("damp_q", k_damp_q, nodes0n(), S.vx, S.vy, S.vz, S.visc3, S.qx, S.qy, S.qz);
    LAUNCH_BOX("damp_apply", k_damp_apply, nodes0n(), S.vx, S.vy, S.vz, S.qx, S.qy, S.qz, S.dt);
  }
  {
    const Box ab = mkbox(-1, N0 + 1, -1, N1 + 2, 0, N2 + 2);
    LAUNCH("alpha", k_alpha, zchunk_grid(ab), box_block(), S.g, ab, S.vx, S.vy, S.vz, S.rho_v, S.cs_v, S.al1, S.al2,
           S.al3, a);
  }
  const Box vdb = mkbox(0, N0 + 1, 0, N1 + 1, 0, N2 + 1);
  dim3 gr = zchunk_grid(vdb);
  if (S.pi_eos)
    LAUNCH("visc_dt", k_visc_dt<true>, gr, box_block(), S.g, vdb, S.vx, S.vy, S.vz, S.rho, S.energy, S.bx, S.by, S.bz,
           S.eta, S.al1, S.al2, S.al3, S.visc_heat, S.p0, S.gamma_boris, S.partial, a, eos_args());
  else
    LAUNCH("visc_dt", k_visc_dt<false>, gr, box_block(), S.g, vdb, S.vx, S.vy, S.vz, S.rho, S.energy, S.bx, S.by, S.bz,
           S.eta, S.al1, S.al2, S.al3, S.visc_heat, S.p0, S.gamma_boris, S.partial, a, eos_args());
  int rc = final_reduce(gr, 3, 7, 0, 0);
  if (rc) return rc;
  double mins[3];
  rc = read_scalars(0, 3, mins);
  if (rc) return rc;
  if (mins[2] < 0.0) {
    // the state entering the step holds a non-finite value or rho <= 0
    static const char *names[8] = {"rho", "energy", "vx", "vy", "vz", "bx", "by", "bz"};
    const double code = -mins[2];
    const int fid = (int)(code / 4294967296.0) - 1;
    const long cell = (long)(code - (double)(fid + 1) * 4294967296.0);
    const int ci = (int)(cell % N0) + 1, cj = (int)((cell / N0) % N1) + 1, ck = (int)(cell / ((long)N0 * N1)) + 1;
    return fail(LARE_F_NONFINITE, "step %d: %s is not finite%s at (%d, %d, %d) entering lagrangian_step", S.step,
                fid >= 0 && fid < 8 ? names[fid] : "?", fid == 0 ? " or not positive" : "", ci, cj, ck);
  }
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
    // eta_calc, rkstep, bstep and the ohmic heating of each substep run in
    // k_resist, which writes the new B to the fxr..fzr arrays; they then
    // swap roles with bx..bz.  The B and energy halos of a substep are
    // refreshed together.
    const Box fb = full_box();
    int tiles_x;
    const dim3 g2 = resist_grid(&tiles_x);
    if ((size_t)g2.x * g2.y * g2.z > S.partial_cap)
      return fail(LARE_E_ARG, "resistive grid of %u blocks exceeds the partial buffer", g2.x * g2.z);
    for (int sc = 1; sc <= substeps; ++sc) {
      // eta_calc runs inside the substep kernels; neutral_fraction and
      // perpendicular_resistivity read nothing it writes (lagran.F90:146-149)
      if (S.pi_eos) {
        rc = neutral_fraction(S.energy, TMODE_DIRECT);
        if (rc) return rc;
      }
      if (S.eta_perp) {
        rc = perpendicular_resistivity();
        if (rc) return rc;
      }
      if (S.p.fourth_order) {
        rc = resist_rk4_substep();
        if (rc) return rc;
        continue;
      }
      if (S.eta_perp)
        LAUNCH("resist", k_resist<true>, g2, dim3(64, RS_ROWS, 1), S.g, fb, tiles_x, S.bx, S.by, S.bz, S.fxr, S.fyr,
               S.fzr, S.eta, S.energy, S.ohmic_dep, S.rho, S.curlb, S.partial, S.dt, S.p.eta_background, S.p.eta0,
               S.p.j_max, S.eta_perp);
      else
        LAUNCH("resist", k_resist<false>, g2, dim3(64, RS_ROWS, 1), S.g, fb, tiles_x, S.bx, S.by, S.bz, S.fxr, S.fyr,
               S.fzr, S.eta, S.energy, S.ohmic_dep, S.rho, S.curlb, S.partial, S.dt, S.p.eta_background, S.p.eta0,
               S.p.j_max, S.eta_perp);
      rc = final_reduce(g2, 1, 0, 9, 1);
      if (rc) return rc;
      if (S.curlb)
        LAUNCH_BOX("cool_b", k_cool_b, interior(), S.curlb, S.rho, S.cool_b, S.energy, S.p.alpha_av, S.dt);
      double *tb;
      tb = S.bx; S.bx = S.fxr; S.fxr = tb;
      tb = S.by; S.by = S.fyr; S.fyr = tb;
      tb = S.bz; S.bz = S.fzr; S.fzr = tb;
      BcField f[4] = {{S.bx, FK_BX}, {S.by, FK_BY}, {S.bz, FK_BZ}, {S.energy, FK_CELL}};
      rc = apply_bcs(f, 4, "bc_resistive");
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
  } else if (!S.p.resistive_mhd || S.p.fourth_order) {
    // energy = MAX(energy0 + 0, 0) (lagran.F90:180-181).  After a
    // first-order resistive step it changes nothing: the last substep's ohmic
    // heating already clamped the whole cell box (:1305) and energy_bcs only
    // copies clamped values, so the launch is left out then.
    LAUNCH_BOX("energy_max", k_energy_max, cell_box(), S.energy);
  }

  // predictor-corrector (lagran.F90:208-478)
  if (S.p.cauchy) {
    // the Cauchy solution for the cell-centred B1 (B1 cv1 in the spare
    // predictor-B arrays), cv1 and the predictor pressure; no predictor remap
    real *b1[3] = {(real *)S.bpx, (real *)S.bpy, (real *)S.bpz};
    LAUNCH_BOX("cauchy_b", k_cauchy_b, cell_box(), S.vx, S.vy, S.vz, S.bx, S.by, S.bz, S.cv1, b1[0], b1[1], b1[2],
               S.energy, S.rho, S.visc_heat, S.p0, S.pp, S.p.gamma, S.dt2, eos_args());
    LAUNCH("force", k_force<true>, zchunk_grid(nodes0n()), box_block(), S.g, nodes0n(), S.vx, S.vy, S.vz, S.pp,
           S.bx, S.by, S.bz, S.rho_v, S.al1, S.al2, S.al3, S.gamma_boris, S.vx1, S.vy1, S.vz1, S.dt2, S.cv1, b1[0],
           b1[1], b1[2]);
  } else {
    const double dt_full = S.dt;
    S.dt = 0.5 * dt_full;
    rc = eulerian_remap(S.step, true);
    S.dt = dt_full;
    if (rc) return rc;
    // the predictor energy and pressure were formed in the predictor
    // remap's last geometry pass
    LAUNCH("force", k_force<false>, zchunk_grid(nodes0n()), box_block(), S.g, nodes0n(), S.vx, S.vy, S.vz, S.pp,
           S.bpx, S.bpy, S.bpz, S.rho_v, S.al1, S.al2, S.al3, S.gamma_boris, S.vx1, S.vy1, S.vz1, S.dt2, S.cv1,
           nullptr, nullptr, nullptr);
  }
  {
    BcField f[3] = {{S.vx1, FK_NODE}, {S.vy1, FK_NODE}, {S.vz1, FK_NODE}};
    rc = apply_bcs(f, 3, "bc_v1");
    if (rc) return rc;
  }
  if (S.any_open) {
    // v1 = v over the whole array before the predictor remap
    // (lagran.F90:1051-1053); the forces then set nodes 0..n and
    // remap_v_bcs the filled sides, so a BC_OPEN side keeps v's ghosts
    LAUNCH_BOX("copy_open", k_copy_open, full_box(), S.vx, S.vx1, 1);
    LAUNCH_BOX("copy_open", k_copy_open, full_box(), S.vy, S.vy1, 1);
    LAUNCH_BOX("copy_open", k_copy_open, full_box(), S.vz, S.vz1, 1);
  }
  if (S.any_open) {
    rc = open_save();
    if (rc) return rc;
  }
  dim3 g3 = zchunk_grid(interior());
  if (S.cool_v)
    LAUNCH("shock_corr", k_shock_corr<true>, g3, box_block(), S.g, interior(), S.vx, S.vy, S.vz, S.vx1, S.vy1, S.vz1,
           S.al1, S.al2, S.al3, S.pp, S.energy, S.rho, S.visc_dep, S.Fx, S.Fy, S.Fz, S.cool_v, S.p.alpha_av,
           S.partial, S.dt);
  else
    LAUNCH("shock_corr", k_shock_corr<false>, g3, box_block(), S.g, interior(), S.vx, S.vy, S.vz, S.vx1, S.vy1, S.vz1,
           S.al1, S.al2, S.al3, S.pp, S.energy, S.rho, S.visc_dep, S.Fx, S.Fy, S.Fz, S.cool_v, S.p.alpha_av,
           S.partial, S.dt);
  rc = final_reduce(g3, 1, 0, 8, 1);
  if (rc) return rc;
  {
    const long nlow = (long)(nx() + 1) * (ny() + 1) + (long)nz() * (nx() + ny() + 1);
    LAUNCH("vupdate_low", k_vupdate_low, dim3((unsigned)((nlow + 255) / 256)), dim3(256), S.g, S.vx, S.vy, S.vz,
           S.vx1, S.vy1, S.vz1, S.Fx, S.Fy, S.Fz);
  }
  // the new velocities went to the momentum-flux arrays, which swap roles
  // with the velocity arrays; the boundary fill below sets their ghosts
  if (S.any_open) {
    LAUNCH_BOX("copy_open", k_copy_open, full_box(), S.vx, S.Fx, 1);
    LAUNCH_BOX("copy_open", k_copy_open, full_box(), S.vy, S.Fy, 1);
    LAUNCH_BOX("copy_open", k_copy_open, full_box(), S.vz, S.Fz, 1);
  }
  {
    real *tv;
    tv = S.vx; S.vx = S.Fx; S.Fx = tv;
    tv = S.vy; S.vy = S.Fy; S.Fy = tv;
    tv = S.vz; S.vz = S.Fz; S.Fz = tv;
  }
  if (S.any_open) {
    // velocity_bcs, open_bcs, then (with the end of lagrangian_step)
    // energy_bcs and density_bcs (lagran.F90:408-412, :196-198)
    BcField fv[3] = {{S.vx, FK_NODE}, {S.vy, FK_NODE}, {S.vz, FK_NODE}};
    rc = apply_bcs(fv, 3, "bc_lagrangian");
    if (rc) return rc;
    rc = open_bcs();
    if (rc) return rc;
    BcField fc[2] = {{S.energy, FK_CELL}, {S.rho, FK_CELL}};
    return apply_bcs(fc, 2, "bc_lagrangian");
  }
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
  for (void *p : S.allocs) (void)hipFree(p);
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

template <class T>
int dalloc(T **p) {
  hipError_t e = hipMalloc((void **)p, sizeof(T) * S.ntot);
  if (e != hipSuccess) return fail(LARE_E_NOMEM, "hipMalloc of %zu elements failed: %s", S.ntot, hipGetErrorString(e));
  S.allocs.push_back((void *)*p);
  e = hipMemsetAsync(*p, 0, sizeof(T) * S.ntot, S.stream);
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

// host staging buffer of one element type
template <class T>
std::vector<T> &stage_buf() {
  static std::vector<T> v;
  return v;
}

// copy one field between a host Fortran array (FP64, reference bounds) and
// its padded device box, converting to the device storage type
template <class T>
int xfer(T *dev, double *host, FieldKind k, bool to_dev) {
  int lo[3];
  lower_of(k, lo);
  const long ex = nx() + 2 - lo[0] + 1, ey = ny() + 2 - lo[1] + 1;
  std::vector<T> &st = stage_buf<T>();
  if (st.size() != S.ntot) st.assign(S.ntot, (T)0);
  HIPCHK(hipMemcpy(st.data(), dev, sizeof(T) * S.ntot, hipMemcpyDeviceToHost));
  for (int kk = lo[2]; kk <= nz() + 2; ++kk)
    for (int jj = lo[1]; jj <= ny() + 2; ++jj)
      for (int ii = lo[0]; ii <= nx() + 2; ++ii) {
        const long h = ((long)(kk - lo[2]) * ey + (jj - lo[1])) * ex + (ii - lo[0]);
        if (to_dev) st[S.g.idx(ii, jj, kk)] = (T)host[h];
        else host[h] = (double)st[S.g.idx(ii, jj, kk)];
      }
  if (to_dev) HIPCHK(hipMemcpy(dev, st.data(), sizeof(T) * S.ntot, hipMemcpyHostToDevice));
  return 0;
}

// ============================================================================
// entry points of this precision (wrapped by the C-ABI in lare_dev.hip)
// ============================================================================


int api_init(const lare_grid *grid, const lare_params *params) {
  if (!grid || !params) return fail(LARE_E_ARG, "lare_dev_init: NULL argument");
  free_all();
  const int n[3] = {grid->nx, grid->ny, grid->nz};
  const int bc[6] = {grid->xbc_min, grid->xbc_max, grid->ybc_min, grid->ybc_max, grid->zbc_min, grid->zbc_max};
  for (int a = 0; a < 3; ++a) {
    if (n[a] < 4) return fail(LARE_E_ARG, "grid size %d on axis %d is below 4", n[a], a);
    const int lo = bc[2 * a], hi = bc[2 * a + 1];
    const bool ok = (lo == LARE_BC_PERIODIC && hi == LARE_BC_PERIODIC) ||
                    ((lo == LARE_BC_USER || lo == LARE_BC_OPEN) && (hi == LARE_BC_USER || hi == LARE_BC_OPEN));
    if (!ok)
      return fail(LARE_E_ARG,
                  "axis %d: boundary codes (%d, %d) unsupported (both periodic, or each side BC_USER or BC_OPEN)", a,
                  lo, hi);
  }
  if (!grid->dxb || !grid->dxc || !grid->dyb || !grid->dyc || !grid->dzb || !grid->dzc)
    return fail(LARE_E_ARG, "lare_dev_init: metric arrays missing");
  memcpy(S.n, n, sizeof(n));
  memcpy(S.bc, bc, sizeof(bc));
  S.any_open = false;
  for (int a = 0; a < 3; ++a) {
    S.per[a] = bc[2 * a] == LARE_BC_PERIODIC;
    S.user[a] = !S.per[a];
    S.any_open = S.any_open || bc[2 * a] == LARE_BC_OPEN || bc[2 * a + 1] == LARE_BC_OPEN;
  }
  const int eos = params->eos_number == 0 ? LARE_EOS_IDEAL : params->eos_number;
  if (eos != LARE_EOS_IDEAL && eos != LARE_EOS_ION && eos != LARE_EOS_PI)
    return fail(LARE_E_ARG, "eos_number %d unknown", params->eos_number);
  const bool pi = eos != LARE_EOS_IDEAL;
  if (params->cowling_resistivity && !pi)
    return fail(LARE_E_ARG, "cowling_resistivity needs a partially ionised eos_number (control.f90:158-161)");
  if (pi && sizeof(real) != sizeof(double))
    return fail(LARE_E_ARG, "the partially ionised equation of state runs in the FP64 mode only");
  for (int q = 0; q < 6; ++q) {
    if (bc[q] == LARE_BC_OPEN && sizeof(real) != sizeof(double))
      return fail(LARE_E_ARG, "BC_OPEN boundaries run in the FP64 mode only");
    // with CAUCHY nothing sets v1 on a BC_OPEN side's ghost nodes before the
    // remap reads them (lagran.F90:1050-1059 copy v to v1 only without it)
    if (bc[q] == LARE_BC_OPEN && params->cauchy)
      return fail(LARE_E_ARG, "BC_OPEN boundaries with the CAUCHY predictor: v1 on the open ghosts is undefined");
  }
  if (pi && !(params->f_bar > 0.0 && params->t_rad > 0.0 && params->r_bar > 0.0 && params->t_bar >= 0.0))
    return fail(LARE_E_ARG, "partially ionised: f_bar, t_rad and r_bar must be > 0, t_bar >= 0");
  if (pi && params->cowling_resistivity && !(params->eta_bar >= 0.0))
    return fail(LARE_E_ARG, "cowling_resistivity: eta_bar must be >= 0");
  S.p = *params;
  S.p.eos_number = eos;
  S.pi_eos = pi;
  S.ip = eos == LARE_EOS_ION ? params->ionise_pot : 0.0;
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
    S.g.open_lo[a] = bc[2 * a] == LARE_BC_OPEN;
    S.g.open_hi[a] = bc[2 * a + 1] == LARE_BC_OPEN;
  }
  real **bulk[] = {&S.rho, &S.rho_alt, &S.energy, &S.vx, &S.vy, &S.vz, &S.vx1, &S.vy1, &S.vz1, &S.eta,
                    &S.visc3, &S.visc_dep, &S.ohmic_dep, &S.rho_v, &S.cs_v, &S.al1, &S.al2, &S.al3,
                    &S.visc_heat, &S.p0, &S.pp, &S.cv1, &S.cv2, &S.db1, &S.Fe, &S.dm, &S.Fx, &S.Fy,
                    &S.Fz, &S.larsen, &S.Lc, &S.Yb[0], &S.Yb[1], &S.Yb[2]};
  for (real **pp : bulk) {
    int rc = dalloc(pp);
    if (rc) return rc;
  }
  double **wide[] = {&S.bx, &S.by, &S.bz, &S.bpx, &S.bpy, &S.bpz, &S.fxr, &S.fyr, &S.fzr, &S.F, &S.F2};
  for (double **pp : wide) {
    int rc = dalloc(pp);
    if (rc) return rc;
  }
  S.gamma_boris = S.cool_v = S.cool_b = S.curlb = nullptr;
  S.delta_ke = nullptr;
  S.rk = RkeOut{};
  S.rk4k = RkK{};
  S.rk4s = RkK{};
  if (params->fourth_order && params->resistive_mhd) {
    int rc = 0;
    for (int q = 0; q < 4 && !rc; ++q) rc = dalloc(&S.rk4k.k[q]);
    for (int q = 0; q < 4 && !rc; ++q) rc = dalloc(&S.rk4s.k[q]);
    if (rc) return rc;
  }
  if (params->rke) {
    int rc = dalloc(&S.delta_ke);
    for (int q = 0; q < 3 && !rc; ++q) rc = dalloc(&S.rk.dk[q]);
    if (rc) return rc;
  }
  S.xi = S.eta_perp = nullptr;
  S.nf_flag = nullptr;
  S.obc_save = nullptr;
  S.obc_face = 0;
  if (S.any_open) {
    const size_t e = (size_t)(nmax + 2) * (size_t)(nmax + 2);
    S.obc_face = 2 * e;
    HIPCHK(hipMalloc((void **)&S.obc_save, sizeof(double) * 6 * S.obc_face));
    S.allocs.push_back(S.obc_save);
  }
  if (pi) {
    // xi_n = 0 until the first neutral_fraction (neutral.f90:38)
    int rc = dalloc(&S.xi);
    if (!rc && params->cowling_resistivity) rc = dalloc(&S.eta_perp);
    if (rc) return rc;
    HIPCHK(hipMalloc((void **)&S.nf_flag, sizeof(int)));
    S.allocs.push_back(S.nf_flag);
    HIPCHK(hipMemsetAsync(S.nf_flag, 0, sizeof(int), S.stream));
  }
  if (params->boris) {
    int rc = dalloc(&S.gamma_boris);
    if (rc) return rc;
    // 1 until the first set_dt (gamma_boris(:,:,:) = 1, lagran.F90:1092)
    LAUNCH_BOX("fill", k_fill, full_box(), S.gamma_boris, 1.0);
  }
  if (params->cooling_term) {
    int rc = dalloc(&S.cool_v);
    if (!rc) rc = dalloc(&S.cool_b);
    if (!rc && params->resistive_mhd) rc = dalloc(&S.curlb);
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
  int tiles_x;
  const dim3 grs = resist_grid(&tiles_x);
  if ((size_t)grs.x * grs.y * grs.z > S.partial_cap) S.partial_cap = (size_t)grs.x * grs.y * grs.z;
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

int api_upload(const lare_fields *f) {
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

int api_download(lare_fields *f) {
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

int api_get_scalars(lare_scalars *o) {
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

int api_set_scalars(const lare_scalars *in) {
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

int api_lagrangian_step(void) {
  REQUIRE_LOADED();
  int rc = lagrangian_step();
  if (rc) return rc;
  return check_neutral_fault("lagrangian_step");
}

int api_eulerian_remap(int32_t step) {
  REQUIRE_LOADED();
  return eulerian_remap(step, false);
}

int api_eta_calc(void) {
  REQUIRE_LOADED();
  return eta_calc();
}

int api_conduct_heat(void) {
  REQUIRE_LOADED();
  if (S.p.kappa_0 <= 0.0) return fail(LARE_E_ARG, "conduct_heat needs kappa_0 > 0");
  int rc = calc_s_stages();
  if (rc) return rc;
  rc = heat_conduct_sts2(0);
  if (rc) return rc;
  return check_neutral_fault("conduct_heat");
}

int api_boundary_conditions(void) {
  REQUIRE_LOADED();
  return boundary_conditions();
}

int api_neutral_fraction(void) {
  REQUIRE_LOADED();
  if (!S.pi_eos) return fail(LARE_E_STATE, "neutral_fraction needs eos_number EOS_ION or EOS_PI");
  int rc = neutral_fraction(S.energy, TMODE_DIRECT);
  if (rc) return rc;
  return check_neutral_fault("neutral_fraction");
}

int api_perpendicular_resistivity(void) {
  REQUIRE_LOADED();
  if (!S.eta_perp) return fail(LARE_E_STATE, "perpendicular_resistivity needs cowling_resistivity");
  return perpendicular_resistivity();
}

// xi_n and eta_perp (0..n, zero elsewhere) in the cell layout
int api_get_neutral(double *xi_h, double *ep_h) {
  REQUIRE_LOADED();
  if ((xi_h && !S.xi) || (ep_h && !S.eta_perp))
    return fail(LARE_E_STATE, "get_neutral: the partially ionised arrays are off");
  HIPCHK(hipStreamSynchronize(S.stream));
  int rc = 0;
  if (xi_h) rc = xfer(S.xi, xi_h, FK_CELL, false);
  if (!rc && ep_h) {
    rc = xfer(S.eta_perp, ep_h, FK_CELL, false);
    if (rc) return rc;
    // the array exists at nodes 0..n only (neutral.f90:73-131)
    const long ex = nx() + 4, ey = ny() + 4, ez = nz() + 4;
    for (long k = 0; k < ez; ++k)
      for (long j = 0; j < ey; ++j)
        for (long i = 0; i < ex; ++i)
          if (i < 1 || i > nx() + 1 || j < 1 || j > ny() + 1 || k < 1 || k > nz() + 1) ep_h[(k * ey + j) * ex + i] = 0.0;
  }
  return rc;
}

int api_energy_account(double en[3], double *divb) {
  REQUIRE_LOADED();
  if (!en) return fail(LARE_E_ARG, "energy_account: NULL output");
  dim3 gr = grid_from0(nx(), ny(), nz());
  LAUNCH("energy", (k_energy<PadLayout, real>), gr, box_block(), S.g, PadLayout{S.g}, S.rho, S.energy, S.vx, S.vy,
         S.vz, S.bx, S.by, S.bz, S.partial);
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

// diagnostics.F90:613-629 energy_correction, then energy_bcs
int energy_correction() {
  LAUNCH_BOX("energy_correction", k_energy_correction, interior(), S.delta_ke, S.rho, S.energy);
  return energy_bcs();
}

int api_energy_correction() {
  REQUIRE_LOADED();
  if (!S.delta_ke) return fail(LARE_E_STATE, "energy_correction needs rke on");
  return energy_correction();
}

// lare3d.f90:82-87 loop body (lagrangian_step, eulerian_remap,
// energy_correction when rke, eta_calc)
int api_step(int32_t step, double *dt, double *time) {
  REQUIRE_LOADED();
  S.step = step;
  int rc = 0;
  if (S.pi_eos) {
    rc = neutral_fraction(S.energy, TMODE_DIRECT);
    if (rc) return rc;
  }
  rc = lagrangian_step();
  if (rc) return rc;
  rc = eulerian_remap(step, false);
  if (rc) return rc;
  if (S.delta_ke) {
    rc = energy_correction();
    if (rc) return rc;
  }
  rc = eta_calc();
  if (rc) return rc;
  rc = check_neutral_fault("step");
  if (rc) return rc;
  if (!isfinite(S.dt) || (S.dt <= 0.0 && !S.p.restart))
    return fail(LARE_F_NONFINITE, "step %d: invalid dt %g", step, S.dt);
  if (dt) *dt = S.dt;
  if (time) *time = S.time;
  return 0;
}

int api_profile(int32_t enable) {
  if (!S.ready) return fail(LARE_E_STATE, "not initialised");
  prof_flush();
  if (enable && !S.prof) S.acc.clear();
  S.prof = enable != 0;
  return 0;
}

int api_kernel_stats(char *names, int32_t name_len, int32_t *counts, double *total_ms, int32_t cap) {
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

int api_synchronize(void) {
  if (!S.stream) return 0;
  HIPCHK(hipStreamSynchronize(S.stream));
  return 0;
}

int api_finalize(void) {
  free_all();
  if (S.stream) {
    (void)hipStreamDestroy(S.stream);
    S.stream = nullptr;
  }
  return 0;
}


// ============================================================================
// learned conduction surrogate
// ============================================================================
// Runtime of the learned conduction surrogate (kernels in surrogate.hpp).
//
// Data flow:
//   collect: features of the current state -> random interior cells -> rows
//            of 27 x 12 raw channels; one plain RKL2 super-step on the same
//            state gives the target dT of each row; the state is restored.
//   fit_norm: channel mean / std and the target scale from the rows; the
//            stored rows are standardised in place.
//   train:   Adam on mini-batches drawn from the rows; forward and backward
//            GEMMs in BF16 MFMA with FP32 accumulation.
//   conduct: features -> MFMA inference over the interior -> E1; residual
//            r = |E1 - E0 - dt/2 (L(E0) + L(E1))| / |E0| (L2 norms over the
//            interior, L = heat-flux divergence / rho of conduct.f90:152-291);
//            r <= tol accepts E1, otherwise the super-step is redone by RKL2.



struct Sur {
  float *w = nullptr, *gr = nullptr, *m = nullptr, *v = nullptr;  // SUR_NPARAM each
  unsigned short *w1t = nullptr, *w2t = nullptr;                   // packed BF16 for inference
  bool have_params = false, have_norm = false;
  SurNorm nm;
  int adam_t = 0;
  float *feat = nullptr;      // ntot * SUR_CH
  real *E1 = nullptr, *trap = nullptr, *E0save = nullptr;
  // dataset
  float *X = nullptr, *y = nullptr;
  int *cells = nullptr;
  int rows = 0, cap = 0;
  bool standardised = false;
  // mini-batch work arrays (capacity bcap columns)
  int bcap = 0;
  float *Xb = nullptr, *yb = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *yo = nullptr,
        *dy = nullptr, *dZ2 = nullptr, *dH1 = nullptr, *dZ1 = nullptr, *lossacc = nullptr, *losshist = nullptr;
  int losscap = 0;
  // use inside lagrangian_step
  int enabled = 0;
  double tol = 1.0e-4;
  int accepted = 0, rejected = 0;
  double last_res = 0.0;
};
Sur SU;

template <class T>
void sfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void sur_free_batch() {
  sfree(SU.Xb); sfree(SU.yb); sfree(SU.Z1); sfree(SU.H1); sfree(SU.Z2); sfree(SU.H2); sfree(SU.yo);
  sfree(SU.dy); sfree(SU.dZ2); sfree(SU.dH1); sfree(SU.dZ1); sfree(SU.lossacc);
  SU.bcap = 0;
}

void sur_free() {
  sfree(SU.w); sfree(SU.gr); sfree(SU.m); sfree(SU.v); sfree(SU.w1t); sfree(SU.w2t);
  sfree(SU.feat); sfree(SU.E1); sfree(SU.trap); sfree(SU.E0save);
  sfree(SU.X); sfree(SU.y); sfree(SU.cells); sfree(SU.losshist);
  sur_free_batch();
  const int en = SU.enabled;
  const double tol = SU.tol;
  SU = Sur();
  SU.enabled = en;  // the mode survives re-initialisation; parameters do not
  SU.tol = tol;
}

template <class T>
int salloc(T **p, size_t count) {
  hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
  if (e != hipSuccess) return fail(LARE_E_NOMEM, "hipMalloc of %zu bytes failed: %s", sizeof(T) * count,
                                   hipGetErrorString(e));
  return 0;
}

int sur_alloc_params() {
  if (SU.w) return 0;
  int rc = salloc(&SU.w, SUR_NPARAM);
  if (!rc) rc = salloc(&SU.gr, SUR_NPARAM);
  if (!rc) rc = salloc(&SU.m, SUR_NPARAM);
  if (!rc) rc = salloc(&SU.v, SUR_NPARAM);
  if (!rc) rc = salloc(&SU.w1t, SUR_H * SUR_KP);
  if (!rc) rc = salloc(&SU.w2t, SUR_H * SUR_HP);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(SU.m, 0, sizeof(float) * SUR_NPARAM, S.stream));
  HIPCHK(hipMemsetAsync(SU.v, 0, sizeof(float) * SUR_NPARAM, S.stream));
  return 0;
}

int sur_alloc_fields() {
  if (SU.feat) return 0;
  int rc = salloc(&SU.feat, S.ntot * SUR_CH);
  if (!rc) rc = salloc(&SU.E1, S.ntot);
  if (!rc) rc = salloc(&SU.trap, S.ntot);
  if (!rc) rc = salloc(&SU.E0save, S.ntot);
  if (rc) return rc;
  // cells never written by the feature kernel are never read by it either;
  // zero them once so no stale bits reach a gather
  HIPCHK(hipMemsetAsync(SU.feat, 0, sizeof(float) * S.ntot * SUR_CH, S.stream));
  return 0;
}

int sur_ensure_batch(int cols) {
  if (cols <= SU.bcap) return 0;
  sur_free_batch();
  const size_t c = (size_t)cols;
  int rc = salloc(&SU.Xb, c * SUR_K);
  if (!rc) rc = salloc(&SU.yb, c);
  if (!rc) rc = salloc(&SU.Z1, c * SUR_H);
  if (!rc) rc = salloc(&SU.H1, c * SUR_H);
  if (!rc) rc = salloc(&SU.Z2, c * SUR_H);
  if (!rc) rc = salloc(&SU.H2, c * SUR_H);
  if (!rc) rc = salloc(&SU.yo, c);
  if (!rc) rc = salloc(&SU.dy, c);
  if (!rc) rc = salloc(&SU.dZ2, c * SUR_H);
  if (!rc) rc = salloc(&SU.dH1, c * SUR_H);
  if (!rc) rc = salloc(&SU.dZ1, c * SUR_H);
  if (!rc) rc = salloc(&SU.lossacc, (c + 255) / 256);
  if (rc) return rc;
  SU.bcap = cols;
  return 0;
}

unsigned blocks_for(size_t n, unsigned bs = 256) { return (unsigned)((n + bs - 1) / bs); }

int gemm(int M, int N, int K, const float *A, int sam, int sak, const float *B, int sbk, int sbn, float *C, int scm,
         int scn) {
  LAUNCH("sur_gemm", k_gemm_bf16, dim3((unsigned)(N / 16), (unsigned)(M / 16)), dim3(64), M, N, K, A, sam, sak, B, sbk,
         sbn, C, scm, scn);
  return 0;
}

// Forward pass on SU.Xb (nb columns, a multiple of 16; the first nreal are
// real rows).  With tgt: loss partials and dy as well.
int sur_forward(int nb, int nreal, const float *tgt) {
  int rc = gemm(SUR_H, nb, SUR_K, SU.w + SUR_OFF_W1, 1, SUR_H, SU.Xb, 1, SUR_K, SU.Z1, nb, 1);
  if (rc) return rc;
  LAUNCH("sur_bias_gelu", k_bias_gelu, dim3(blocks_for((size_t)SUR_H * nb)), dim3(256), SU.Z1, SU.H1,
         SU.w + SUR_OFF_B1, SUR_H, nb);
  rc = gemm(SUR_H, nb, SUR_H, SU.w + SUR_OFF_W2, 1, SUR_H, SU.H1, nb, 1, SU.Z2, nb, 1);
  if (rc) return rc;
  LAUNCH("sur_bias_gelu", k_bias_gelu, dim3(blocks_for((size_t)SUR_H * nb)), dim3(256), SU.Z2, SU.H2,
         SU.w + SUR_OFF_B2, SUR_H, nb);
  LAUNCH("sur_out", k_out_layer, dim3(blocks_for(nb)), dim3(256), SU.H2, SU.w, tgt, nb, nreal, SU.yo, SU.dy,
         SU.lossacc);
  return 0;
}

// Backward pass after sur_forward with a target: gradient into SU.gr.
int sur_backward(int nb) {
  LAUNCH("sur_rowdot", k_rowdot, dim3(SUR_H), dim3(256), SU.H2, SU.dy, nb, SU.gr + SUR_OFF_W3);
  LAUNCH("sur_rowdot", k_rowdot, dim3(1), dim3(256), SU.dy, (const float *)nullptr, nb, SU.gr + SUR_OFF_B3);
  LAUNCH("sur_back_out", k_back_out, dim3(blocks_for((size_t)SUR_H * nb)), dim3(256), SU.Z2, SU.w, SU.dy, nb,
         SU.dZ2);
  LAUNCH("sur_rowdot", k_rowdot, dim3(SUR_H), dim3(256), SU.dZ2, (const float *)nullptr, nb, SU.gr + SUR_OFF_B2);
  // dW2[k][f] = sum_b H1[k][b] dZ2[f][b]
  int rc = gemm(SUR_H, SUR_H, nb, SU.H1, nb, 1, SU.dZ2, 1, nb, SU.gr + SUR_OFF_W2, SUR_H, 1);
  if (rc) return rc;
  // dH1[k][b] = sum_f W2[k][f] dZ2[f][b]
  rc = gemm(SUR_H, nb, SUR_H, SU.w + SUR_OFF_W2, SUR_H, 1, SU.dZ2, nb, 1, SU.dH1, nb, 1);
  if (rc) return rc;
  LAUNCH("sur_gelu_back", k_gelu_back, dim3(blocks_for((size_t)SUR_H * nb)), dim3(256), SU.Z1, SU.dH1,
         SUR_H * nb, SU.dZ1);
  LAUNCH("sur_rowdot", k_rowdot, dim3(SUR_H), dim3(256), SU.dZ1, (const float *)nullptr, nb, SU.gr + SUR_OFF_B1);
  // dW1[k][f] = sum_b X[b][k] dZ1[f][b]
  return gemm(SUR_K, SUR_H, nb, SU.Xb, 1, SUR_K, SU.dZ1, 1, nb, SU.gr + SUR_OFF_W1, SUR_H, 1);
}

int sur_enabled() { return SU.enabled; }

int sur_pack() {
  LAUNCH("sur_pack", k_sur_pack, dim3(blocks_for((size_t)SUR_H * SUR_KP)), dim3(256), SU.w, SU.w1t, SU.w2t);
  return 0;
}

int sur_features(int normalise) {
  int rc = sur_alloc_fields();
  if (rc) return rc;
  const double logdt = log(S.dt > 0.0 ? S.dt : 1.0e-300);
  LAUNCH_BOX("sur_features", k_sur_features, mkbox(0, nx() + 1, 0, ny() + 1, 0, nz() + 1), S.energy, S.rho, S.bx,
             S.by, S.bz, SU.feat, SU.nm, S.tfac, logdt, normalise);
  return 0;
}

// Trapezoidal residual |E1 - E0 - dt/2 (L(E0) + L(E1))| / |E0| (L2 over the
// interior) of a candidate super-step E1 from E0 = S.energy, using the
// Larsen factor of the last calc_s_stages.  The stage kernels with all
// history weights zero evaluate the trapezoid into SU.trap.
int sur_residual(const real *E1, double *res) {
  CondArgs ca = {S.p.kappa_0, S.tfac, S.dt};
  int rc0 = rkl2_stage1(S.energy, S.Yb[0], ca, 0.0);
  if (rc0) return rc0;
  rc0 = rkl2_stage(TMODE_WRAP, E1, E1, S.energy, SU.trap, ca, 0.0, 0.0, 0.0, 0.5, 0.5);
  if (rc0) return rc0;
  dim3 gr = box_grid(interior());
  LAUNCH("sur_residual", k_sur_residual, gr, box_block(), S.g, S.energy, E1, SU.trap, S.partial);
  int rc = final_reduce(gr, 2, 0, 0, 0);
  if (rc) return rc;
  double v[2];
  rc = read_scalars(0, 2, v);
  if (rc) return rc;
  *res = v[1] > 0.0 ? sqrt(v[0] / v[1]) : 0.0;
  return 0;
}

// One surrogate super-step of length S.dt on S.energy (stages and Larsen
// factor from calc_s_stages), checked and, if rejected, redone by RKL2.
int sur_conduct_step(int delta_mode, double *res_out, int *acc_out) {
  if (!SU.have_params || !SU.have_norm) return fail(LARE_E_STATE, "surrogate: no trained parameters loaded");
  int rc = calc_s_stages();
  if (rc) return rc;
  rc = sur_features(1);
  if (rc) return rc;
  const dim3 gi((unsigned)((nx() + SIT_WX - 1) / SIT_WX), (unsigned)((ny() + SIT_NW - 1) / SIT_NW),
                (unsigned)((nz() + SIT_ZC - 1) / SIT_ZC));
  LAUNCH("sur_infer", k_sur_infer_t, gi, dim3(64 * SIT_NW), S.g, SU.feat, SU.w1t, SU.w2t, SU.w, S.energy, SU.E1,
         S.tfac, SU.nm.y_scale);
  double r;
  rc = sur_residual(SU.E1, &r);
  if (rc) return rc;
  SU.last_res = r;
  const int ok = (r == r) && r <= SU.tol;
  if (res_out) *res_out = r;
  if (acc_out) *acc_out = ok;
  if (ok) {
    ++SU.accepted;
    LAUNCH_BOX("rkl2_final", k_rkl2_final, cell_box(), SU.E1, S.energy, delta_mode);
    return 0;
  }
  ++SU.rejected;
  return heat_conduct_sts2(delta_mode);
}

// rows of standardised inputs (SUR_K floats each) from the host into SU.Xb,
// zero-padded to a multiple of 16 columns
int sur_stage_rows(const float *X, const float *t, int rows, int *nb_out) {
  const int nb = (rows + 15) / 16 * 16;
  int rc = sur_ensure_batch(nb);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(SU.Xb, 0, sizeof(float) * (size_t)nb * SUR_K, S.stream));
  HIPCHK(hipMemcpyAsync(SU.Xb, X, sizeof(float) * (size_t)rows * SUR_K, hipMemcpyHostToDevice, S.stream));
  if (t) {
    HIPCHK(hipMemsetAsync(SU.yb, 0, sizeof(float) * nb, S.stream));
    HIPCHK(hipMemcpyAsync(SU.yb, t, sizeof(float) * rows, hipMemcpyHostToDevice, S.stream));
  }
  *nb_out = nb;
  return 0;
}


int api_sur_set_params(const float *params, int32_t n) {
  if (!S.ready) return fail(LARE_E_STATE, "not initialised");
  if (!params || n != SUR_NPARAM) return fail(LARE_E_ARG, "surrogate parameters: need %d floats, got %d", SUR_NPARAM, n);
  int rc = sur_alloc_params();
  if (rc) return rc;
  HIPCHK(hipMemcpy(SU.w, params, sizeof(float) * SUR_NPARAM, hipMemcpyHostToDevice));
  HIPCHK(hipMemsetAsync(SU.m, 0, sizeof(float) * SUR_NPARAM, S.stream));
  HIPCHK(hipMemsetAsync(SU.v, 0, sizeof(float) * SUR_NPARAM, S.stream));
  SU.adam_t = 0;
  SU.have_params = true;
  return sur_pack();
}

int api_sur_get_params(float *params, int32_t n) {
  if (!SU.have_params) return fail(LARE_E_STATE, "surrogate: no parameters");
  if (!params || n != SUR_NPARAM) return fail(LARE_E_ARG, "surrogate parameters: need %d floats, got %d", SUR_NPARAM, n);
  HIPCHK(hipStreamSynchronize(S.stream));
  HIPCHK(hipMemcpy(params, SU.w, sizeof(float) * SUR_NPARAM, hipMemcpyDeviceToHost));
  return 0;
}

int api_sur_set_norm(const float *mu, const float *sd, float y_scale) {
  if (!mu || !sd || !(y_scale > 0.0f)) return fail(LARE_E_ARG, "surrogate norm: bad arguments");
  for (int q = 0; q < SUR_CH; ++q) {
    SU.nm.mu[q] = q < SUR_NCH ? mu[q] : 0.0f;
    SU.nm.sd[q] = q < SUR_NCH ? sd[q] : 1.0f;
    if (!(SU.nm.sd[q] > 0.0f)) return fail(LARE_E_ARG, "surrogate norm: std of channel %d is not positive", q);
  }
  SU.nm.y_scale = y_scale;
  SU.have_norm = true;
  return 0;
}

int api_sur_get_norm(float *mu, float *sd, float *y_scale) {
  if (!SU.have_norm) return fail(LARE_E_STATE, "surrogate: no normalisation");
  for (int q = 0; q < SUR_NCH; ++q) {
    if (mu) mu[q] = SU.nm.mu[q];
    if (sd) sd[q] = SU.nm.sd[q];
  }
  if (y_scale) *y_scale = SU.nm.y_scale;
  return 0;
}

int api_sur_collect(int32_t count, uint64_t seed, int32_t *total) {
  REQUIRE_LOADED();
  if (count <= 0) return fail(LARE_E_ARG, "surrogate collect: count %d", count);
  if (SU.standardised) return fail(LARE_E_STATE, "surrogate collect: rows already standardised");
  if (S.p.kappa_0 <= 0.0 || !(S.dt > 0.0)) return fail(LARE_E_ARG, "surrogate collect needs kappa_0 > 0 and dt > 0");
  if (SU.rows + count > SU.cap) {  // grow the row store
    const int ncap = (SU.rows + count) * 2;
    float *X = nullptr, *y = nullptr;
    int *cells = nullptr;
    int rc = salloc(&X, (size_t)ncap * SUR_K);
    if (!rc) rc = salloc(&y, (size_t)ncap);
    if (!rc) rc = salloc(&cells, (size_t)ncap);
    if (rc) return rc;
    if (SU.rows) {
      HIPCHK(hipMemcpyAsync(X, SU.X, sizeof(float) * SU.rows * SUR_K, hipMemcpyDeviceToDevice, S.stream));
      HIPCHK(hipMemcpyAsync(y, SU.y, sizeof(float) * SU.rows, hipMemcpyDeviceToDevice, S.stream));
    }
    HIPCHK(hipStreamSynchronize(S.stream));
    sfree(SU.X); sfree(SU.y); sfree(SU.cells);
    SU.X = X; SU.y = y; SU.cells = cells;
    SU.cap = ncap;
  }
  int rc = calc_s_stages();
  if (rc) return rc;
  rc = sur_features(0);
  if (rc) return rc;
  int *cells = SU.cells + SU.rows;
  LAUNCH("sur_sample", k_sur_sample_cells, dim3(blocks_for(count)), dim3(256), S.g, SU.feat, (unsigned long long)seed,
         count, cells);
  LAUNCH("sur_gather", k_sur_gather, dim3(blocks_for((size_t)count * SUR_K)), dim3(256), S.g, SU.feat, cells, count,
         SU.X + (size_t)SU.rows * SUR_K);
  HIPCHK(hipMemcpyAsync(SU.E0save, S.energy, sizeof(real) * S.ntot, hipMemcpyDeviceToDevice, S.stream));
  rc = heat_conduct_sts2(0);
  if (rc) return rc;
  LAUNCH("sur_target", k_sur_target, dim3(blocks_for(count)), dim3(256), SU.E0save, S.energy, cells, count, S.tfac,
         SU.y + SU.rows);
  HIPCHK(hipMemcpyAsync(S.energy, SU.E0save, sizeof(real) * S.ntot, hipMemcpyDeviceToDevice, S.stream));
  HIPCHK(hipStreamSynchronize(S.stream));
  SU.rows += count;
  if (total) *total = SU.rows;
  return 0;
}

int api_sur_rows(float *X, float *y, int32_t rows) {
  if (rows < 0 || rows > SU.rows) return fail(LARE_E_ARG, "surrogate rows: %d requested, %d held", rows, SU.rows);
  HIPCHK(hipStreamSynchronize(S.stream));
  if (X && rows) HIPCHK(hipMemcpy(X, SU.X, sizeof(float) * (size_t)rows * SUR_K, hipMemcpyDeviceToHost));
  if (y && rows) HIPCHK(hipMemcpy(y, SU.y, sizeof(float) * (size_t)rows, hipMemcpyDeviceToHost));
  return 0;
}

int api_sur_fit_norm(void) {
  if (SU.rows < 2) return fail(LARE_E_STATE, "surrogate fit_norm: %d rows collected", SU.rows);
  if (SU.standardised) return fail(LARE_E_STATE, "surrogate fit_norm: rows already standardised");
  std::vector<float> X((size_t)SU.rows * SUR_K), y(SU.rows);
  HIPCHK(hipStreamSynchronize(S.stream));
  HIPCHK(hipMemcpy(X.data(), SU.X, sizeof(float) * X.size(), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(y.data(), SU.y, sizeof(float) * y.size(), hipMemcpyDeviceToHost));
  // channel statistics of the centre cell (neighbour 13) of every row
  for (int ch = 0; ch < SUR_CH; ++ch) {
    double s = 0.0, s2 = 0.0;
    for (int r = 0; r < SU.rows; ++r) {
      const double x = X[(size_t)r * SUR_K + 13 * SUR_CH + ch];
      s += x;
      s2 += x * x;
    }
    const double mu = s / SU.rows;
    double var = s2 / SU.rows - mu * mu;
    double sd = var > 1.0e-12 * (1.0 + mu * mu) ? sqrt(var) : 1.0;
    // log dt (channel 5) is nearly constant within one training run; a floor
    // of 1 keeps a step with a somewhat different dt inside the input range
    if (ch == 5 && sd < 1.0) sd = 1.0;
    SU.nm.mu[ch] = ch < SUR_NCH ? (float)mu : 0.0f;
    SU.nm.sd[ch] = ch < SUR_NCH ? (float)sd : 1.0f;
  }
  double sy2 = 0.0;
  for (int r = 0; r < SU.rows; ++r) sy2 += (double)y[r] * y[r];
  const double ys = sqrt(sy2 / SU.rows);
  SU.nm.y_scale = ys > 0.0 ? (float)ys : 1.0f;
  SU.have_norm = true;
  LAUNCH("sur_standardise", k_sur_standardise, dim3(blocks_for((size_t)SU.rows * SUR_K)), dim3(256), SU.X, SU.y,
         SU.rows, SU.nm);
  SU.standardised = true;
  return 0;
}

int api_sur_train(int32_t steps, int32_t batch, float lr, uint64_t seed, float *loss_out) {
  if (!S.ready) return fail(LARE_E_STATE, "not initialised");
  if (!SU.have_params) return fail(LARE_E_STATE, "surrogate train: no initial parameters");
  if (!SU.standardised) return fail(LARE_E_STATE, "surrogate train: call fit_norm after collecting rows");
  if (steps <= 0 || batch < 16 || batch % 16 != 0) return fail(LARE_E_ARG, "surrogate train: steps %d batch %d", steps, batch);
  int rc = sur_ensure_batch(batch);
  if (rc) return rc;
  if (steps > SU.losscap) {
    sfree(SU.losshist);
    rc = salloc(&SU.losshist, (size_t)steps);
    if (rc) return rc;
    SU.losscap = steps;
  }
  const float b1 = 0.9f, b2 = 0.999f;
  for (int s = 0; s < steps; ++s) {
    const unsigned long long bseed = mix64(seed * 0x9E3779B97F4A7C15ull + (unsigned long long)s);
    LAUNCH("sur_batch", k_sur_batch, dim3(blocks_for((size_t)batch * SUR_K)), dim3(256), SU.X, SU.y, SU.rows, bseed,
           batch, SU.Xb, SU.yb);
    rc = sur_forward(batch, batch, SU.yb);
    if (rc) return rc;
    LAUNCH("sur_loss", k_sum_into, dim3(1), dim3(64), SU.lossacc, (int)blocks_for(batch), 1.0f / (float)batch,
           SU.losshist + s);
    rc = sur_backward(batch);
    if (rc) return rc;
    SU.adam_t += 1;
    const float c1 = 1.0f - powf(b1, (float)SU.adam_t), c2 = 1.0f - powf(b2, (float)SU.adam_t);
    LAUNCH("sur_adam", k_adam, dim3(blocks_for(SUR_NPARAM)), dim3(256), SU.w, SU.gr, SU.m, SU.v, SUR_NPARAM, lr, b1,
           b2, 1.0e-8f, c1, c2);
  }
  rc = sur_pack();
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(S.stream));
  if (loss_out) HIPCHK(hipMemcpy(loss_out, SU.losshist, sizeof(float) * steps, hipMemcpyDeviceToHost));
  return 0;
}

int api_sur_forward(const float *X, int32_t rows, float *y) {
  if (!S.ready) return fail(LARE_E_STATE, "not initialised");
  if (!SU.have_params) return fail(LARE_E_STATE, "surrogate forward: no parameters");
  if (!X || !y || rows <= 0) return fail(LARE_E_ARG, "surrogate forward: bad arguments");
  int nb;
  int rc = sur_stage_rows(X, nullptr, rows, &nb);
  if (rc) return rc;
  rc = sur_forward(nb, rows, nullptr);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(S.stream));
  HIPCHK(hipMemcpy(y, SU.yo, sizeof(float) * rows, hipMemcpyDeviceToHost));
  return 0;
}

int api_sur_grad(const float *X, const float *t, int32_t rows, float *grad, float *loss) {
  if (!S.ready) return fail(LARE_E_STATE, "not initialised");
  if (!SU.have_params) return fail(LARE_E_STATE, "surrogate grad: no parameters");
  if (!X || !t || !grad || rows <= 0) return fail(LARE_E_ARG, "surrogate grad: bad arguments");
  int nb;
  int rc = sur_stage_rows(X, t, rows, &nb);
  if (rc) return rc;
  rc = sur_forward(nb, rows, SU.yb);
  if (rc) return rc;
  rc = sur_backward(nb);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(S.stream));
  HIPCHK(hipMemcpy(grad, SU.gr, sizeof(float) * SUR_NPARAM, hipMemcpyDeviceToHost));
  if (loss) {
    const int nblk = (int)blocks_for(nb);
    std::vector<float> part(nblk);
    HIPCHK(hipMemcpy(part.data(), SU.lossacc, sizeof(float) * nblk, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (float p : part) s += p;
    *loss = (float)(s / rows);
  }
  return 0;
}

int api_sur_enable(int32_t on, double tol) {
  if (!(tol > 0.0)) return fail(LARE_E_ARG, "surrogate tolerance must be positive");
  if (on && S.pi_eos) return fail(LARE_E_ARG, "the conduction surrogate is trained for EOS_IDEAL only");
  if (on && S.any_open) return fail(LARE_E_ARG, "the conduction surrogate does not handle BC_OPEN sides");
  if (on && (!SU.have_params || !SU.have_norm)) return fail(LARE_E_STATE, "surrogate enable: no trained parameters");
  SU.enabled = on != 0;
  SU.tol = tol;
  return 0;
}

int api_sur_conduct(double *residual, int32_t *accepted) {
  REQUIRE_LOADED();
  if (S.p.kappa_0 <= 0.0) return fail(LARE_E_ARG, "surrogate conduct needs kappa_0 > 0");
  int acc = 0;
  int rc = sur_conduct_step(0, residual, &acc);
  if (accepted) *accepted = acc;
  return rc;
}

int api_sur_plain_residual(double *residual) {
  REQUIRE_LOADED();
  if (S.p.kappa_0 <= 0.0) return fail(LARE_E_ARG, "plain residual needs kappa_0 > 0");
  if (!residual) return fail(LARE_E_ARG, "plain residual: NULL output");
  int rc = sur_alloc_fields();
  if (!rc) rc = calc_s_stages();
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(SU.E0save, S.energy, sizeof(real) * S.ntot, hipMemcpyDeviceToDevice, S.stream));
  rc = heat_conduct_sts2(0);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(SU.E1, S.energy, sizeof(real) * S.ntot, hipMemcpyDeviceToDevice, S.stream));
  HIPCHK(hipMemcpyAsync(S.energy, SU.E0save, sizeof(real) * S.ntot, hipMemcpyDeviceToDevice, S.stream));
  return sur_residual(SU.E1, residual);
}

int api_sur_stats(int32_t *accepted, int32_t *rejected, double *last_residual) {
  if (accepted) *accepted = SU.accepted;
  if (rejected) *rejected = SU.rejected;
  if (last_residual) *last_residual = SU.last_res;
  return 0;
}

// lare_dev_step followed by the scalars a driver reads after it, and
// optionally energy_account of the new state, read back in one copy.
int api_step_info(int32_t step, int32_t with_energy, lare_step_info *o) {
  REQUIRE_LOADED();
  if (!o) return fail(LARE_E_ARG, "step_info: NULL output");
  double dt = 0.0, t = 0.0;
  int rc = api_step(step, &dt, &t);
  if (rc) return rc;
  if (with_energy) {
    dim3 gr = grid_from0(nx(), ny(), nz());
    LAUNCH("energy", (k_energy<PadLayout, real>), gr, box_block(), S.g, PadLayout{S.g}, S.rho, S.energy, S.vx, S.vy,
           S.vz, S.bx, S.by, S.bz, S.partial);
    rc = final_reduce(gr, 4, 8, 0, 0);
    if (rc) return rc;
  }
  double v[10];
  rc = read_scalars(0, 10, v);
  if (rc) return rc;
  o->dt = S.dt;
  o->dtr = S.dtr;
  o->time = S.time;
  o->total_visc_heating = v[8];
  o->total_ohmic_heating = v[9];
  o->en_b = with_energy ? v[0] : 0.0;
  o->en_ke = with_energy ? v[1] : 0.0;
  o->en_int = with_energy ? v[2] : 0.0;
  o->max_divb = with_energy ? -v[3] : 0.0;
  o->n_s_stages = S.n_s_stages;
  o->substeps = S.last_substeps;
  o->peak_substeps = S.peak_substeps;
  o->sur_accepted = SU.accepted;
  o->sur_rejected = SU.rejected;
  o->with_energy = with_energy ? 1 : 0;
  return 0;
}

// initial_conditions.f90:114-296 potential_field.  phi lives in the FP64
// scratch array F, the flux patch in F2.
int api_potential_field(const lare_potential *pf, int32_t *iterations, double *residual, double *bz_max) {
  REQUIRE_LOADED();
  if (!pf || !pf->flux0) return fail(LARE_E_ARG, "potential_field: NULL argument");
  const int N0 = nx(), N1 = ny(), N2 = nz();
  if (S.per[2]) return fail(LARE_E_ARG, "potential_field: the z axis must be BC_USER (flux patch below)");
  if (S.any_open) return fail(LARE_E_ARG, "potential_field: BC_OPEN sides are not supported");
  if (N2 > N1 + 2)
    return fail(LARE_E_ARG, "potential_field: nz = %d > ny + 2 = %d (the reference reads bz(1:nx, 1:nz, 0))", N2,
                N1 + 2);
  if (pf->max_iterations < 1) return fail(LARE_E_ARG, "potential_field: max_iterations < 1");
  double *phi = S.F, *flux0 = S.F2;
  HIPCHK(hipStreamSynchronize(S.stream));
  HIPCHK(hipMemcpy(flux0, pf->flux0, sizeof(double) * (size_t)N0 * N1, hipMemcpyHostToDevice));
  LAUNCH_BOX("fill", k_fill64, full_box(), phi, 0.0);
  LAUNCH("sor_reset", k_sor_reset, dim3(1), dim3(64), S.dscal);
  const int ncol = (N0 + 1) / 2;
  const dim3 gs((unsigned)(((long)ncol * N1 + BX * BY - 1) / (BX * BY)));
  const Box zb = mkbox(-1, N0 + 2, -1, N1 + 2, 0, 0);
  const double w = pf->omega, dx1 = pf->dx1, dy1 = pf->dy1, dz1 = pf->dz1;
  auto phi_bcs = [&]() -> int {
    // phi_mpi: exchanges on periodic axes, the z boundaries, then the
    // zero-gradient x and y walls
    for (int ax = 0; ax < 3; ++ax) {
      if (!S.per[ax]) continue;
      BcSet b = {};
      b.nf = 1;
      b.f[0] = phi;
      b.op[0] = BC_PER_CELL;
      b.wide[0] = 1;
      int rc = launch_bc_axis(b, ax, "bc_phi");
      if (rc) return rc;
    }
    LAUNCH("phi_zbc", k_phi_zbc, box_grid(zb), box_block(), S.g, zb, phi, flux0);
    for (int ax = 0; ax < 2; ++ax) {
      if (S.per[ax]) continue;
      BcSet b = {};
      b.nf = 1;
      b.f[0] = phi;
      b.op[0] = BC_PHI_NEU;
      b.wide[0] = 1;
      int rc = launch_bc_axis(b, ax, "bc_phi");
      if (rc) return rc;
    }
    return 0;
  };
  const int batch = 32;
  double h[5] = {0, 0, 0, 0, 0};
  for (;;) {
    for (int q = 0; q < batch; ++q) {
      for (int z1 = 1; z1 <= 2; ++z1) {
        LAUNCH("sor", k_sor_pass, gs, box_block(), S.g, phi, z1, w, dx1, dy1, dz1, S.dscal + 12, S.partial);
        int rc = final_reduce(gs, 1, 1, 9 + z1, 0);
        if (rc) return rc;
        rc = phi_bcs();
        if (rc) return rc;
      }
      LAUNCH("sor_check", k_sor_check, dim3(1), dim3(64), S.dscal, pf->fractional_error, pf->max_iterations);
    }
    int rc = read_scalars(10, 5, h);
    if (rc) return rc;
    if (h[2] != 0.0) break;
  }
  // B = -grad phi, bfield_bcs, scale by the largest bz below, bfield_bcs
  LAUNCH_BOX("phi_to_b", k_phi_to_b, full_box(), phi, S.bx, S.by, S.bz);
  int rc = bfield_bcs(S.bx, S.by, S.bz);
  if (rc) return rc;
  const Box lb = mkbox(1, N0, 1, N2, 0, 0);
  const dim3 gb = box_grid(lb);
  LAUNCH("bz_lower", k_bz_lower, gb, box_block(), S.g, lb, S.bz, S.partial);
  rc = final_reduce(gb, 1, 1, 15, 0);
  if (rc) return rc;
  LAUNCH_BOX("scale_b", k_scale_b, full_box(), S.bx, S.by, S.bz, S.dscal);
  rc = bfield_bcs(S.bx, S.by, S.bz);
  if (rc) return rc;
  double bm;
  rc = read_scalars(15, 1, &bm);
  if (rc) return rc;
  if (iterations) *iterations = (int32_t)h[3];
  if (residual) *residual = h[4];
  if (bz_max) *bz_max = -bm;
  return 0;
}
