// bc.hpp -- boundary / halo fill on one GPU.
//
// Single-rank semantics of boundary.f90 + mpiboundary.f90: on a periodic
// axis every MPI_SENDRECV is a self-exchange of 2- or 3-layer slabs that
// span the full extent of the other two axes (so corners wrap); on a
// BC_USER axis the MPI exchange is a no-op and the user mirrors / zero
// walls of boundary.f90 apply.  The reference runs all exchanges (x, y, z)
// before all user fills (x, y, z); the host launches one kernel per axis in
// that order.  One launch handles up to 8 fields so that a whole
// boundary_conditions call is at most three launches.  A BC_OPEN side
// (openboundary.f90) gets no fill at all: the MPI exchange is a no-op there
// and the user fills of boundary.f90 apply to BC_USER sides only, so the
// ghosts keep the far-field values until open_bcs sets the first layer.

enum BcOp {
  BC_NONE = 0,
  BC_PER_CELL,   // (n+1:n+2) <- (1:2), (-1:0) <- (n-1:n)      cell_xface
  BC_PER_FACE,   // (n+1:n+2) <- (1:2), (-2:0) <- (n-2:n)      bx_xface / bx_xface1
  BC_PER_NODE,   // (n:n+2) <- (0:2), (-2:-1) <- (n-2:n-1)     node_xface1 / node_xface
  BC_MIR_CELL,   // 0 <- 1, -1 <- 2, n+1 <- n, n+2 <- n-1     boundary.f90:134-160 etc.
  BC_MIR_FACE,   // -1 <- 1, -2 <- 2, n+1 <- n-1, n+2 <- n-2  boundary.f90:68-69,77-78
  BC_ZERO_NODE,  // -2..0 = 0, n..n+2 = 0                      boundary.f90:260-292
  BC_DM_CLAMP,   // dm(n+1) <- dm(n), dm(-1) <- dm(0) over transverse 0..n+1 (xremap.f90:906-915)
  BC_PHI_NEU     // 0 <- 1, -1 <- 1, n+1 <- n, n+2 <- n     phi_mpi walls (initial_conditions.f90:276-291)
};

#define BC_MAXF 8
struct BcSet {
  void *f[BC_MAXF];
  int op[BC_MAXF];
  int wide[BC_MAXF];  // 1: FP64 array, 0: FP32 array
  int nf;
  int skip_lo, skip_hi;  // leave the low / high side alone (a BC_OPEN side)
};

template <class T>
__device__ __forceinline__ void bc_line(T *a, int op, int base, int st, int n, bool dm_ok, bool lo, bool hi) {
  switch (op) {
    case BC_PER_CELL:
      a[base + (n + 1) * st] = a[base + 1 * st];
      a[base + (n + 2) * st] = a[base + 2 * st];
      a[base - 1 * st] = a[base + (n - 1) * st];
      a[base + 0 * st] = a[base + n * st];
      break;
    case BC_PER_FACE:
      a[base + (n + 1) * st] = a[base + 1 * st];
      a[base + (n + 2) * st] = a[base + 2 * st];
      a[base - 2 * st] = a[base + (n - 2) * st];
      a[base - 1 * st] = a[base + (n - 1) * st];
      a[base + 0 * st] = a[base + n * st];
      break;
    case BC_PER_NODE: {
      const T v0 = a[base], v1 = a[base + st], v2 = a[base + 2 * st];
      a[base + n * st] = v0;
      a[base + (n + 1) * st] = v1;
      a[base + (n + 2) * st] = v2;
      a[base - 2 * st] = a[base + (n - 2) * st];
      a[base - 1 * st] = a[base + (n - 1) * st];
      break;
    }
    case BC_MIR_CELL:
      if (lo) {
        a[base + 0 * st] = a[base + 1 * st];
        a[base - 1 * st] = a[base + 2 * st];
      }
      if (hi) {
        a[base + (n + 1) * st] = a[base + n * st];
        a[base + (n + 2) * st] = a[base + (n - 1) * st];
      }
      break;
    case BC_MIR_FACE:
      if (lo) {
        a[base - 1 * st] = a[base + 1 * st];
        a[base - 2 * st] = a[base + 2 * st];
      }
      if (hi) {
        a[base + (n + 1) * st] = a[base + (n - 1) * st];
        a[base + (n + 2) * st] = a[base + (n - 2) * st];
      }
      break;
    case BC_ZERO_NODE:
      if (lo) {
        a[base - 2 * st] = (T)0;
        a[base - 1 * st] = (T)0;
        a[base + 0 * st] = (T)0;
      }
      if (hi) {
        a[base + n * st] = (T)0;
        a[base + (n + 1) * st] = (T)0;
        a[base + (n + 2) * st] = (T)0;
      }
      break;
    case BC_PHI_NEU:
      a[base + 0 * st] = a[base + 1 * st];
      a[base - 1 * st] = a[base + 1 * st];
      a[base + (n + 1) * st] = a[base + n * st];
      a[base + (n + 2) * st] = a[base + n * st];
      break;
    case BC_DM_CLAMP:
      if (dm_ok) {
        a[base + (n + 1) * st] = a[base + n * st];
        a[base - 1 * st] = a[base + 0 * st];
      }
      break;
    default:
      break;
  }
}

__global__ __launch_bounds__(64) void k_bc_axis(Geo g, BcSet s, int axis) {
  // plane coordinates: u runs along the contiguous axis where possible
  const int u = axis == 0 ? 1 : 0;
  const int w = axis == 2 ? 1 : 2;
  const int iu = -2 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int iw = -2 + (int)blockIdx.y;
  if (iu > g.n(u) + 2) return;
  const int n = g.n(axis);
  const int st = g.stride(axis);
  int c[3];
  c[axis] = 0;
  c[u] = iu;
  c[w] = iw;
  const int base = g.idx(c[0], c[1], c[2]);  // index of layer 0 on this line
  const bool dm_ok = iu >= 0 && iu <= g.n(u) + 1 && iw >= 0 && iw <= g.n(w) + 1;
  const bool lo = !s.skip_lo, hi = !s.skip_hi;
  for (int q = 0; q < s.nf; ++q) {
    if (s.wide[q]) bc_line((double *)s.f[q], s.op[q], base, st, n, dm_ok, lo, hi);
    else bc_line((float *)s.f[q], s.op[q], base, st, n, dm_ok, lo, hi);
  }
}

// The x-axis fill with the work spread over the ghost slots of each line:
// thread (line, slot) does the slot-th (destination <- source) copy of the
// operation, slots 0..7, lines over the full (y, z) extent -2..n+2.  The
// copies of one operation touch disjoint destination and source layers
// (n >= 4), so doing them in parallel gives the sequential result, and the
// eight lanes of a line share its two end cache lines.
// Destination and source offsets (along the line) of one slot of a fill.
struct BcMove {
  int dst, src;
  bool zero, act;
};
__device__ __forceinline__ BcMove bc_slot(int op, int n, int slot, bool dm_ok) {
  int dst = 0, src = 0;
  bool zero = false, act = false;
  switch (op) {
    case BC_PER_CELL: {
      const int d[4] = {n + 1, n + 2, -1, 0}, r[4] = {1, 2, n - 1, n};
      if (slot < 4) { dst = d[slot]; src = r[slot]; act = true; }
      break;
    }
    case BC_PER_FACE: {
      const int d[5] = {n + 1, n + 2, -2, -1, 0}, r[5] = {1, 2, n - 2, n - 1, n};
      if (slot < 5) { dst = d[slot]; src = r[slot]; act = true; }
      break;
    }
    case BC_PER_NODE: {
      const int d[5] = {n, n + 1, n + 2, -2, -1}, r[5] = {0, 1, 2, n - 2, n - 1};
      if (slot < 5) { dst = d[slot]; src = r[slot]; act = true; }
      break;
    }
    case BC_MIR_CELL: {
      const int d[4] = {0, -1, n + 1, n + 2}, r[4] = {1, 2, n, n - 1};
      if (slot < 4) { dst = d[slot]; src = r[slot]; act = true; }
      break;
    }
    case BC_MIR_FACE: {
      const int d[4] = {-1, -2, n + 1, n + 2}, r[4] = {1, 2, n - 1, n - 2};
      if (slot < 4) { dst = d[slot]; src = r[slot]; act = true; }
      break;
    }
    case BC_ZERO_NODE: {
      const int d[6] = {-2, -1, 0, n, n + 1, n + 2};
      if (slot < 6) { dst = d[slot]; zero = true; act = true; }
      break;
    }
    case BC_PHI_NEU: {
      const int d[4] = {0, -1, n + 1, n + 2}, r[4] = {1, 1, n, n};
      if (slot < 4) { dst = d[slot]; src = r[slot]; act = true; }
      break;
    }
    case BC_DM_CLAMP: {
      const int d[2] = {n + 1, -1}, r[2] = {n, 0};
      if (slot < 2 && dm_ok) { dst = d[slot]; src = r[slot]; act = true; }
      break;
    }
    default:
      break;
  }
  BcMove m = {dst, src, zero, act};
  return m;
}
// the slot's move, unless it writes a side that is left alone
__device__ __forceinline__ BcMove bc_slot_side(int op, int n, int slot, bool dm_ok, const BcSet &s) {
  BcMove m = bc_slot(op, n, slot, dm_ok);
  if (m.act && (m.dst <= 0 ? s.skip_lo : s.skip_hi)) m.act = false;
  return m;
}

__global__ __launch_bounds__(256) void k_bc_x(Geo g, BcSet s) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int slot = t & 7, line = t >> 3;
  const int ly = g.ny + 5;
  if (line >= ly * (g.nz + 5)) return;
  const int j = -2 + line % ly, k = -2 + line / ly;
  const int base = g.idx(0, j, k);
  const bool dm_ok = j >= 0 && j <= g.ny + 1 && k >= 0 && k <= g.nz + 1;
  // The fields are distinct arrays, so all loads are issued before any
  // store (values pass through double, which is exact for FP32 too).
  BcMove mv[BC_MAXF];
  double v[BC_MAXF];
#pragma unroll
  for (int q = 0; q < BC_MAXF; ++q) {
    mv[q] = q < s.nf ? bc_slot_side(s.op[q], g.nx, slot, dm_ok, s) : BcMove{0, 0, false, false};
    v[q] = 0.0;
    if (mv[q].act && !mv[q].zero)
      v[q] = s.wide[q] ? ((const double *)s.f[q])[base + mv[q].src] : (double)((const float *)s.f[q])[base + mv[q].src];
  }
#pragma unroll
  for (int q = 0; q < BC_MAXF; ++q) {
    if (!mv[q].act) continue;
    if (s.wide[q]) ((double *)s.f[q])[base + mv[q].dst] = v[q];
    else ((float *)s.f[q])[base + mv[q].dst] = (float)v[q];
  }
}
