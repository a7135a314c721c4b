// common.hpp -- device layout and launch helpers shared by all kernels.
//
// Every 3-D field lives in one padded box covering Fortran indices
// -2..n+2 on each axis (the union of the cell, node and face bounds of
// mpi_routines.F90:295-334), x fastest.  Index (i, j, k) maps to
//     (k + 2) * sz + (j + 2) * sy + (i + OX)
// with OX = 16 so that i = 0 sits on a 128-byte boundary for FP64 rows and
// the row pitch sy is a multiple of 16 elements.  All fields share the
// same index arithmetic, which lets one kernel address cell, node and face
// arrays with one offset.


#define OX 16

struct Geo {
  int nx, ny, nz;
  int sy, sz;
  // 1-D metrics, indexable by Fortran index (pointer already offset)
  const double *dxb, *dxc, *dyb, *dyc, *dzb, *dzc, *grav;
  int per[3];     // axis periodic (both sides)
  int user[3];    // axis not periodic (each side BC_USER or BC_OPEN)
  int open_lo[3], open_hi[3];  // that side BC_OPEN: no fill, ghosts as stored
  __host__ __device__ __forceinline__ int idx(int i, int j, int k) const {
    return (k + 2) * sz + (j + 2) * sy + (i + OX);
  }
  __host__ __device__ __forceinline__ int n(int a) const { return a == 0 ? nx : (a == 1 ? ny : nz); }
  __host__ __device__ __forceinline__ int stride(int a) const { return a == 0 ? 1 : (a == 1 ? sy : sz); }
  __device__ __forceinline__ double db(int a, int i) const { return a == 0 ? dxb[i] : (a == 1 ? dyb[i] : dzb[i]); }
  __device__ __forceinline__ double dc(int a, int i) const { return a == 0 ? dxc[i] : (a == 1 ? dyc[i] : dzc[i]); }
  // cell volume of setup.F90:242-246, recomputed from the metrics
  __device__ __forceinline__ double cv(int i, int j, int k) const { return dxb[i] * dyb[j] * dzb[k]; }
};

// Index box [lo, hi] inclusive per axis.
struct Box {
  int x0, x1, y0, y1, z0, z1;
};

#define BX 64
#define BY 4

// In-plane position (i, j) of a thread of a box launch.  The (x, y) extent
// of the box is walked as one flattened index, BX*BY consecutive points per
// block, so a box a few points wider than a multiple of 64 (n+1 .. n+4 for
// n = 256) does not launch a column of blocks whose waves are almost empty.
// i always lies inside [x0, x1]; j may pass y1 in the last block.
//
// Blocks are dispatched to the eight XCDs round-robin by linear block id,
// and each XCD has its own L2.  Tile t of a plane goes to block
// xcd_tile(t): the tiles that share an XCD (block ids congruent mod 8) get
// consecutive plane positions, so the rows a tile's stencil reads from its
// neighbour tiles are mostly in the same L2.
__device__ __forceinline__ unsigned xcd_tile(unsigned bx, unsigned gx) {
  const unsigned q = gx >> 3, r = gx & 7u, t = bx & 7u;
  return t * q + (t < r ? t : r) + (bx >> 3);
}
#define PLANE_POS(b, i, j)                                                                  \
  const unsigned pw_ = (unsigned)((b).x1 - (b).x0 + 1);                                     \
  const unsigned pp_ = xcd_tile(blockIdx.x, gridDim.x) * (unsigned)(BX * BY) + threadIdx.y * (unsigned)BX + threadIdx.x; \
  const unsigned pj_ = pp_ / pw_;                                                           \
  const int i = (b).x0 + (int)(pp_ - pj_ * pw_);                                            \
  const int j = (b).y0 + (int)pj_;

// Thread (i, j, k) of a box launch; returns false outside the box.
#define BOX_INDEX(b, i, j, k)                  \
  PLANE_POS(b, i, j)                           \
  const int k = (b).z0 + (int)blockIdx.z;      \
  if (j > (b).y1 || k > (b).z1) return;

static inline unsigned plane_blocks(const Box &b) {
  const long w = b.x1 - b.x0 + 1, h = b.y1 - b.y0 + 1;
  return (unsigned)((w * h + BX * BY - 1) / (BX * BY));
}
static inline dim3 box_grid(const Box &b) { return dim3(plane_blocks(b), 1, (unsigned)(b.z1 - b.z0 + 1)); }
static inline dim3 box_block() { return dim3(BX, BY, 1); }
static inline Box mkbox(int x0, int x1, int y0, int y1, int z0, int z1) {
  Box b = {x0, x1, y0, y1, z0, z1};
  return b;
}

__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ float dmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float dmax(float a, float b) { return a > b ? a : b; }

// Fortran MAX(x, 0.0_num) as used on energy and visc_heat.
__device__ __forceinline__ double max0(double x) { return x > 0.0 ? x : 0.0; }
__device__ __forceinline__ float max0(float x) { return x > 0.0f ? x : 0.0f; }

// Wavefront (64-lane) reductions.
__device__ __forceinline__ double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_down(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block-level reduction of NV values into partial[blockid * NV + q].
// op 0 = min, 1 = sum.  Block size BX*BY = 256 (4 waves).
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], const int *ops, double *partial) {
  __shared__ double red[4][NV];
  const int lane = threadIdx.x & 63;
  const int wave = (threadIdx.y * BX + threadIdx.x) >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = ops[q] == 0 ? wave_min(v[q]) : wave_sum(v[q]);
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) red[wave][q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int bid = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double r = red[0][q];
      for (int w = 1; w < 4; ++w) r = ops[q] == 0 ? dmin(r, red[w][q]) : r + red[w][q];
      partial[(size_t)bid * NV + q] = r;
    }
  }
}

// Columns of a z-marching launch: thread (i, j) of the flattened plane tile
// blockIdx.x walks planes k0 .. k1 of its chunk (ZCHUNK planes per block).
#define ZCHUNK 8
#define ZRANGE(b, k0, k1)                                             \
  const int k0 = (b).z0 + (int)blockIdx.z * ZCHUNK;                    \
  const int k1 = k0 + ZCHUNK - 1 < (b).z1 ? k0 + ZCHUNK - 1 : (b).z1;
#define ZCOLUMN(b, i, j, k0, k1) \
  PLANE_POS(b, i, j)             \
  ZRANGE(b, k0, k1)              \
  if (j > (b).y1) return;

// Columns of a launch that marches along axis M (1 = y, 2 = z) for ZCHUNK
// points: the flattened tile covers x and the other transverse axis, whose
// index is p; MJ/MK give j and k at march coordinate m.  The grid is the
// zchunk grid of the box with its y and z ranges swapped when M = 1.
#define MCOLUMN(M, b, i, p, m0, m1)                                                         \
  const Box pb_ = (M) == 2 ? (b) : Box{(b).x0, (b).x1, (b).z0, (b).z1, (b).y0, (b).y1};   \
  PLANE_POS(pb_, i, p)                                                                     \
  ZRANGE(pb_, m0, m1)                                                                      \
  if (p > pb_.y1) return;
#define MJ(M, m, p) ((M) == 1 ? (m) : (p))
#define MK(M, m, p) ((M) == 2 ? (m) : (p))

// A field at the four in-plane points (i, j), (i-1, j), (i, j-1), (i-1, j-1).
struct Quad {
  real p0, px, py, pxy;
};
__device__ __forceinline__ Quad quad_back(const real *__restrict__ u, int c, int Y) {
  Quad q = {u[c], u[c - 1], u[c - Y], u[c - 1 - Y]};
  return q;
}

// A field at the four in-plane points (i, j), (i+1, j), (i, j+1), (i+1, j+1).
__device__ __forceinline__ Quad quad_fwd(const real *__restrict__ u, int c, int Y) {
  Quad q = {u[c], u[c + 1], u[c + Y], u[c + 1 + Y]};
  return q;
}
