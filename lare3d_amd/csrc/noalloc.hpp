// noalloc.hpp -- lare_dev_energy_account_noalloc: energy_account on the
// caller's own device arrays, with no device allocation by the library.
//
// For callers that cannot call hipMalloc (code under HIP graph capture,
// tight serving loops, harnesses that forbid it).  The kernels are those of
// lare_dev_energy_account (k_energy with the Fortran layout, then the same
// two-level fold of the per-block partials), so the results are the same
// bits for the same field values.  Instead of the partial buffer that
// lare_dev_init allocates, they use `scratch`, a fixed array in device
// memory that each GPU holds from the library's load to the end of the
// process:
//   [0, NA_MET)            dxb, dyb, dzb copied from the host grid
//   [NA_MET, +NA_PART)     per-block partial sums, 4 doubles per block
//   [.., +NA_LEVEL1)       level-1 results of the fold
//   [.., +NA_OUT)          the four final sums
// Every call on one GPU uses that GPU's one copy of the array and leaves
// its last contents there: calls that overlap on one GPU (several host
// threads, or streams the caller runs concurrently) race.  Nothing in it
// needs to start at zero -- every value read is written earlier in the
// same call.
//
// Included once by lare_dev.hip, after the fp64 runtime.

#define NA_NMAX 4096                              // cells per axis
#define NA_LEN (NA_NMAX + 8)                       // one metric array, index i + 2
#define NA_MET (3 * NA_LEN)
#define NA_MAX_BLOCKS 262144                       // energy grid blocks (n+1)^3 / 256
#define NA_PART (NA_MAX_BLOCKS * 4)
#define NA_LEVEL1 (RED_ROWS * RED_MAXV)
#define NA_OUT 8
#define NA_TOTAL (NA_MET + NA_PART + NA_LEVEL1 + NA_OUT)

static __device__ __align__(256) double scratch[NA_TOTAL];

namespace {

// Device address of `scratch` on each GPU (the host symbol is only a
// stand-in): looked up once per device with that device current.
#define NA_MAXDEV 64
double *g_na_ptr[NA_MAXDEV] = {};

int na_scratch(double **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAXDEV) return fail(LARE_E_ARG, "energy_account_noalloc: device %d above %d", dev, NA_MAXDEV);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_na_ptr[dev] = (double *)p;
  }
  *out = g_na_ptr[dev];
  return 0;
}

int api_energy_account_noalloc(const lare_grid *grid, const lare_fields *f, double en[3], double *divb) {
  // every argument and size is checked before the first HIP call
  if (!grid || !f || !en) return fail(LARE_E_ARG, "energy_account_noalloc: NULL argument");
  if (!f->rho || !f->energy || !f->vx || !f->vy || !f->vz || !f->bx || !f->by || !f->bz)
    return fail(LARE_E_ARG, "energy_account_noalloc: missing field pointer");
  if (!grid->dxb || !grid->dyb || !grid->dzb) return fail(LARE_E_ARG, "energy_account_noalloc: metric arrays missing");
  const int n[3] = {grid->nx, grid->ny, grid->nz};
  for (int a = 0; a < 3; ++a)
    if (n[a] < 4 || n[a] > NA_NMAX)
      return fail(LARE_E_ARG, "energy_account_noalloc: grid size %d on axis %d outside 4..%d", n[a], a, NA_NMAX);
  const long nodes = (long)(n[0] + 5) * (n[1] + 5) * (n[2] + 5);
  if (nodes > 2147483647L) return fail(LARE_E_ARG, "energy_account_noalloc: grid too large for 32-bit indexing");
  const fp64::Box b = fp64::mkbox(0, n[0], 0, n[1], 0, n[2]);
  const dim3 gr = fp64::box_grid(b);
  const long nb = (long)gr.x * gr.y * gr.z;
  if (nb > NA_MAX_BLOCKS)
    return fail(LARE_E_ARG,
                "energy_account_noalloc: %ld reduction blocks exceed the scratch maximum of %d "
                "(about (n+1)^3 <= 6.7e7 nodes)", nb, NA_MAX_BLOCKS);

  double *sp = nullptr;
  int rc = na_scratch(&sp);
  if (rc) return rc;
  double *met = sp, *partial = sp + NA_MET, *level1 = partial + NA_PART, *out = level1 + NA_LEVEL1;
  // dxb(-1:n+2) of each axis to met[a * NA_LEN + (i + 2)]
  const double *src[3] = {grid->dxb, grid->dyb, grid->dzb};
  for (int a = 0; a < 3; ++a)
    HIPCHK(hipMemcpy(met + (size_t)a * NA_LEN + 1, src[a], sizeof(double) * (n[a] + 4), hipMemcpyHostToDevice));
  fp64::Geo g = {};
  g.nx = n[0];
  g.ny = n[1];
  g.nz = n[2];
  g.dxb = met + 2;
  g.dyb = met + NA_LEN + 2;
  g.dzb = met + 2 * NA_LEN + 2;
  const fp64::FtnLayout lay = {n[0], n[1], n[2]};
  hipLaunchKernelGGL((fp64::k_energy<fp64::FtnLayout, double>), gr, fp64::box_block(), 0, 0, g, lay, f->rho, f->energy,
                     f->vx, f->vy, f->vz, f->bx, f->by, f->bz, partial);
  HIPCHK(hipGetLastError());
  // the fold of fp64::final_reduce, on the scratch array
  const int nbi = (int)nb, nv = 4, minmask = 8;
  const int chunk = (nbi + RED_ROWS - 1) / RED_ROWS;
  const int rows = (nbi + chunk - 1) / chunk;
  hipLaunchKernelGGL(fp64::k_reduce_rows, dim3(rows), dim3(256), 0, 0, partial, nbi, nv, minmask, chunk, level1);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(fp64::k_final_reduce, dim3(1), dim3(256), 0, 0, level1, rows, nv, minmask, 0, out);
  HIPCHK(hipGetLastError());
  double v[4];
  HIPCHK(hipMemcpy(v, out, sizeof(v), hipMemcpyDeviceToHost));
  en[0] = v[0];
  en[1] = v[1];
  en[2] = v[2];
  if (divb) *divb = -v[3];
  return 0;
}

}  // namespace
