// initial.hpp -- potential_field (initial_conditions.f90:114-296) on the
// device: red-black SOR for the potential phi of a current-free field,
// B = -grad phi, scaled by the largest bz on the lower z boundary.
//
// The reference's colouring depends on (ix + iy) only: x1 restarts at z1
// for every iz (:141-160), so the two passes are checkerboards in x-y and
// each z column of a colour is relaxed from iz = 1 upwards, the point below
// already updated in the same pass.  One thread per column of the colour
// does exactly that; the x and y neighbours belong to the other colour and
// do not change during the pass, so the result is that of the reference's
// sequential loops, bit for bit.  Iterations run in batches without a host
// synchronisation: a flag in device memory stops the relaxation once the
// largest residual falls below the tolerance (or the iteration limit is
// reached), and the remaining launches of the batch change nothing (the
// boundary fills are idempotent).

// One colour pass: z1 = 1 relaxes (ix + iy) even, z1 = 2 (ix + iy) odd.
// flag[0] != 0 means converged: the pass does nothing.  Per-block partial
// of -max |residual| (for the min reduction).
__global__ __launch_bounds__(BX * BY) void k_sor_pass(Geo g, double *__restrict__ phi, int z1, double w, double dx1,
                                                      double dy1, double dz1, const double *__restrict__ flag,
                                                      double *__restrict__ partial) {
  const int ncol = (g.nx + 1) / 2;  // most points of one colour in a row
  const int q = (int)blockIdx.x * (BX * BY) + (int)threadIdx.y * BX + (int)threadIdx.x;
  const int iy = q / ncol + 1, t = q - (iy - 1) * ncol;
  const int x1 = ((iy - 1) & 1) ? 3 - z1 : z1;  // first ix of row iy
  const int ix = x1 + 2 * t;
  double v[1] = {-0.0};
  if (flag[0] == 0.0 && iy <= g.ny && ix <= g.nx) {
    const int Y = g.sy, Z = g.sz;
    int c = g.idx(ix, iy, 1);
    double below = phi[c - Z], cur = phi[c];
    double em = 0.0;
    const double den = dx1 + dy1 + dz1;
    for (int iz = 1; iz <= g.nz; ++iz, c += Z) {
      const double above = phi[c + Z];
      const double r = (((phi[c + 1] - 2.0 * cur) + phi[c - 1]) * dx1 +
                        ((phi[c + Y] - 2.0 * cur) + phi[c - Y]) * dy1 + ((above - 2.0 * cur) + below) * dz1) /
                       2.0 / den;
      cur = cur + w * r;
      phi[c] = cur;
      em = dmax(em, fabs(r));
      below = cur;
      cur = above;
    }
    v[0] = -em;
  }
  const int ops[1] = {0};
  block_reduce_store<1>(v, ops, partial);
}

// The lower and upper z boundaries of phi_mpi (:268-275): the flux patch
// phi(ix, iy, 0) = phi(ix, iy, 1) + flux0(ix, iy), phi(ix, iy, -1) =
// phi(ix, iy, 0) on 1..nx x 1..ny, and phi = 0 on planes nz+1, nz+2 over
// the whole x-y extent.
__global__ __launch_bounds__(BX * BY) void k_phi_zbc(Geo g, Box b, double *__restrict__ phi,
                                                     const double *__restrict__ flux0) {
  PLANE_POS(b, i, j)
  if (j > b.y1) return;
  if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny) {
    const int c0 = g.idx(i, j, 0);
    const double p0 = phi[c0 + g.sz] + flux0[(i - 1) + g.nx * (j - 1)];
    phi[c0] = p0;
    phi[c0 - g.sz] = p0;
  }
  phi[g.idx(i, j, g.nz + 1)] = 0.0;
  phi[g.idx(i, j, g.nz + 2)] = 0.0;
}

// End of one SOR iteration: the largest residual of both passes; set the
// flag when it is below the tolerance or the iteration limit is reached.
// s[10], s[11]: -max of the passes; s[12] flag; s[13] iterations; s[14]
// the last largest residual.
__global__ void k_sor_check(double *__restrict__ s, double tol, int maxit) {
  if (threadIdx.x != 0 || s[12] != 0.0) return;
  const double err = dmax(-s[10], -s[11]);
  const double it = s[13] + 1.0;
  s[13] = it;
  s[14] = err;
  if (err < tol || it >= (double)maxit) s[12] = 1.0;
}

__global__ void k_sor_reset(double *__restrict__ s) {
  if (threadIdx.x < 5) s[10 + threadIdx.x] = 0.0;
}

__global__ __launch_bounds__(BX * BY) void k_fill64(Geo g, Box b, double *__restrict__ a, double v) {
  BOX_INDEX(b, i, j, k);
  a[g.idx(i, j, k)] = v;
}

// B = -grad phi on the faces the reference sets (:198-218): bx on
// 0..nx x 1..ny x 1..nz, by on 1..nx x 0..ny x 1..nz, bz on
// 1..nx x 1..ny x 0..nz; every other face keeps its value.
__global__ __launch_bounds__(BX * BY) void k_phi_to_b(Geo g, Box b, const double *__restrict__ phi,
                                                      double *__restrict__ bx, double *__restrict__ by,
                                                      double *__restrict__ bz) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const bool in_x = i >= 1 && i <= g.nx, in_y = j >= 1 && j <= g.ny, in_z = k >= 1 && k <= g.nz;
  if (i >= 0 && i <= g.nx && in_y && in_z) bx[c] = -(phi[c + 1] - phi[c]) / g.dxc[i];
  if (in_x && j >= 0 && j <= g.ny && in_z) by[c] = -(phi[c + g.sy] - phi[c]) / g.dyc[j];
  if (in_x && in_y && k >= 0 && k <= g.nz) bz[c] = -(phi[c + g.sz] - phi[c]) / g.dzc[k];
}

// -bz on the lower boundary plane bz(1:nx, 1:nz, 0) (the y range is 1:nz
// in the reference, :221): per-block partial of the minimum.
__global__ __launch_bounds__(BX * BY) void k_bz_lower(Geo g, Box b, const double *__restrict__ bz,
                                                      double *__restrict__ partial) {
  PLANE_POS(b, i, j)
  double v[1] = {LARGEST};
  if (j <= b.y1) v[0] = -bz[g.idx(i, j, 0)];
  const int ops[1] = {0};
  block_reduce_store<1>(v, ops, partial);
}

// B / bz_max over whole arrays (:224-226); s[15] holds -bz_max.
__global__ __launch_bounds__(BX * BY) void k_scale_b(Geo g, Box b, double *__restrict__ bx, double *__restrict__ by,
                                                     double *__restrict__ bz, const double *__restrict__ s) {
  BOX_INDEX(b, i, j, k);
  const int c = g.idx(i, j, k);
  const double bm = -s[15];
  bx[c] = bx[c] / bm;
  by[c] = by[c] / bm;
  bz[c] = bz[c] / bm;
}
