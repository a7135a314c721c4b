// lare_dev.hip -- C-ABI of the MI355X Lare3d hot path.
//
// One translation unit.  The runtime (runtime.hpp) and the kernels it
// includes are compiled twice, once per storage precision of the bulk
// fields: namespace fp64 (LARE_PREC_FP64, bit-compatible with the FP64
// reference arithmetic) and namespace fp32 (LARE_PREC_FP32: FP32 bulk
// storage, FP64 face B, CT fluxes, metrics, reductions and register
// arithmetic).  lare_dev_init selects the mode; every other entry point is
// forwarded to the initialised one.  The runtime owns every device array from
// lare_dev_init to lare_dev_finalize, queues all work on one HIP stream, and
// synchronises only where the reference's control flow needs a device scalar
// on the host (dt after set_dt, the RKL2 stage count, energy diagnostics).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/lare3d_dev.h"

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

}  // namespace

#define LARE_REAL double
#define LARE_ACC double
namespace fp64 {
#include "runtime.hpp"
}  // namespace fp64
#undef LARE_REAL
#undef LARE_ACC

#define LARE_REAL float
#define LARE_ACC float
namespace fp32 {
#include "runtime.hpp"
}  // namespace fp32
#undef LARE_REAL
#undef LARE_ACC

#include "noalloc.hpp"

namespace {
int g_prec = LARE_PREC_FP64;  // storage precision of the initialised state
}  // namespace

#define FWD(call) (g_prec == LARE_PREC_FP32 ? fp32::call : fp64::call)

extern "C" {

const char *lare_dev_last_error(void) { return g_err; }

int lare_dev_init(const lare_grid *grid, const lare_params *params) {
  if (!grid || !params) return fail(LARE_E_ARG, "lare_dev_init: NULL argument");
  if (params->precision == LARE_PREC_FP32) {
    fp64::api_finalize();
    g_prec = LARE_PREC_FP32;
    return fp32::api_init(grid, params);
  }
  if (params->precision != LARE_PREC_FP64) return fail(LARE_E_ARG, "precision mode %d not available", params->precision);
  fp32::api_finalize();
  g_prec = LARE_PREC_FP64;
  return fp64::api_init(grid, params);
}

int lare_dev_upload(const lare_fields *f) { return FWD(api_upload(f)); }
int lare_dev_download(lare_fields *f) { return FWD(api_download(f)); }
int lare_dev_get_scalars(lare_scalars *o) { return FWD(api_get_scalars(o)); }
int lare_dev_set_scalars(const lare_scalars *in) { return FWD(api_set_scalars(in)); }
int lare_dev_lagrangian_step(void) { return FWD(api_lagrangian_step()); }
int lare_dev_eulerian_remap(int32_t step) { return FWD(api_eulerian_remap(step)); }
int lare_dev_eta_calc(void) { return FWD(api_eta_calc()); }
int lare_dev_energy_correction(void) { return FWD(api_energy_correction()); }
int lare_dev_conduct_heat(void) { return FWD(api_conduct_heat()); }
int lare_dev_boundary_conditions(void) { return FWD(api_boundary_conditions()); }
int lare_dev_neutral_fraction(void) { return FWD(api_neutral_fraction()); }
int lare_dev_perpendicular_resistivity(void) { return FWD(api_perpendicular_resistivity()); }
int lare_dev_get_neutral(double *xi_n, double *eta_perp) { return FWD(api_get_neutral(xi_n, eta_perp)); }
int lare_dev_energy_account(double en[3], double *divb) { return FWD(api_energy_account(en, divb)); }
int lare_dev_step(int32_t step, double *dt, double *time) { return FWD(api_step(step, dt, time)); }
int lare_dev_step_info(int32_t step, int32_t with_energy, lare_step_info *info) {
  return FWD(api_step_info(step, with_energy, info));
}
int lare_dev_potential_field(const lare_potential *pf, int32_t *iterations, double *residual, double *bz_max) {
  return FWD(api_potential_field(pf, iterations, residual, bz_max));
}
int lare_dev_energy_account_noalloc(const lare_grid *grid, const lare_fields *dev_fields, double en[3], double *divb) {
  return api_energy_account_noalloc(grid, dev_fields, en, divb);
}
int lare_dev_profile(int32_t enable) { return FWD(api_profile(enable)); }
int lare_dev_kernel_stats(char *names, int32_t name_len, int32_t *counts, double *total_ms, int32_t cap) {
  return FWD(api_kernel_stats(names, name_len, counts, total_ms, cap));
}
int lare_dev_synchronize(void) { return FWD(api_synchronize()); }
int lare_dev_finalize(void) {
  const int a = fp64::api_finalize();
  const int b = fp32::api_finalize();
  return a ? a : b;
}

int lare_dev_sur_set_params(const float *params, int32_t n) { return FWD(api_sur_set_params(params, n)); }
int lare_dev_sur_get_params(float *params, int32_t n) { return FWD(api_sur_get_params(params, n)); }
int lare_dev_sur_set_norm(const float *mu, const float *sd, float y_scale) {
  return FWD(api_sur_set_norm(mu, sd, y_scale));
}
int lare_dev_sur_get_norm(float *mu, float *sd, float *y_scale) { return FWD(api_sur_get_norm(mu, sd, y_scale)); }
int lare_dev_sur_collect(int32_t count, uint64_t seed, int32_t *total) { return FWD(api_sur_collect(count, seed, total)); }
int lare_dev_sur_rows(float *X, float *y, int32_t rows) { return FWD(api_sur_rows(X, y, rows)); }
int lare_dev_sur_fit_norm(void) { return FWD(api_sur_fit_norm()); }
int lare_dev_sur_train(int32_t steps, int32_t batch, float lr, uint64_t seed, float *loss_out) {
  return FWD(api_sur_train(steps, batch, lr, seed, loss_out));
}
int lare_dev_sur_forward(const float *X, int32_t rows, float *y) { return FWD(api_sur_forward(X, rows, y)); }
int lare_dev_sur_grad(const float *X, const float *t, int32_t rows, float *grad, float *loss) {
  return FWD(api_sur_grad(X, t, rows, grad, loss));
}
int lare_dev_sur_enable(int32_t on, double tol) { return FWD(api_sur_enable(on, tol)); }
int lare_dev_sur_conduct(double *residual, int32_t *accepted) { return FWD(api_sur_conduct(residual, accepted)); }
int lare_dev_sur_plain_residual(double *residual) { return FWD(api_sur_plain_residual(residual)); }
int lare_dev_sur_stats(int32_t *accepted, int32_t *rejected, double *last_residual) {
  return FWD(api_sur_stats(accepted, rejected, last_residual));
}

}  // extern "C"
