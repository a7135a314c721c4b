// surrogate_rt.hpp -- runtime and C-ABI of the learned conduction surrogate
// (kernels in surrogate.hpp).  Included once, at the end of lare_dev.hip.
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
#pragma once

namespace {

struct Sur {
  float *w = nullptr, *gr = nullptr, *m = nullptr, *v = nullptr;  // SUR_NPARAM each
  unsigned short *w1t = nullptr, *w2t = nullptr;                   // packed BF16 for inference
  bool have_params = false, have_norm = false;
  SurNorm nm;
  int adam_t = 0;
  float *feat = nullptr;      // ntot * SUR_CH
  double *E1 = nullptr, *trap = nullptr, *E0save = nullptr;
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

// One surrogate super-step of length S.dt on S.energy (stages and Larsen
// factor from calc_s_stages), checked and, if rejected, redone by RKL2.
int sur_conduct_step(int delta_mode, double *res_out, int *acc_out) {
  if (!SU.have_params || !SU.have_norm) return fail(LARE_E_STATE, "surrogate: no trained parameters loaded");
  int rc = calc_s_stages();
  if (rc) return rc;
  rc = sur_features(1);
  if (rc) return rc;
  const int gpr = (nx() + 15) / 16;
  const int ngroups = gpr * ny() * nz();
  const unsigned nblk = (unsigned)((ngroups + SUR_INFER_WAVES - 1) / SUR_INFER_WAVES);
  LAUNCH("sur_infer", k_sur_infer, dim3(nblk < 2048u ? nblk : 2048u), dim3(64 * SUR_INFER_WAVES), S.g, SU.feat,
         SU.w1t, SU.w2t, SU.w, S.energy, SU.E1, S.tfac, SU.nm.y_scale, ngroups);
  // trapezoid dt/2 (L(E0) + L(E1)): stage kernels with all history weights zero
  CondArgs ca = {S.p.kappa_0, S.tfac, S.dt};
  Box in = interior();
  LAUNCH_BOX("rkl2_stage1", k_rkl2_stage1, in, S.energy, S.rho, S.bx, S.by, S.bz, S.larsen, S.Lc, S.Yb[0], ca, 0.0);
  LAUNCH_BOX("rkl2_stage", k_rkl2_stage, in, SU.E1, SU.E1, S.energy, S.rho, S.bx, S.by, S.bz, S.larsen, S.Lc,
             SU.trap, ca, TMODE_WRAP, 0.0, 0.0, 0.0, 0.5, 0.5);
  dim3 gr = box_grid(interior());
  LAUNCH("sur_residual", k_sur_residual, gr, box_block(), S.g, S.energy, SU.E1, SU.trap, S.partial);
  rc = final_reduce(gr, 2, 0, 0, 0);
  if (rc) return rc;
  double v[2];
  rc = read_scalars(0, 2, v);
  if (rc) return rc;
  const double r = v[1] > 0.0 ? sqrt(v[0] / v[1]) : 0.0;
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

}  // namespace

extern "C" {

int lare_dev_sur_set_params(const float *params, int32_t n) {
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

int lare_dev_sur_get_params(float *params, int32_t n) {
  if (!SU.have_params) return fail(LARE_E_STATE, "surrogate: no parameters");
  if (!params || n != SUR_NPARAM) return fail(LARE_E_ARG, "surrogate parameters: need %d floats, got %d", SUR_NPARAM, n);
  HIPCHK(hipStreamSynchronize(S.stream));
  HIPCHK(hipMemcpy(params, SU.w, sizeof(float) * SUR_NPARAM, hipMemcpyDeviceToHost));
  return 0;
}

int lare_dev_sur_set_norm(const float *mu, const float *sd, float y_scale) {
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

int lare_dev_sur_get_norm(float *mu, float *sd, float *y_scale) {
  if (!SU.have_norm) return fail(LARE_E_STATE, "surrogate: no normalisation");
  for (int q = 0; q < SUR_NCH; ++q) {
    if (mu) mu[q] = SU.nm.mu[q];
    if (sd) sd[q] = SU.nm.sd[q];
  }
  if (y_scale) *y_scale = SU.nm.y_scale;
  return 0;
}

int lare_dev_sur_collect(int32_t count, uint64_t seed, int32_t *total) {
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
  LAUNCH("sur_sample", k_sur_sample_cells, dim3(blocks_for(count)), dim3(256), S.g, (unsigned long long)seed, count,
         cells);
  LAUNCH("sur_gather", k_sur_gather, dim3(blocks_for((size_t)count * SUR_K)), dim3(256), S.g, SU.feat, cells, count,
         SU.X + (size_t)SU.rows * SUR_K);
  HIPCHK(hipMemcpyAsync(SU.E0save, S.energy, sizeof(double) * S.ntot, hipMemcpyDeviceToDevice, S.stream));
  rc = heat_conduct_sts2(0);
  if (rc) return rc;
  LAUNCH("sur_target", k_sur_target, dim3(blocks_for(count)), dim3(256), SU.E0save, S.energy, cells, count, S.tfac,
         SU.y + SU.rows);
  HIPCHK(hipMemcpyAsync(S.energy, SU.E0save, sizeof(double) * S.ntot, hipMemcpyDeviceToDevice, S.stream));
  HIPCHK(hipStreamSynchronize(S.stream));
  SU.rows += count;
  if (total) *total = SU.rows;
  return 0;
}

int lare_dev_sur_rows(float *X, float *y, int32_t rows) {
  if (rows < 0 || rows > SU.rows) return fail(LARE_E_ARG, "surrogate rows: %d requested, %d held", rows, SU.rows);
  HIPCHK(hipStreamSynchronize(S.stream));
  if (X && rows) HIPCHK(hipMemcpy(X, SU.X, sizeof(float) * (size_t)rows * SUR_K, hipMemcpyDeviceToHost));
  if (y && rows) HIPCHK(hipMemcpy(y, SU.y, sizeof(float) * (size_t)rows, hipMemcpyDeviceToHost));
  return 0;
}

int lare_dev_sur_fit_norm(void) {
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
    const double sd = var > 1.0e-12 * (1.0 + mu * mu) ? sqrt(var) : 1.0;
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

int lare_dev_sur_train(int32_t steps, int32_t batch, float lr, uint64_t seed, float *loss_out) {
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

int lare_dev_sur_forward(const float *X, int32_t rows, float *y) {
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

int lare_dev_sur_grad(const float *X, const float *t, int32_t rows, float *grad, float *loss) {
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

int lare_dev_sur_enable(int32_t on, double tol) {
  if (!(tol > 0.0)) return fail(LARE_E_ARG, "surrogate tolerance must be positive");
  if (on && (!SU.have_params || !SU.have_norm)) return fail(LARE_E_STATE, "surrogate enable: no trained parameters");
  SU.enabled = on != 0;
  SU.tol = tol;
  return 0;
}

int lare_dev_sur_conduct(double *residual, int32_t *accepted) {
  REQUIRE_LOADED();
  if (S.p.kappa_0 <= 0.0) return fail(LARE_E_ARG, "surrogate conduct needs kappa_0 > 0");
  int acc = 0;
  int rc = sur_conduct_step(0, residual, &acc);
  if (accepted) *accepted = acc;
  return rc;
}

int lare_dev_sur_stats(int32_t *accepted, int32_t *rejected, double *last_residual) {
  if (accepted) *accepted = SU.accepted;
  if (rejected) *rejected = SU.rejected;
  if (last_residual) *last_residual = SU.last_res;
  return 0;
}

}  // extern "C"
