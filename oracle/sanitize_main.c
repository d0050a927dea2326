/* Test driver of the oracle under AddressSanitizer and UndefinedBehaviorSanitizer (test
   infrastructure, built by `make -C oracle sanitize`): every entry point the tests use, on a
   64-water box (all-pairs list) and a 500-water box (cell list), including loss blocks with
   truncation and snapshots, the MSD loss, the MLP term, the energy, the reweighting estimator and
   the barostat. Exits 0 when every call succeeds; the sanitizers abort on the first error. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/revsim.h"

typedef struct or_ctx or_ctx;
int or_create(const rs_config* cfg, or_ctx** out);
int or_destroy(or_ctx* o);
int or_set_theta(or_ctx* o, const double* th);
int or_set_target_rdf(or_ctx* o, const double* g, int32_t n);
int or_forces(or_ctx* o, const double* x, double* F, double* energy);
int or_energy(or_ctx* o, const double* x, double* U, double* dU, double* F);
int or_adjoint_terms(or_ctx* o, const double* x, const double* c, const double* w, double* F, double* h,
                     double* vjp);
int or_rdf(or_ctx* o, const double* x, double* g, int64_t* hist, double* loss, double* dLdx);
int or_value_and_grad(or_ctx* o, const double* x0, const double* v0, uint64_t seed, int32_t n_steps,
                      double* loss, double* dLdth);
int or_value_and_grad_stored(or_ctx* o, const double* x0, const double* v0, uint64_t seed, int32_t n_steps,
                             double* loss, double* dLdth);
int or_reweight(or_ctx* o, const double* x0, const double* v0, uint64_t seed, int32_t n_steps, double* loss,
                double* dLdth, double* x, double* v);
int or_forward(or_ctx* o, const double* x0, const double* v0, uint64_t seed, int32_t n_steps, double* x,
               double* v, double* loss);
int or_set_mlp(or_ctx* o, const float* w, int64_t n);
int or_mlp_size(void);
int or_loss_terms(const or_ctx* o, double* t);

static void base_config(rs_config* c, int n, double box, double rc) {
  memset(c, 0, sizeof(*c));
  c->abi_version = RS_ABI_VERSION;
  c->n_molecules = n;
  c->box[0] = c->box[1] = c->box[2] = box;
  c->cutoff = rc;
  c->skin = 0.06;
  c->nl_every = 10;
  c->precision = RS_PRECISION_FP64;
  c->dt = 0.001;
  c->gamma = 1.0;
  c->temperature = 300.0;
  c->eps_rf = 78.3;
  c->coulomb_k = 138.935457644;
  c->kB = 0.00831446262;
  c->mass_O = 15.99943;
  c->mass_H = 1.007947;
  c->theta0 = 1.82421813418;
  c->rdf_bins = 200;
  c->rdf_rmax = rc < 0.8 ? rc : 0.8;
  c->rdf_width = 0.0005;
  c->rdf_window = 0.04;
  c->w_rdf = 1.0;
  c->msd_target = 2.3e-9;
  c->clip_norm = INFINITY;
  c->dhvap_target = 44.12 + 2.8;
  c->dhvap_u_gas = 3.69;
}

/* waters on a cubic lattice, a fixed bent geometry turned by the lattice index; small velocities */
static void lattice(int n, double box, double* x, double* v) {
  int k = (int)ceil(cbrt((double)n));
  double sp = box / k;
  for (int m = 0; m < n; ++m) {
    int i = m % k, j = (m / k) % k, l = m / (k * k);
    double O[3] = {(i + 0.5) * sp, (j + 0.5) * sp, (l + 0.5) * sp};
    double a = 0.7 * m;
    double u[3] = {cos(a), sin(a), 0.0}, w[3] = {-sin(a) * 0.6, cos(a) * 0.6, 0.8};
    double t = 1.82421813418 / 2;
    for (int d = 0; d < 3; ++d) {
      x[9 * m + d] = O[d];
      x[9 * m + 3 + d] = O[d] + 0.09572 * (cos(t) * w[d] + sin(t) * u[d]);
      x[9 * m + 6 + d] = O[d] + 0.09572 * (cos(t) * w[d] - sin(t) * u[d]);
    }
    for (int d = 0; d < 9; ++d) v[9 * m + d] = 0.3 * sin(1.7 * (9 * m + d));
  }
}

#define CHECK(e)                                              \
  do {                                                        \
    int rc_ = (e);                                            \
    if (rc_) {                                                \
      fprintf(stderr, "%s failed: %d\n", #e, rc_);            \
      return 1;                                               \
    }                                                         \
  } while (0)

static int run(int n, double box, double rc, int mlp) {
  rs_config c;
  base_config(&c, n, box, rc);
  c.n_steps_loss = 10;
  c.n_steps_trunc = 6;
  c.n_steps_log = 10;
  c.msd_window_steps = mlp ? 0 : 10;
  c.msd_every = mlp ? 0 : 5;
  c.w_msd = mlp ? 0.0 : 1e-3;
  c.w_dhvap = mlp ? 0.0 : 0.5;
  c.use_mlp = mlp;
  or_ctx* o;
  CHECK(or_create(&c, &o));
  const double th[6] = {0.315, 0.636, 0.417, 462750.4, 0.09572, 836.8};
  CHECK(or_set_theta(o, th));
  const int N3 = 9 * n;
  double *x = malloc(8 * N3), *v = malloc(8 * N3), *F = malloc(8 * N3), *h = malloc(8 * N3);
  double *cc = malloc(8 * N3), *ww = malloc(8 * N3), *g = malloc(8 * 200);
  int64_t hist[200];
  lattice(n, box, x, v);
  for (int i = 0; i < N3; ++i) {
    cc[i] = 1e-3 * sin(0.3 * i);
    ww[i] = 1e-3 * cos(0.7 * i);
  }
  if (mlp) {
    int nw = or_mlp_size();
    float* w = malloc(4 * nw);
    for (int k = 0; k < nw; ++k) w[k] = 0.05f * (float)sin(0.37 * k);
    CHECK(or_set_mlp(o, w, nw));
    free(w);
  }
  double e, U, dU[6], vj[6], loss, dth[6], terms[4];
  CHECK(or_forces(o, x, F, &e));
  CHECK(or_adjoint_terms(o, x, cc, ww, F, h, vj));
  CHECK(or_rdf(o, x, g, hist, &loss, h));
  for (int b = 0; b < 200; ++b) g[b] *= 1.01;
  CHECK(or_set_target_rdf(o, g, 200));
  CHECK(or_value_and_grad(o, x, v, 13, 20, &loss, dth));
  CHECK(or_loss_terms(o, terms));
  if (!mlp) {
    CHECK(or_energy(o, x, &U, dU, F));
    CHECK(or_forward(o, x, v, 13, 20, F, h, &loss));
    rs_config c2;
    base_config(&c2, n, box, rc);
    c2.w_dhvap = 0.5;
    c2.n_steps_loss = 5;
    c2.barostat_every = 5;
    c2.barostat_pressure = 1.0;
    or_ctx* o2;
    CHECK(or_create(&c2, &o2));
    CHECK(or_set_theta(o2, th));
    CHECK(or_set_target_rdf(o2, g, 200));
    CHECK(or_reweight(o2, x, v, 13, 10, &loss, dth, F, h));
    or_destroy(o2);
    rs_config c3;
    base_config(&c3, n, box, rc);
    or_ctx* o3;
    CHECK(or_create(&c3, &o3));
    CHECK(or_set_theta(o3, th));
    CHECK(or_set_target_rdf(o3, g, 200));
    CHECK(or_value_and_grad_stored(o3, x, v, 13, 8, &loss, dth));
    or_destroy(o3);
  }
  or_destroy(o);
  free(x);
  free(v);
  free(F);
  free(h);
  free(cc);
  free(ww);
  free(g);
  return 0;
}

int main(void) {
  if (run(64, 1.3, 0.6, 0) || run(1000, 3.25, 0.8, 0) || run(64, 1.3, 0.6, 1)) return 1;
  printf("sanitized oracle run ok\n");
  return 0;
}
