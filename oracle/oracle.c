/*
 * oracle.c — FP64 CPU restatement of rev-sim's reversible-simulation gradient loop for flexible
 * 3-site water.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker for librevsim.so: only the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark load it (through oracle/oracle.py). The
 * product library never links, loads or calls it.
 *
 * What each function restates (greener-group/rev-sim):
 *   or_forces           calc_F!                      water/water.jl:877-975 (LJ force :160-170)
 *   or_adjoint_terms    calc_F_and_grads!            water/water.jl:1212-1433 (LJ grads :172-193),
 *                       weighted dF/dθ sum           water/water.jl:1517-1519
 *   or_noise            generate_noise! semantics    water/water.jl:867-873 (Philox4x64-10 stream)
 *   or_forward          forward_sim loop             water/water.jl:1029-1045 (e, s at :1019-1020)
 *   or_reverse          reverse_sim (accumulators)   water/water.jl:1435-1558; additive injection of
 *                       intermediate-frame gradients (diffusion/diffusion.jl:1534-1539 reset form)
 *   or_rdf              rdf_differentiable + rrule   water/water.jl:631-736 (binning per config)
 *   or_msd              calculate_D + loss_fn_final  diffusion/diffusion.jl:861-911
 *   or_energy           potential_energy + its θ-gradient water/water.jl:743,1183-1191 (reaction
 *                       field shifted to zero at r_c, truncated LJ, harmonic bond and angle)
 *   dhvap_block         enthalpy_vaporization + loss_enth_vap water/water.jl:738-754
 *   or_reweight         reweighting_sim (ensemble reweighting)  water/water.jl:1107-1210
 *   or_value_and_grad_stored
 *                       independent check: textbook (λx, λv) adjoint over the stored forward
 *                       trajectory — plays the role of the DMS gradient (water/water.jl:1052-1099)
 *
 * Parity pins (tests/): energy finite differences for forces, force finite differences for the
 * Hessian-vector product and the θ-VJP, numpy Philox fixtures for the noise bits, brute-force and
 * closed-form RDF/MSD cases, reversibility, and central finite differences of the whole
 * trajectory loss for dL/dθ.
 *
 * Pair membership (which pairs interact) is defined exactly so that the GPU library and this
 * oracle agree bit for bit:
 *   FP64 rule  dx = x_b - x_a; dx -= L*rint(dx*(1/L)); r2 = fma(dz,dz,fma(dy,dy,dx*dx)); r2 <= rc*rc
 *   FP32 rule  (cubic box) q = (uint32)llrint(x * 2^32/L) per dimension; dxi = (int32)(q_b - q_a);
 *              dxq = (float)dxi; r2q = fmaf(dzq,dzq,fmaf(dyq,dyq,dxq*dxq));
 *              r2q <= (float)(rc*rc*qscale*qscale), qscale = 2^32/L (lengths in units of L/2^32)
 * The RDF bins use FP64 distances in both precisions.
 * Compile with -ffp-contract=off (oracle/Makefile) so these expressions are evaluated as written.
 */
#include "../include/revsim.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NTH 6 /* theta length */

typedef struct {
  int32_t step;
  double* x;
  double* v;
} snap_t;

/* growable list of atom pairs (a, b): one per thread while a neighbour list is built */
typedef struct pairbuf {
  int32_t *a, *b;
  int64_t n, cap;
} pairbuf_t;

typedef struct or_ctx {
  rs_config cfg;
  int M, N, N3;
  double L[3], invL[3], qscale[3];
  float unitf[3];
  double theta[NTH];
  double krf, rc2;
  float rc2f;
  float rc2q; /* FP32 pair rule threshold in units of (L/2^32)^2 */
  double crf; /* reaction-field energy shift (1/r_c) 3 eps_rf / (2 eps_rf + 1) */
  int rule64; /* 1 while evaluating the energy: FP64 membership rule in both precisions */
  double* mass3; /* per degree of freedom */
  /* atom-pair half list (a<b, different molecules), FP64 mic distance <= rlist */
  double rlist;
  int32_t* pa;
  int32_t* pb;
  int64_t npairs, cap;
  int64_t* cnt;
  /* loss */
  double* target;
  /* external noise */
  double* ext;
  int ext_steps;
  /* forward records for the reverse pass */
  snap_t* snaps;
  int nsnaps, capsnaps;
  double* frames; /* msd frames: nf * 3M */
  int nframes;
  int32_t rec_steps;
  uint64_t rec_seed;
  /* radial-basis MLP on O-O pairs (BF16-valued weights held as doubles) */
  int use_mlp;
  double* mlp;      /* NW weights: W1[64][16], b1[64], W2[64][64], b2[64], w3[64], b3 */
  double* dW;       /* dL/dW of the last reverse pass */
  /* per-thread force / HVP buffers of eval_terms and the u = c/m scratch, allocated once per
     context (grown when the thread count grows) and zeroed by their own threads */
  double* tbuf;
  int tbuf_threads;
  pairbuf_t* qbuf; /* per-thread list-build buffers, kept between builds */
  int qbuf_threads;
  double* ubuf;
  int64_t* qbufF; /* per-thread fixed-point pair-force sums */
  int float_sums;  /* 1: pair forces summed in FP64 instead (finite-difference tests only) */
  double* fbond;  /* bonded forces of the last evaluation */
  /* loss terms of the last trajectory call: RDF, enthalpy of vaporization, MSD (averaged over
     blocks, weights applied) and D (nm^2/ps) */
  double terms[4];
  char err[256];
} or_ctx;

/* fixed-point unit of the pair-force sums: 2^-33 kJ/mol/nm */
#define FX_SCALE 8589934592.0
#define FX_INV (1.0 / 8589934592.0)

#define MLP_NF 16
#define MLP_NH 64
#define MLP_NW (MLP_NH * MLP_NF + MLP_NH + MLP_NH * MLP_NH + MLP_NH + MLP_NH + 1)
#define MLP_W1 0
#define MLP_B1 (MLP_NH * MLP_NF)
#define MLP_W2 (MLP_B1 + MLP_NH)
#define MLP_B2 (MLP_W2 + MLP_NH * MLP_NH)
#define MLP_W3 (MLP_B2 + MLP_NH)
#define MLP_B3 (MLP_W3 + MLP_NH)
/* oracle-internal copies after the packed weights: W1 and W2 transposed, so the layer loops run
   over output neurons innermost (vectorizable, each sum still taken in input order) */
#define MLP_W1T MLP_NW
#define MLP_W2T (MLP_W1T + MLP_NF * MLP_NH)
#define MLP_NWX (MLP_W2T + MLP_NH * MLP_NH)

static void set_err(or_ctx* o, const char* m) {
  if (o) snprintf(o->err, sizeof(o->err), "%s", m);
}

const char* or_last_error(const or_ctx* o) { return o ? o->err : "null context"; }

/* ---------------------------------------------------------------- geometry helpers */

static inline double mic1(double d, double L, double invL) { return d - L * rint(d * invL); }

static inline void mic3(const or_ctx* o, const double* xa, const double* xb, double dr[3]) {
  for (int k = 0; k < 3; ++k) dr[k] = mic1(xb[k] - xa[k], o->L[k], o->invL[k]);
}

static inline double r2_of(const double dr[3]) {
  return fma(dr[2], dr[2], fma(dr[1], dr[1], dr[0] * dr[0]));
}

static inline uint32_t quant1(double x, double s) { return (uint32_t)(int64_t)llrint(x * s); }

/* FP32 pair membership in quantized units (no scaling of the integer difference). */
static inline float r2q_quant(const or_ctx* o, const double* xa, const double* xb) {
  float d[3];
  for (int k = 0; k < 3; ++k) {
    uint32_t qa = quant1(xa[k], o->qscale[k]);
    uint32_t qb = quant1(xb[k], o->qscale[k]);
    d[k] = (float)(int32_t)(qb - qa);
  }
  return fmaf(d[2], d[2], fmaf(d[1], d[1], d[0] * d[0]));
}

static inline int pair_in(const or_ctx* o, const double* xa, const double* xb, double r2) {
  if (o->cfg.precision == RS_PRECISION_FP32 && !o->rule64) return r2q_quant(o, xa, xb) <= o->rc2q;
  return r2 <= o->rc2;
}

static inline double wrap1(double x, double L) { return x - floor(x / L) * L; }

/* ---------------------------------------------------------------- context */

void or_default_config(rs_config* c) {
  memset(c, 0, sizeof(*c));
  c->abi_version = RS_ABI_VERSION;
  c->n_molecules = 216;
  c->box[0] = c->box[1] = c->box[2] = 1.862;
  c->cutoff = 0.8;
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
  c->rdf_rmax = 0.8;
  c->rdf_width = 0.0005;
  c->rdf_window = 0.04;
  c->w_rdf = 1.0;
  c->clip_norm = INFINITY;
}

static void free_records(or_ctx* o) {
  for (int i = 0; i < o->nsnaps; ++i) {
    free(o->snaps[i].x);
    free(o->snaps[i].v);
  }
  o->nsnaps = 0;
  free(o->frames);
  o->frames = NULL;
  o->nframes = 0;
}

int or_destroy(or_ctx* o) {
  if (!o) return 0;
  free_records(o);
  free(o->mlp);
  free(o->dW);
  free(o->snaps);
  free(o->mass3);
  free(o->pa);
  free(o->pb);
  free(o->cnt);
  free(o->tbuf);
  free(o->ubuf);
  free(o->qbufF);
  free(o->fbond);
  for (int t = 0; t < o->qbuf_threads; ++t) {
    free(o->qbuf[t].a);
    free(o->qbuf[t].b);
  }
  free(o->qbuf);
  free(o->target);
  free(o->ext);
  free(o);
  return 0;
}

int or_create(const rs_config* cfg, or_ctx** out) {
  *out = NULL;
  if (!cfg || cfg->abi_version != RS_ABI_VERSION || cfg->n_molecules < 2) return RS_E_CONFIG;
  if (cfg->precision == RS_PRECISION_FP32 && !(cfg->box[0] == cfg->box[1] && cfg->box[1] == cfg->box[2]))
    return RS_E_CONFIG;
  or_ctx* o = (or_ctx*)calloc(1, sizeof(or_ctx));
  o->cfg = *cfg;
  o->M = cfg->n_molecules;
  o->N = 3 * o->M;
  o->N3 = 3 * o->N;
  for (int k = 0; k < 3; ++k) {
    o->L[k] = cfg->box[k];
    o->invL[k] = 1.0 / cfg->box[k];
    o->qscale[k] = 4294967296.0 / cfg->box[k];
    o->unitf[k] = (float)(cfg->box[k] / 4294967296.0);
    if (cfg->cutoff > 0.5 * cfg->box[k]) {
      free(o);
      return RS_E_CONFIG;
    }
  }
  double rc = cfg->cutoff;
  o->krf = (1.0 / (rc * rc * rc)) * ((cfg->eps_rf - 1.0) / (2.0 * cfg->eps_rf + 1.0));
  o->rc2 = rc * rc;
  o->rc2f = (float)(rc * rc);
  o->rc2q = (float)(rc * rc * o->qscale[0] * o->qscale[0]);
  o->crf = (1.0 / rc) * ((3.0 * cfg->eps_rf) / (2.0 * cfg->eps_rf + 1.0));
  o->mass3 = (double*)malloc(sizeof(double) * o->N3);
  for (int a = 0; a < o->N; ++a)
    for (int k = 0; k < 3; ++k) o->mass3[3 * a + k] = (a % 3 == 0) ? cfg->mass_O : cfg->mass_H;
  /* atom list radius as in the reference (dist_neighbors = cutoff + 0.2, water/water.jl:38-39) */
  o->rlist = rc + 0.2;
  o->theta[0] = 0.315;
  o->theta[1] = 0.636;
  o->theta[2] = 0.417;
  o->theta[3] = 462750.4;
  o->theta[4] = 0.09572;
  o->theta[5] = 836.8;
  o->target = (double*)calloc(cfg->rdf_bins > 0 ? cfg->rdf_bins : 1, sizeof(double));
  *out = o;
  return 0;
}

int or_set_theta(or_ctx* o, const double* th) {
  memcpy(o->theta, th, sizeof(double) * NTH);
  return 0;
}

int or_set_target_rdf(or_ctx* o, const double* g, int32_t n) {
  if (n != o->cfg.rdf_bins) return RS_E_CONFIG;
  memcpy(o->target, g, sizeof(double) * n);
  return 0;
}

int or_set_external_noise(or_ctx* o, const double* xi, int32_t n_steps) {
  free(o->ext);
  o->ext = NULL;
  o->ext_steps = 0;
  if (!xi) return 0;
  o->ext = (double*)malloc(sizeof(double) * (size_t)o->N3 * n_steps);
  memcpy(o->ext, xi, sizeof(double) * (size_t)o->N3 * n_steps);
  o->ext_steps = n_steps;
  return 0;
}

/* ---------------------------------------------------------------- neighbour list (atoms) */

/* Half list of atom pairs in different molecules with FP64 minimum-image distance <= rlist.
   Plays the role of Molly's CellListMap finder (water/water.jl:1025,1038). Built by molecule: an
   atom pair within rlist has an O-O distance within R = rlist + 2 h, h the largest O-H distance
   at this state (triangle inequality), so molecule pairs i<j with O-O <= R (cells of edge >= R/2,
   a 5^3 stencil; 3^3 or all pairs in small boxes) are expanded to their nine atom pairs, each
   kept if within rlist. One pass: each thread takes a contiguous range of i (static schedule) and
   appends to its own buffer; the buffers are joined in thread order. Order: by i, stencil order,
   j ascending within a cell, then atom order — deterministic for a given thread count. */
static void pairbuf_grow(pairbuf_t* q, int64_t need) {
  if (need <= q->cap) return;
  int64_t cap = q->cap ? 2 * q->cap : 65536;
  while (cap < need) cap *= 2;
  q->a = (int32_t*)realloc(q->a, sizeof(int32_t) * cap);
  q->b = (int32_t*)realloc(q->b, sizeof(int32_t) * cap);
  q->cap = cap;
}

static void build_list(or_ctx* o, const double* x) {
  const int M = o->M;
  double hmax = 0.0;
#pragma omp parallel for schedule(static) reduction(max : hmax)
  for (int m = 0; m < M; ++m)
    for (int a = 1; a <= 2; ++a) {
      double dr[3];
      mic3(o, x + 9 * (size_t)m, x + 9 * (size_t)m + 3 * a, dr);
      const double r = sqrt(r2_of(dr));
      if (r > hmax) hmax = r;
    }
  const double R = o->rlist + 2.0 * hmax * (1.0 + 1e-12);
  const double R2 = R * R, rl2 = o->rlist * o->rlist;
  int nc[3], all = 0, S = 2;
  for (int k = 0; k < 3; ++k)
    if ((int)floor(o->L[k] * 2.0 / R) < 5) S = 1;
  for (int k = 0; k < 3; ++k) {
    nc[k] = (int)floor(o->L[k] * S / R);
    if (nc[k] < 2 * S + 1) all = 1;
  }
  int* cell_of = NULL;
  int *cstart = NULL, *cidx = NULL;
  int ncell = 1;
  if (!all) {
    ncell = nc[0] * nc[1] * nc[2];
    cell_of = (int*)malloc(sizeof(int) * M);
    cstart = (int*)calloc(ncell + 1, sizeof(int));
    cidx = (int*)malloc(sizeof(int) * M);
#pragma omp parallel for schedule(static)
    for (int m = 0; m < M; ++m) {
      int c[3];
      for (int k = 0; k < 3; ++k) {
        double w = wrap1(x[9 * (size_t)m + k], o->L[k]);
        c[k] = (int)(w / o->L[k] * nc[k]);
        if (c[k] >= nc[k]) c[k] = nc[k] - 1;
        if (c[k] < 0) c[k] = 0;
      }
      cell_of[m] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
    }
    for (int m = 0; m < M; ++m) cstart[cell_of[m] + 1]++;
    for (int c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    int* fill = (int*)malloc(sizeof(int) * ncell);
    memcpy(fill, cstart, sizeof(int) * ncell);
    for (int m = 0; m < M; ++m) cidx[fill[cell_of[m]]++] = m; /* ascending within a cell */
    free(fill);
  }
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  if (o->qbuf_threads < nth) {
    o->qbuf = (pairbuf_t*)realloc(o->qbuf, sizeof(pairbuf_t) * nth);
    memset(o->qbuf + o->qbuf_threads, 0, sizeof(pairbuf_t) * (nth - o->qbuf_threads));
    o->qbuf_threads = nth;
  }
  pairbuf_t* qb = o->qbuf;
  for (int t = 0; t < nth; ++t) qb[t].n = 0;
#pragma omp parallel num_threads(nth)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    pairbuf_t* q = qb + t;
#pragma omp for schedule(static)
    for (int i = 0; i < M; ++i) {
      const double* xi = x + 9 * (size_t)i;
      const int ci = all ? 0 : cell_of[i];
      const int cx = ci % nc[0], cy = (ci / nc[0]) % nc[1], cz = ci / (nc[0] * nc[1]);
      const int span = all ? 0 : S;
      for (int dz = -span; dz <= span; ++dz)
        for (int dy = -span; dy <= span; ++dy)
          for (int dx = -span; dx <= span; ++dx) {
            int lo = i + 1, hi = M;
            if (!all) {
              const int ex = (cx + dx + nc[0]) % nc[0], ey = (cy + dy + nc[1]) % nc[1],
                        ez = (cz + dz + nc[2]) % nc[2];
              const int e = (ez * nc[1] + ey) * nc[0] + ex;
              lo = cstart[e];
              hi = cstart[e + 1];
            }
            for (int u = lo; u < hi; ++u) {
              const int j = all ? u : cidx[u];
              if (j <= i) continue;
              const double* xj = x + 9 * (size_t)j;
              double dr[3];
              mic3(o, xi, xj, dr);
              if (r2_of(dr) > R2) continue;
              pairbuf_grow(q, q->n + 9);
              int64_t k = q->n;
              for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                  mic3(o, xi + 3 * a, xj + 3 * b, dr);
                  q->a[k] = 3 * i + a;
                  q->b[k] = 3 * j + b;
                  k += r2_of(dr) <= rl2;
                }
              q->n = k;
            }
          }
    }
  }
  int64_t tot = 0;
  for (int t = 0; t < nth; ++t) tot += qb[t].n;
  if (tot > o->cap) {
    free(o->pa);
    free(o->pb);
    o->cap = tot + tot / 4 + 16;
    o->pa = (int32_t*)malloc(sizeof(int32_t) * o->cap);
    o->pb = (int32_t*)malloc(sizeof(int32_t) * o->cap);
  }
  /* join the thread buffers in thread order, each thread copying its own */
#pragma omp parallel num_threads(nth)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    int T = 1;
#ifdef _OPENMP
    T = omp_get_num_threads();
#endif
    for (int w = t; w < nth; w += T) {
      int64_t off = 0;
      for (int u = 0; u < w; ++u) off += qb[u].n;
      memcpy(o->pa + off, qb[w].a, sizeof(int32_t) * qb[w].n);
      memcpy(o->pb + off, qb[w].b, sizeof(int32_t) * qb[w].n);
    }
  }
  o->npairs = tot;
  free(cell_of);
  free(cstart);
  free(cidx);
}

/* Molecule pairs i<j with FP64 O-O minimum-image distance <= r (brute force, for list parity). */
int or_molecule_pairs(or_ctx* o, const double* x, double r, int32_t* pairs, int64_t capacity,
                      int64_t* n_pairs) {
  int64_t n = 0;
  double r2 = r * r;
  for (int i = 0; i < o->M; ++i)
    for (int j = i + 1; j < o->M; ++j) {
      double dr[3];
      mic3(o, x + 9 * i, x + 9 * j, dr);
      if (r2_of(dr) <= r2) {
        if (n < capacity) {
          pairs[2 * n] = i;
          pairs[2 * n + 1] = j;
        }
        ++n;
      }
    }
  *n_pairs = n;
  return n > capacity ? RS_E_OVERFLOW : 0;
}

/* ---------------------------------------------------------------- dual numbers (angle HVP) */

typedef struct {
  double v, d;
} dual;

static inline dual D(double v, double d) {
  dual r = {v, d};
  return r;
}
static inline dual dadd(dual a, dual b) { return D(a.v + b.v, a.d + b.d); }
static inline dual dsub(dual a, dual b) { return D(a.v - b.v, a.d - b.d); }
static inline dual dmul(dual a, dual b) { return D(a.v * b.v, a.d * b.v + a.v * b.d); }
static inline dual ddiv(dual a, dual b) {
  return D(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
static inline dual dsqrt(dual a) {
  double s = sqrt(a.v);
  return D(s, a.d / (2.0 * s));
}
static inline dual dacos(dual a) { return D(acos(a.v), -a.d / sqrt(1.0 - a.v * a.v)); }
static inline dual dneg(dual a) { return D(-a.v, -a.d); }

static inline void dcross(const dual a[3], const dual b[3], dual c[3]) {
  c[0] = dsub(dmul(a[1], b[2]), dmul(a[2], b[1]));
  c[1] = dsub(dmul(a[2], b[0]), dmul(a[0], b[2]));
  c[2] = dsub(dmul(a[0], b[1]), dmul(a[1], b[0]));
}
static inline dual ddot(const dual a[3], const dual b[3]) {
  return dadd(dadd(dmul(a[0], b[0]), dmul(a[1], b[1])), dmul(a[2], b[2]));
}

/* Harmonic angle force at (i, j=centre, k), restating water/water.jl:936-968 with dual numbers:
   the value parts are the forces, the tangent parts are the directional derivative of the forces
   along (ui, uj, uk), i.e. (dF/dx) u = (dF/dx)^T u since dF/dx is the negative Hessian.
   Also returns the θ-derivative weights (disp) and the unit directions. Returns 0 if skipped. */
static int angle_dual(const or_ctx* o, const double* xi, const double* xj, const double* xk,
                      const double* ui, const double* uj, const double* uk, double ka,
                      double theta0, double fi[3], double fj[3], double fk[3], double hi[3],
                      double hj[3], double hk[3], double* disp_out, double p1d_out[3],
                      double p3d_out[3], double* energy) {
  double dji[3], djk[3];
  mic3(o, xj, xi, dji);
  mic3(o, xj, xk, djk);
  dual a[3], b[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = D(dji[k], ui ? ui[k] - uj[k] : 0.0);
    b[k] = D(djk[k], uk ? uk[k] - uj[k] : 0.0);
  }
  dual cr[3];
  dcross(a, b, cr);
  if (cr[0].v == 0.0 && cr[1].v == 0.0 && cr[2].v == 0.0) return 0;
  dual r2ji = ddot(a, a), r2jk = ddot(b, b);
  dual rji = dsqrt(r2ji), rjk = dsqrt(r2jk);
  dual p1u[3], p3u[3], mb[3];
  dcross(a, cr, p1u);
  for (int k = 0; k < 3; ++k) mb[k] = dneg(b[k]);
  dcross(mb, cr, p3u);
  dual p1n = dsqrt(ddot(p1u, p1u)), p3n = dsqrt(ddot(p3u, p3u));
  dual p1d[3], p3d[3];
  for (int k = 0; k < 3; ++k) {
    p1d[k] = ddiv(ddiv(p1u[k], p1n), rji);
    p3d[k] = ddiv(ddiv(p3u[k], p3n), rjk);
  }
  dual dotp = ddot(a, b);
  dual cs = ddiv(dotp, dmul(rji, rjk));
  dual th = dacos(cs);
  dual disp = dsub(D(theta0, 0.0), th);
  dual F = dmul(D(ka, 0.0), disp);
  for (int k = 0; k < 3; ++k) {
    dual gi = dmul(F, p1d[k]);
    dual gk = dmul(F, p3d[k]);
    dual gj = dneg(dadd(gi, gk));
    fi[k] = gi.v;
    fk[k] = gk.v;
    fj[k] = gj.v;
    hi[k] = gi.d;
    hk[k] = gk.d;
    hj[k] = gj.d;
    p1d_out[k] = p1d[k].v;
    p3d_out[k] = p3d[k].v;
  }
  *disp_out = disp.v;
  if (energy) *energy = 0.5 * ka * disp.v * disp.v;
  return 1;
}

/* ---------------------------------------------------------------- radial-basis MLP (O-O pairs) */

/* Round a float to bfloat16 (round to nearest, ties to even) and return it as a double: the MLP's
   weights are BF16 values. */
static double bf16_round(float f) {
  uint32_t b;
  memcpy(&b, &f, 4);
  uint32_t lsb = (b >> 16) & 1u;
  b += 0x7FFFu + lsb;
  b &= 0xFFFF0000u;
  float r;
  memcpy(&r, &b, 4);
  return (double)r;
}

static void mlp_transpose(double* W) {
  for (int n = 0; n < MLP_NH; ++n)
    for (int f = 0; f < MLP_NF; ++f) W[MLP_W1T + f * MLP_NH + n] = W[MLP_W1 + n * MLP_NF + f];
  for (int m = 0; m < MLP_NH; ++m)
    for (int n = 0; n < MLP_NH; ++n) W[MLP_W2T + n * MLP_NH + m] = W[MLP_W2 + m * MLP_NH + n];
}

int or_set_mlp(or_ctx* o, const float* w, int64_t n) {
  if (n != MLP_NW) return RS_E_CONFIG;
  if (!o->mlp) o->mlp = (double*)malloc(sizeof(double) * MLP_NWX);
  for (int i = 0; i < MLP_NW; ++i) o->mlp[i] = bf16_round(w[i]);
  mlp_transpose(o->mlp);
  o->use_mlp = 1;
  return 0;
}

int or_mlp_size(void) { return MLP_NW; }

/* Unrounded weights (finite-difference tests of the weight gradient). */
int or_set_mlp_f64(or_ctx* o, const double* w, int64_t n) {
  if (n != MLP_NW) return RS_E_CONFIG;
  if (!o->mlp) o->mlp = (double*)malloc(sizeof(double) * MLP_NWX);
  memcpy(o->mlp, w, sizeof(double) * MLP_NW);
  mlp_transpose(o->mlp);
  o->use_mlp = 1;
  return 0;
}

/* Features φ_k(r) = exp(-(r-μ_k)^2/(2w^2)) fc(r), μ_k = 0.2 + k 0.8/15, w = 0.05 nm,
   fc(r) = (cos(π r/r_c) + 1)/2, with their first and second r-derivatives. */
static void rbf(double r, double rc, double p[MLP_NF], double p1[MLP_NF], double p2[MLP_NF]) {
  const double w = 0.05, iw2 = 1.0 / (w * w);
  const double a = M_PI / rc;
  double fc = 0.5 * (cos(a * r) + 1.0), fc1 = -0.5 * a * sin(a * r), fc2 = -0.5 * a * a * cos(a * r);
  for (int k = 0; k < MLP_NF; ++k) {
    double mu = 0.2 + k * (0.8 / 15.0);
    double d = r - mu;
    double g = exp(-0.5 * d * d * iw2);
    double g1 = -d * iw2 * g;
    double g2 = (d * d * iw2 - 1.0) * iw2 * g;
    p[k] = g * fc;
    p1[k] = g1 * fc + g * fc1;
    p2[k] = g2 * fc + 2.0 * g1 * fc1 + g * fc2;
  }
}

/* SiLU σ(z) = z s(z) and its first three derivatives. */
static void silu4(double z, double* f0, double* f1, double* f2, double* f3) {
  double s = 1.0 / (1.0 + exp(-z));
  double s1 = s * (1.0 - s), s2 = s1 * (1.0 - 2.0 * s), s3 = s2 * (1.0 - 2.0 * s) - 2.0 * s1 * s1;
  *f0 = z * s;
  *f1 = s + z * s1;
  *f2 = 2.0 * s1 + z * s2;
  *f3 = 3.0 * s2 + z * s3;
}

typedef struct {
  double p[MLP_NF], p1[MLP_NF], p2[MLP_NF];
  double z1[MLP_NH], z1a[MLP_NH], z1b[MLP_NH], s1[4][MLP_NH];
  double a1[MLP_NH], a1a[MLP_NH], a1b[MLP_NH];
  double z2[MLP_NH], z2a[MLP_NH], z2b[MLP_NH], s2[4][MLP_NH];
  double a2[MLP_NH], a2a[MLP_NH], a2b[MLP_NH];
  double u, u1, u2;
} mlp_jet;

/* u(r), u'(r), u''(r) by forward-mode (second-order jets) through 16 -> 64 -> 64 -> 1 SiLU. */
static void mlp_forward(const double* W, double r, double rc, mlp_jet* J) {
  rbf(r, rc, J->p, J->p1, J->p2);
  for (int n = 0; n < MLP_NH; ++n) {
    J->z1[n] = W[MLP_B1 + n];
    J->z1a[n] = J->z1b[n] = 0.0;
  }
  for (int f = 0; f < MLP_NF; ++f) {
    const double* wt = W + MLP_W1T + f * MLP_NH;
    const double p0 = J->p[f], p1 = J->p1[f], p2 = J->p2[f];
    for (int n = 0; n < MLP_NH; ++n) {
      J->z1[n] += wt[n] * p0;
      J->z1a[n] += wt[n] * p1;
      J->z1b[n] += wt[n] * p2;
    }
  }
  for (int n = 0; n < MLP_NH; ++n) {
    const double z = J->z1[n], za = J->z1a[n], zb = J->z1b[n];
    silu4(z, &J->s1[0][n], &J->s1[1][n], &J->s1[2][n], &J->s1[3][n]);
    J->a1[n] = J->s1[0][n];
    J->a1a[n] = J->s1[1][n] * za;
    J->a1b[n] = J->s1[2][n] * za * za + J->s1[1][n] * zb;
  }
  J->u = W[MLP_B3];
  J->u1 = 0.0;
  J->u2 = 0.0;
  for (int m = 0; m < MLP_NH; ++m) {
    J->z2[m] = W[MLP_B2 + m];
    J->z2a[m] = J->z2b[m] = 0.0;
  }
  for (int n = 0; n < MLP_NH; ++n) {
    const double* wt = W + MLP_W2T + n * MLP_NH;
    const double a0 = J->a1[n], aa = J->a1a[n], ab = J->a1b[n];
    for (int m = 0; m < MLP_NH; ++m) {
      J->z2[m] += wt[m] * a0;
      J->z2a[m] += wt[m] * aa;
      J->z2b[m] += wt[m] * ab;
    }
  }
  for (int m = 0; m < MLP_NH; ++m) {
    const double z = J->z2[m], za = J->z2a[m], zb = J->z2b[m];
    silu4(z, &J->s2[0][m], &J->s2[1][m], &J->s2[2][m], &J->s2[3][m]);
    J->a2[m] = J->s2[0][m];
    J->a2a[m] = J->s2[1][m] * za;
    J->a2b[m] = J->s2[2][m] * za * za + J->s2[1][m] * zb;
    J->u += W[MLP_W3 + m] * J->a2[m];
    J->u1 += W[MLP_W3 + m] * J->a2a[m];
    J->u2 += W[MLP_W3 + m] * J->a2b[m];
  }
}

/* dW += g1 ∂u'/∂W + g2 ∂u''/∂W (reverse mode through the jets of mlp_forward). */
static void mlp_backward(const double* W, const mlp_jet* J, double g1, double g2, double* dW) {
  double zb2[MLP_NH], zb2a[MLP_NH], zb2b[MLP_NH];
  for (int m = 0; m < MLP_NH; ++m) {
    double w3 = W[MLP_W3 + m];
    dW[MLP_W3 + m] += g1 * J->a2a[m] + g2 * J->a2b[m];
    double ab = g1 * w3, abb = g2 * w3; /* ā2', ā2'' */
    double sp1 = J->s2[1][m], sp2 = J->s2[2][m], sp3 = J->s2[3][m];
    double za = J->z2a[m], zbv = J->z2b[m];
    zb2b[m] = abb * sp1;
    zb2a[m] = abb * sp2 * 2.0 * za + ab * sp1;
    zb2[m] = abb * (sp3 * za * za + sp2 * zbv) + ab * sp2 * za;
  }
  double ab1[MLP_NH], ab1a[MLP_NH], ab1b[MLP_NH];
  for (int n = 0; n < MLP_NH; ++n) ab1[n] = ab1a[n] = ab1b[n] = 0.0;
  for (int m = 0; m < MLP_NH; ++m) {
    dW[MLP_B2 + m] += zb2[m];
    for (int n = 0; n < MLP_NH; ++n) {
      double wv = W[MLP_W2 + m * MLP_NH + n];
      dW[MLP_W2 + m * MLP_NH + n] += zb2[m] * J->a1[n] + zb2a[m] * J->a1a[n] + zb2b[m] * J->a1b[n];
      ab1[n] += wv * zb2[m];
      ab1a[n] += wv * zb2a[m];
      ab1b[n] += wv * zb2b[m];
    }
  }
  for (int n = 0; n < MLP_NH; ++n) {
    double sp1 = J->s1[1][n], sp2 = J->s1[2][n], sp3 = J->s1[3][n];
    double za = J->z1a[n], zbv = J->z1b[n];
    double zb1b = ab1b[n] * sp1;
    double zb1a = ab1b[n] * sp2 * 2.0 * za + ab1a[n] * sp1;
    double zb1 = ab1b[n] * (sp3 * za * za + sp2 * zbv) + ab1a[n] * sp2 * za + ab1[n] * sp1;
    dW[MLP_B1 + n] += zb1;
    for (int f = 0; f < MLP_NF; ++f)
      dW[MLP_W1 + n * MLP_NF + f] += zb1 * J->p[f] + zb1a * J->p1[f] + zb1b * J->p2[f];
  }
}

/* u, u', u'' at r (for tests of the library's tabulation). */
int or_mlp_eval(or_ctx* o, const double* r, int32_t n, double* u, double* u1, double* u2) {
  if (!o->use_mlp) return RS_E_CONFIG;
  mlp_jet J;
  for (int i = 0; i < n; ++i) {
    mlp_forward(o->mlp, r[i], o->cfg.cutoff, &J);
    if (u) u[i] = J.u;
    if (u1) u1[i] = J.u1;
    if (u2) u2[i] = J.u2;
  }
  return 0;
}

/* dW of Σ_i (g1_i u'(r_i) + g2_i u''(r_i)) (for tests of the library's weight backprop). */
int or_mlp_vjp(or_ctx* o, const double* r, const double* g1, const double* g2, int32_t n, double* dW) {
  if (!o->use_mlp) return RS_E_CONFIG;
  memset(dW, 0, sizeof(double) * MLP_NW);
  mlp_jet J;
  for (int i = 0; i < n; ++i) {
    mlp_forward(o->mlp, r[i], o->cfg.cutoff, &J);
    mlp_backward(o->mlp, &J, g1[i], g2 ? g2[i] : 0.0, dW);
  }
  return 0;
}

/* ---------------------------------------------------------------- deterministic log, sincos, acos */

/* log, sin/cos of 2πu and acos from IEEE operations only (+, -, *, /, sqrt, fma, rint), the same
   fixed polynomials and evaluation order as the GPU library's (the file is built with
   -ffp-contract=off, so only the explicit fma calls fuse): the Box-Muller noise and the angle force
   then agree bit for bit, and so does the FP64 trajectory.
     log(x): x = m 2^k, m in (1/√2, √2], log m = 2 atanh((m-1)/(m+1)) to s^25, k ln2 (fdlibm split)
     sincos(2πu): 4u = q + r, q = rint(4u), θ = r π/2, Taylor to θ^19 / θ^20, quadrant q
     acos(x): π/2 - asin x (|x| <= 1/2), 2 asin √((1-x)/2), π - 2 asin √((1+x)/2); asin series to
              n = 25 */
static inline double rs_log(double x) {
  uint64_t bits;
  memcpy(&bits, &x, 8);
  int k = (int)((bits >> 52) & 0x7ffULL) - 1023;
  const uint64_t mb = (bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m;
  memcpy(&m, &mb, 8);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    k += 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double p = 0.04;
  p = fma(p, z, 0.043478260869565216);
  p = fma(p, z, 0.047619047619047616);
  p = fma(p, z, 0.05263157894736842);
  p = fma(p, z, 0.058823529411764705);
  p = fma(p, z, 0.06666666666666667);
  p = fma(p, z, 0.07692307692307693);
  p = fma(p, z, 0.09090909090909091);
  p = fma(p, z, 0.1111111111111111);
  p = fma(p, z, 0.14285714285714285);
  p = fma(p, z, 0.2);
  p = fma(p, z, 0.3333333333333333);
  p = fma(p, z, 1.0);
  const double lm = (2.0 * s) * p;
  const double kd = (double)k;
  return fma(kd, 6.93147180369123816490e-01, fma(kd, 1.90821492927058770002e-10, lm));
}

static inline void rs_sincos2pi(double u, double* sn, double* cs) {
  const double t = 4.0 * u;
  const double q = rint(t);
  const double th = (t - q) * 1.5707963267948966;
  const double z = th * th;
  double ps = -8.22063524662433e-18;
  ps = fma(ps, z, 2.8114572543455206e-15);
  ps = fma(ps, z, -7.647163731819816e-13);
  ps = fma(ps, z, 1.6059043836821613e-10);
  ps = fma(ps, z, -2.505210838544172e-08);
  ps = fma(ps, z, 2.7557319223985893e-06);
  ps = fma(ps, z, -0.0001984126984126984);
  ps = fma(ps, z, 0.008333333333333333);
  ps = fma(ps, z, -0.16666666666666666);
  ps = fma(ps, z, 1.0);
  double pc = 4.110317623312165e-19;
  pc = fma(pc, z, -1.5619206968586225e-16);
  pc = fma(pc, z, 4.779477332387385e-14);
  pc = fma(pc, z, -1.1470745597729725e-11);
  pc = fma(pc, z, 2.08767569878681e-09);
  pc = fma(pc, z, -2.755731922398589e-07);
  pc = fma(pc, z, 2.48015873015873e-05);
  pc = fma(pc, z, -0.001388888888888889);
  pc = fma(pc, z, 0.041666666666666664);
  pc = fma(pc, z, -0.5);
  pc = fma(pc, z, 1.0);
  const double sv = th * ps;
  switch (((int)q) & 3) {
    case 0: *sn = sv; *cs = pc; break;
    case 1: *sn = pc; *cs = -sv; break;
    case 2: *sn = -sv; *cs = -pc; break;
    default: *sn = -pc; *cs = sv; break;
  }
}

static inline double rs_asin_series(double y) {
  const double z = y * y;
  double p = 0.0022014739737101384;
  p = fma(p, z, 0.002338091892111975);
  p = fma(p, z, 0.0024894486782468836);
  p = fma(p, z, 0.00265787063820729);
  p = fma(p, z, 0.002846178401108942);
  p = fma(p, z, 0.0030578216492580306);
  p = fma(p, z, 0.003297059503473485);
  p = fma(p, z, 0.0035692053938259347);
  p = fma(p, z, 0.003880964558837669);
  p = fma(p, z, 0.004240907093679363);
  p = fma(p, z, 0.004660143486915096);
  p = fma(p, z, 0.005153309682319905);
  p = fma(p, z, 0.005740037670841924);
  p = fma(p, z, 0.006447210311889649);
  p = fma(p, z, 0.0073125258735988454);
  p = fma(p, z, 0.008390335809616815);
  p = fma(p, z, 0.009761609529194078);
  p = fma(p, z, 0.011551800896139705);
  p = fma(p, z, 0.01396484375);
  p = fma(p, z, 0.017352764423076924);
  p = fma(p, z, 0.022372159090909092);
  p = fma(p, z, 0.030381944444444444);
  p = fma(p, z, 0.044642857142857144);
  p = fma(p, z, 0.075);
  p = fma(p, z, 0.16666666666666666);
  p = fma(p, z, 1.0);
  return y * p;
}

static inline double rs_acos(double x) {
  const double pio2_hi = 1.57079632679489655800e+00, pio2_lo = 6.12323399573676603587e-17;
  if (x > 0.5) return 2.0 * rs_asin_series(sqrt(0.5 * (1.0 - x)));
  if (x < -0.5) return 2.0 * pio2_hi - (2.0 * rs_asin_series(sqrt(0.5 * (1.0 + x))) - 2.0 * pio2_lo);
  return pio2_hi - (rs_asin_series(x) - pio2_lo);
}

/* 1/√x (x > 0 normal) as the GPU library's FP64 pair terms compute it: the seed
   0x5FE6EB50C7B537A9 - (bits >> 1) (Lomont's / Robertson's 64-bit constant; relative error below
   3.5 %), then four Newton steps y ← y (3/2 - (x/2) y²); within 2 units in the last place. */
static inline double rs_rsqrt(double x) {
  int64_t b;
  memcpy(&b, &x, 8);
  b = 0x5FE6EB50C7B537A9LL - (b >> 1);
  double y;
  memcpy(&y, &b, 8);
  const double hx = 0.5 * x;
  for (int k = 0; k < 4; ++k) y = y * fma(-hx, y * y, 1.5);
  return y;
}


/* ---------------------------------------------------------------- bonded forces, library order */

static inline void cross3c(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double dot3c(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Bond and angle forces of one water (calc_F!, water/water.jl:915-968) in the GPU library's form
   and operation order, so that the FP64 forces agree bit for bit: n = d (1/r) for the bonds; for
   the angle (a = H1 - O, b = H2 - O, c = a x b) F_H1 = k_a (θ0 - θ) (a x c)/(|a|^2 |c|),
   F_H2 = k_a (θ0 - θ) (c x b)/(|b|^2 |c|), F_O = -(F_H1 + F_H2), with θ = acos(a.b / (|a||b|))
   from rs_acos. (The HVP and the θ-VJP, which do not steer the trajectory, come from the
   dual-number form in eval_terms.) */
static void bonded_force_mol(const or_ctx* o, const double* xm, double F[9]) {
  const double r0 = o->theta[4], kb = o->theta[3], ka = o->theta[5];
  for (int t = 0; t < 9; ++t) F[t] = 0.0;
  double a[3], b[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = mic1(xm[3 + k] - xm[k], o->L[k], o->invL[k]);
    b[k] = mic1(xm[6 + k] - xm[k], o->L[k], o->invL[k]);
  }
  for (int hh = 1; hh <= 2; ++hh) {
    const double* d = hh == 1 ? a : b;
    const double r = sqrt(r2_of(d));
    const double ir = 1.0 / r;
    const double n[3] = {d[0] * ir, d[1] * ir, d[2] * ir};
    const double disp = r - r0;
    const double Fv = -kb * disp;
    for (int k = 0; k < 3; ++k) {
      F[k] -= Fv * n[k];
      F[3 * hh + k] += Fv * n[k];
    }
  }
  double c[3];
  cross3c(a, b, c);
  if (c[0] == 0.0 && c[1] == 0.0 && c[2] == 0.0) return;
  const double aa = dot3c(a, a), bb = dot3c(b, b), ab = dot3c(a, b), cc = dot3c(c, c);
  const double P = aa * bb;
  const double iP = 1.0 / P;
  const double rab = sqrt(P);
  const double irab = rab * iP;
  const double rc = sqrt(cc), irc = 1.0 / rc;
  const double cs = ab * irab;
  const double th = rs_acos(cs);
  const double disp = o->cfg.theta0 - th;
  const double Fa = ka * disp;
  const double s1 = bb * iP * irc, s3 = aa * iP * irc;
  double p1u[3], p3u[3];
  cross3c(a, c, p1u);
  cross3c(c, b, p3u);
  const double c1 = Fa * s1, c3 = Fa * s3;
  for (int k = 0; k < 3; ++k) {
    const double gi = c1 * p1u[k], gk = c3 * p3u[k];
    F[3 + k] += gi;
    F[6 + k] += gk;
    F[k] -= gi + gk;
  }
}

/* ---------------------------------------------------------------- forces and adjoint terms */

/* One pass over all interactions. With c == NULL only forces (and energy) are produced
   (calc_F!, water/water.jl:877-975); otherwise the adjoint quantities of calc_F_and_grads!
   (water/water.jl:1212-1433): h = (dF/dx)^T (c/m) and vjp[k] = sum_dof w_dof dF_dof/dθ_k with
   θ = (σ_O, ε_O, q_H, k_b, r0, k_a). Per-thread buffers followed by a reduction, as in the
   reference (water/water.jl:879-885,970-972,1424-1430). */
static void eval_terms(or_ctx* o, const double* x, const double* c, const double* w, double* F,
                       double* h, double* vjp, double* energy, double* dW, double* dU) {
  const int N3 = o->N3;
  const double sig = o->theta[0], eps = o->theta[1], qH = o->theta[2];
  const double kb = o->theta[3], r0 = o->theta[4], ka = o->theta[5];
  const double qO = -2.0 * qH;
  const double kc = o->cfg.coulomb_k, krf = o->krf;
  const int adj = (c != NULL);
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  if (o->tbuf_threads < nth) {
    free(o->tbuf);
    free(o->qbufF);
    o->tbuf = (double*)malloc(sizeof(double) * 2 * (size_t)N3 * nth);
    o->qbufF = (int64_t*)malloc(sizeof(int64_t) * (size_t)N3 * nth);
    o->tbuf_threads = nth;
  }
  if (!o->ubuf) o->ubuf = (double*)malloc(sizeof(double) * N3);
  if (!o->fbond) o->fbond = (double*)malloc(sizeof(double) * N3);
  /* the pair-force constants exactly as the library forms them (per pair type O-O, O-H, H-H) */
  const double kqq[3] = {kc * qO * qO, kc * qO * qH, kc * qH * qH};
  const double kBq[3] = {kqq[0] * 2.0 * krf, kqq[1] * 2.0 * krf, kqq[2] * 2.0 * krf};
  const double sig2c = sig * sig, eps24c = 24.0 * eps;
  double* Fb = o->tbuf;
  double* hb = adj ? o->tbuf + (size_t)N3 * nth : NULL;
  double* vb = (double*)calloc((size_t)NTH * nth, sizeof(double));
  double* eb = (double*)calloc((size_t)nth, sizeof(double));
  double* db = (double*)calloc((size_t)NTH * nth, sizeof(double)); /* dU/dθ per thread */
  const int mlp_on = o->use_mlp;
  double* dWb = (adj && mlp_on && dW) ? (double*)calloc((size_t)MLP_NW * nth, sizeof(double)) : NULL;
  double* u = NULL;
  if (adj) {
    u = o->ubuf;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N3; ++i) u[i] = c[i] / o->mass3[i];
  }
  const int64_t np = o->npairs;
#pragma omp parallel num_threads(nth)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    double* Ft = Fb + (size_t)N3 * t;
    double* ht = adj ? hb + (size_t)N3 * t : NULL;
    int64_t* Fqt = o->qbufF + (size_t)N3 * t;
    {
      /* each thread zeroes its own buffer (and any of a team smaller than requested) */
      int T = 1;
#ifdef _OPENMP
      T = omp_get_num_threads();
#endif
      for (int w = t; w < nth; w += T) {
        memset(Fb + (size_t)N3 * w, 0, sizeof(double) * N3);
        memset(o->qbufF + (size_t)N3 * w, 0, sizeof(int64_t) * N3);
        if (adj) memset(hb + (size_t)N3 * w, 0, sizeof(double) * N3);
      }
    }
    double* vt = vb + NTH * t;
    double* dUt = db + NTH * t;
    double* dWt = dWb ? dWb + (size_t)MLP_NW * t : NULL;
    double et = 0.0;
    mlp_jet J;
#pragma omp for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
      int a = o->pa[p], b = o->pb[p];
      double dr[3];
      mic3(o, x + 3 * a, x + 3 * b, dr);
      double r2 = r2_of(dr);
      if (!pair_in(o, x + 3 * a, x + 3 * b, r2)) continue;
      int aO = (a % 3 == 0), bO = (b % 3 == 0);
      double qa = aO ? qO : qH, qb = bO ? qO : qH;
      double r = sqrt(r2);
      double invr2 = 1.0 / r2;
      double invr3 = invr2 / r;
      double ndr[3] = {dr[0] / r, dr[1] / r, dr[2] / r};
      double cp = kc * qa * qb;
      double Fp = -2.0 * cp * (invr3 + krf);
      double Frf = cp * (invr2 - 2.0 * krf * r);
      double Fv = Frf;
      double dFds = 0.0, dFde = 0.0;
      /* reaction-field pair energy, shifted to zero at r_c; -d/dr of it is Frf */
      const double erf_ = cp * (1.0 / r + krf * r2 - o->crf);
      if (energy) et += erf_;
      if (dU) dUt[2] += 2.0 * erf_ / qH; /* every charge product is q_H^2 times a constant */
      if (aO && bO && eps != 0.0) {
        double six = pow(sig * sig * invr2, 3);
        double twelve = six * six;
        double Flj = (24.0 * eps / r) * (2.0 * twelve - six);
        Fp += -24.0 * eps * invr2 * (26.0 * twelve - 7.0 * six);
        dFds = (144.0 * eps * invr2 / sig) * (4.0 * twelve - six); /* times dr */
        dFde = Flj / eps;                                            /* times ndr */
        Fv += Flj;
        if (energy) et += 4.0 * eps * (twelve - six);
        if (dU) {
          dUt[0] += (24.0 * eps / sig) * (2.0 * twelve - six);
          dUt[1] += 4.0 * (twelve - six);
        }
      }
      if (aO && bO && mlp_on) {
        /* MLP pair energy u(r): f = -u'(r), f' = -u''(r) */
        mlp_forward(o->mlp, r, o->cfg.cutoff, &J);
        Fv += -J.u1;
        Fp += -J.u2;
        if (energy) et += J.u;
        if (dWt) {
          double wd = 0.0;
          for (int k = 0; k < 3; ++k) wd += ndr[k] * (w[3 * b + k] - w[3 * a + k]);
          mlp_backward(o->mlp, &J, -wd, 0.0, dWt); /* dF/dW . w = -(du'/dW) r̂.(w_b - w_a) */
        }
      }
      {
        /* the force in the library's form and operation order: f/r from powers of 1/r,
           F_a -= (f/r) d, one term round((f/r) d_t 2^33) per component summed as 64-bit integers,
           so the sum is exact and independent of pair and thread order */
        const int ty = (!aO) + (!bO);
        const double ri = rs_rsqrt(r2), ri2 = ri * ri, ri3 = ri2 * ri;
        double ftot = fma(kqq[ty], ri3, -kBq[ty]);
        if (aO && bO) {
          const double s2 = sig2c * ri2, s6 = s2 * s2 * s2, s12 = s6 * s6;
          ftot = ftot + (eps24c * ri2) * fma(2.0, s12, -s6);
        }
        if (aO && bO && mlp_on) ftot = ftot + (-J.u1) * ri;
        if (o->float_sums) {
          for (int k = 0; k < 3; ++k) {
            Ft[3 * a + k] -= ftot * dr[k];
            Ft[3 * b + k] += ftot * dr[k];
          }
        } else {
          for (int k = 0; k < 3; ++k) {
            const int64_t q = (int64_t)rint((ftot * dr[k]) * FX_SCALE);
            Fqt[3 * a + k] += q;
            Fqt[3 * b + k] -= q;
          }
        }
      }
      if (adj) {
        /* K = F' r̂r̂ᵀ + (F/r)(I - r̂r̂ᵀ); h_a += K (u_a - u_b); h_b -= K (u_a - u_b) */
        double du[3], dot = 0.0;
        for (int k = 0; k < 3; ++k) {
          du[k] = u[3 * a + k] - u[3 * b + k];
          dot += ndr[k] * du[k];
        }
        double Fr = Fv / r;
        for (int k = 0; k < 3; ++k) {
          double hv = Fr * du[k] + (Fp - Fr) * dot * ndr[k];
          ht[3 * a + k] += hv;
          ht[3 * b + k] -= hv;
        }
        double wd_n = 0.0, wd_r = 0.0; /* ndr.(w_b - w_a), dr.(w_b - w_a) */
        for (int k = 0; k < 3; ++k) {
          double dw = w[3 * b + k] - w[3 * a + k];
          wd_n += ndr[k] * dw;
          wd_r += dr[k] * dw;
        }
        vt[0] += dFds * wd_r;
        vt[1] += dFde * wd_n;
        vt[2] += (2.0 * Frf / qH) * wd_n;
      }
    }
    /* bonded terms, per molecule: O-H1, O-H2 bonds (water/water.jl:915-934,1279-1305) and the
       H1-O-H2 angle (water/water.jl:936-968,1307-1422). */
#pragma omp for schedule(static)
    for (int m = 0; m < o->M; ++m) {
      int O = 3 * m;
      bonded_force_mol(o, x + 9 * (size_t)m, o->fbond + 9 * (size_t)m);
      for (int hh = 1; hh <= 2; ++hh) {
        int i = O, j = O + hh;
        double dr[3];
        mic3(o, x + 3 * i, x + 3 * j, dr);
        double r2 = r2_of(dr);
        double r = sqrt(r2);
        double ndr[3] = {dr[0] / r, dr[1] / r, dr[2] / r};
        double disp = r - r0;
        double Fv = -kb * disp;
        double Fp = -kb;
        if (energy) et += 0.5 * kb * disp * disp;
        if (dU) {
          dUt[3] += 0.5 * disp * disp;
          dUt[4] += -kb * disp;
        }
        if (adj) {
          double du[3], dot = 0.0;
          for (int k = 0; k < 3; ++k) {
            du[k] = u[3 * i + k] - u[3 * j + k];
            dot += ndr[k] * du[k];
          }
          double Fr = Fv / r;
          for (int k = 0; k < 3; ++k) {
            double hv = Fr * du[k] + (Fp - Fr) * dot * ndr[k];
            ht[3 * i + k] += hv;
            ht[3 * j + k] -= hv;
          }
          /* dF_i/dkb = disp*ndr, dF_j/dkb = -disp*ndr; dF_i/dr0 = -kb*ndr, dF_j/dr0 = kb*ndr */
          for (int k = 0; k < 3; ++k) {
            double dw = w[3 * i + k] - w[3 * j + k];
            vt[3] += disp * ndr[k] * dw;
            vt[4] += -kb * ndr[k] * dw;
          }
        }
      }
      double fi[3], fj[3], fk[3], hi[3], hj[3], hk[3], p1d[3], p3d[3], disp, ea = 0.0;
      int ok = angle_dual(o, x + 3 * (O + 1), x + 3 * O, x + 3 * (O + 2), adj ? u + 3 * (O + 1) : NULL,
                          adj ? u + 3 * O : NULL, adj ? u + 3 * (O + 2) : NULL, ka, o->cfg.theta0,
                          fi, fj, fk, hi, hj, hk, &disp, p1d, p3d, energy ? &ea : NULL);
      if (ok) {
        et += ea;
        if (dU) dUt[5] += 0.5 * disp * disp;
        for (int k = 0; k < 3; ++k) {
          if (adj) {
            ht[3 * (O + 1) + k] += hi[k];
            ht[3 * O + k] += hj[k];
            ht[3 * (O + 2) + k] += hk[k];
            vt[5] += disp * (p1d[k] * w[3 * (O + 1) + k] + p3d[k] * w[3 * (O + 2) + k] -
                             (p1d[k] + p3d[k]) * w[3 * O + k]);
          }
        }
      }
    }
    eb[t] = et;
  }
  /* forces: bonded + the exact integer pair sum; the HVP's thread buffers summed in thread order
     (deterministic for a given thread count) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N3; ++i) {
    int64_t q = 0;
    double fs = 0.0;
    for (int t = 0; t < nth; ++t) {
      q += o->qbufF[(size_t)N3 * t + i];
      fs += Fb[(size_t)N3 * t + i];
    }
    F[i] = o->fbond[i] + (o->float_sums ? fs : -((double)q * FX_INV));
    if (adj) {
      double sh = 0.0;
      for (int t = 0; t < nth; ++t) sh += hb[(size_t)N3 * t + i];
      h[i] = sh;
    }
  }
  if (vjp) {
    for (int k = 0; k < NTH; ++k) {
      double s = 0.0;
      for (int t = 0; t < nth; ++t) s += vb[NTH * t + k];
      vjp[k] = s;
    }
  }
  if (energy) {
    double s = 0.0;
    for (int t = 0; t < nth; ++t) s += eb[t];
    *energy = s;
  }
  if (dU) {
    for (int k = 0; k < NTH; ++k) {
      double s = 0.0;
      for (int t = 0; t < nth; ++t) s += db[NTH * t + k];
      dU[k] = s;
    }
  }
  if (dWb) {
    for (int k = 0; k < MLP_NW; ++k) {
      double s = 0.0;
      for (int t = 0; t < nth; ++t) s += dWb[(size_t)MLP_NW * t + k];
      dW[k] += s;
    }
    free(dWb);
  }
  free(vb);
  free(eb);
  free(db);
}

int or_forces(or_ctx* o, const double* x, double* F, double* energy) {
  build_list(o, x);
  eval_terms(o, x, NULL, NULL, F, NULL, NULL, energy, NULL, NULL);
  return 0;
}

/* Potential energy, its θ-gradient and the forces (potential_energy, water/water.jl:743;
   gradient(potential_energy, ...), water/water.jl:1183-1191). */
int or_energy(or_ctx* o, const double* x, double* U, double* dU, double* F) {
  double* Ft = F ? F : (double*)malloc(sizeof(double) * o->N3);
  double e = 0.0, d[NTH];
  build_list(o, x);
  o->rule64 = 1; /* the energy is FP64 throughout, its pair set the FP64 rule's */
  eval_terms(o, x, NULL, NULL, Ft, NULL, NULL, &e, NULL, d);
  o->rule64 = 0;
  if (U) *U = e;
  if (dU) memcpy(dU, d, sizeof(d));
  if (!F) free(Ft);
  return 0;
}

int or_adjoint_terms(or_ctx* o, const double* x, const double* c, const double* w, double* F,
                     double* h, double* vjp6) {
  build_list(o, x);
  eval_terms(o, x, c, w, F, h, vjp6, NULL, NULL, NULL);
  return 0;
}

/* As or_adjoint_terms, plus dW = sum_dof w_dof dF_dof/dW for the MLP weights. */
int or_adjoint_terms_mlp(or_ctx* o, const double* x, const double* c, const double* w, double* F,
                         double* h, double* vjp6, double* dW) {
  build_list(o, x);
  memset(dW, 0, sizeof(double) * MLP_NW);
  eval_terms(o, x, c, w, F, h, vjp6, NULL, dW, NULL);
  return 0;
}

/* ---------------------------------------------------------------- Philox noise */

#define PHILOX_M0 0xD2E7470EE14C6C93ULL
#define PHILOX_M1 0xCA5A826395121157ULL
#define PHILOX_W0 0x9E3779B97F4A7C15ULL
#define PHILOX_W1 0xBB67AE8584CAA73BULL

/* Philox4x64-10 (Salmon et al., SC'11), the generator numpy exposes as np.random.Philox. */
void or_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
  uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint64_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    __uint128_t p0 = (__uint128_t)PHILOX_M0 * c0;
    __uint128_t p1 = (__uint128_t)PHILOX_M1 * c2;
    uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
    uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += PHILOX_W0;
    k1 += PHILOX_W1;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* Standard normals of atom a at step: counter (step, a, 0, 0), key (seed, 0); Box-Muller on
   53-bit uniforms u1 in (0,1], u2 in [0,1). */
static void normals3(uint64_t seed, int32_t step, int a, double n[3]) {
  uint64_t ctr[4] = {(uint64_t)(uint32_t)step, (uint64_t)a, 0, 0}, key[2] = {seed, 0}, r[4];
  or_philox4x64(ctr, key, r);
  const double two53 = 1.0 / 9007199254740992.0;
  double u1 = (double)((r[0] >> 11) + 1) * two53, u2 = (double)(r[1] >> 11) * two53;
  double u3 = (double)((r[2] >> 11) + 1) * two53, u4 = (double)(r[3] >> 11) * two53;
  const double rad1 = sqrt(-2.0 * rs_log(u1)), rad2 = sqrt(-2.0 * rs_log(u3));
  double s2, c2, s4, c4;
  rs_sincos2pi(u2, &s2, &c2);
  rs_sincos2pi(u4, &s4, &c4);
  n[0] = rad1 * c2;
  n[1] = rad1 * s2;
  n[2] = rad2 * c4;
}

/* generate_noise! (water/water.jl:867-873): ξ_a = sqrt(kB T / m_a) · N(0, I3). */
int or_noise(or_ctx* o, uint64_t seed, int32_t step, double* xi) {
  if (o->ext) {
    if (step < 1 || step > o->ext_steps) return RS_E_CONFIG;
    memcpy(xi, o->ext + (size_t)(step - 1) * o->N3, sizeof(double) * o->N3);
    return 0;
  }
  double kT = o->cfg.kB * o->cfg.temperature;
#pragma omp parallel for schedule(static)
  for (int a = 0; a < o->N; ++a) {
    double n[3];
    normals3(seed, step, a, n);
    double s = sqrt(kT / o->mass3[3 * a]);
    for (int k = 0; k < 3; ++k) xi[3 * a + k] = s * n[k];
  }
  return 0;
}

/* ---------------------------------------------------------------- exp used by the RDF */

/* exp(t) for t <= 0 by Cody-Waite reduction t = k ln2 + f (fdlibm's split of ln2) and a Taylor
   polynomial in f evaluated with fused multiply-adds; exact scaling by 2^k. The GPU library
   evaluates the same sequence so that smoothed-histogram contributions agree bit for bit. */
static inline double rs_exp(double t) {
  double k = rint(t * 1.4426950408889634);
  double f = fma(-k, 6.93147180369123816490e-01, t);
  f = fma(-k, 1.90821492927058770002e-10, f);
  static const double inv_fact[14] = {1.0,
                                      1.0,
                                      1.0 / 2.0,
                                      1.0 / 6.0,
                                      1.0 / 24.0,
                                      1.0 / 120.0,
                                      1.0 / 720.0,
                                      1.0 / 5040.0,
                                      1.0 / 40320.0,
                                      1.0 / 362880.0,
                                      1.0 / 3628800.0,
                                      1.0 / 39916800.0,
                                      1.0 / 479001600.0,
                                      1.0 / 6227020800.0};
  double p = inv_fact[13];
  for (int i = 12; i >= 0; --i) p = fma(p, f, inv_fact[i]);
  return ldexp(p, (int)k);
}

double or_exp(double t) { return rs_exp(t); }

/* ---------------------------------------------------------------- RDF loss */

typedef struct {
  int lo, hi;
} win_t;

/* Smoothed-histogram contributions of one O-O distance (rdf_differentiable,
   water/water.jl:631-691, re-binned per the config: bin centres x_b = (b+1/2)Δ, Δ = rmax/bins,
   window |x_b - r| <= window, weights exp(-(x_b-r)^2/width) normalised over the window).
   Adds llrint(φ_b · 2^32) into hist. Returns the window. FP64 in both precisions: the narrow
   smoothing makes φ_b about 160/nm sensitive to r, so FP32 distances would put ~1e-5 relative
   noise into the histogram. */
static win_t rdf_contrib(const or_ctx* o, double r64, int64_t* hist) {
  const int nb = o->cfg.rdf_bins;
  win_t w = {0, -1};
  double invd = nb / o->cfg.rdf_rmax, c = o->cfg.rdf_window;
  double dx = o->cfg.rdf_rmax / nb, iw = 1.0 / o->cfg.rdf_width;
  int lo = (int)ceil((r64 - c) * invd - 0.5), hi = (int)floor((r64 + c) * invd - 0.5);
  if (lo < 0) lo = 0;
  if (hi > nb - 1) hi = nb - 1;
  w.lo = lo;
  w.hi = hi;
  if (lo > hi) return w;
  double e[512];
  double den = 0.0;
  for (int b = lo; b <= hi; ++b) {
    double t = ((double)b + 0.5) * dx - r64;
    e[b - lo] = rs_exp(-(t * t) * iw);
    den += e[b - lo];
  }
  for (int b = lo; b <= hi; ++b) hist[b] += (int64_t)llrint((e[b - lo] / den) * 4294967296.0);
  return w;
}

/* g_b, L = sum_b (g_b - target_b)^2 and dL/dx (water/water.jl:693-736 with a squared loss). The
   gradient differentiates φ_b(r) = e_b(r)/den(r) inside the fixed window, as the reference's
   Enzyme rrule does. */
int or_rdf(or_ctx* o, const double* x, double* g, int64_t* hist_out, double* loss, double* dLdx) {
  const int nb = o->cfg.rdf_bins, M = o->M;
  if (nb <= 0 || 2.0 * o->cfg.rdf_window * nb / o->cfg.rdf_rmax + 3.0 > 512.0) return RS_E_CONFIG;
  build_list(o, x);
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  int64_t* hb = (int64_t*)calloc((size_t)nb * nth, sizeof(int64_t));
  const int64_t np = o->npairs;
  const double rmaxw = o->cfg.rdf_rmax + o->cfg.rdf_window;
#pragma omp parallel
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    int64_t* ht = hb + (size_t)nb * t;
#pragma omp for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
      int a = o->pa[p], b = o->pb[p];
      if (a % 3 || b % 3) continue;
      double dr[3];
      mic3(o, x + 3 * a, x + 3 * b, dr);
      double r64 = sqrt(r2_of(dr));
      if (r64 > rmaxw + 0.01) continue;
      rdf_contrib(o, r64, ht);
    }
  }
  int64_t* H = (int64_t*)calloc(nb, sizeof(int64_t));
  for (int b = 0; b < nb; ++b)
    for (int t = 0; t < nth; ++t) H[b] += hb[(size_t)nb * t + b];
  free(hb);
  double V = o->L[0] * o->L[1] * o->L[2];
  double dx = o->cfg.rdf_rmax / nb;
  double pre = V / (4.0 * M_PI * (double)M * (double)M * dx);
  double* G = (double*)malloc(sizeof(double) * nb);
  double L = 0.0;
  for (int b = 0; b < nb; ++b) {
    double xb = ((double)b + 0.5) * dx;
    double gb = 2.0 * pre / (xb * xb) * ((double)H[b] / 4294967296.0);
    if (g) g[b] = gb;
    double d = gb - o->target[b];
    L += d * d;
    G[b] = 2.0 * d * 2.0 * pre / (xb * xb);
  }
  if (hist_out) memcpy(hist_out, H, sizeof(int64_t) * nb);
  if (loss) *loss = L;
  if (dLdx) {
    memset(dLdx, 0, sizeof(double) * o->N3);
    double iw = 1.0 / o->cfg.rdf_width;
    for (int64_t p = 0; p < np; ++p) {
      int a = o->pa[p], b = o->pb[p];
      if (a % 3 || b % 3) continue;
      double dr[3];
      mic3(o, x + 3 * a, x + 3 * b, dr);
      double r = sqrt(r2_of(dr));
      if (r > rmaxw + 0.01) continue;
      win_t w;
      /* window exactly as used for the histogram */
      {
        double invd = nb / o->cfg.rdf_rmax, c = o->cfg.rdf_window;
        w.lo = (int)ceil((r - c) * invd - 0.5);
        w.hi = (int)floor((r + c) * invd - 0.5);
      }
      if (w.lo < 0) w.lo = 0;
      if (w.hi > nb - 1) w.hi = nb - 1;
      if (w.lo > w.hi) continue;
      double den = 0.0, dden = 0.0, s1 = 0.0, s2 = 0.0;
      for (int bb = w.lo; bb <= w.hi; ++bb) {
        double t = ((double)bb + 0.5) * dx - r;
        double e = rs_exp(-(t * t) * iw);
        double de = e * 2.0 * t * iw; /* d e / d r */
        den += e;
        dden += de;
        s1 += G[bb] * de;
        s2 += G[bb] * e;
      }
      double dLdr = s1 / den - dden * s2 / (den * den);
      for (int k = 0; k < 3; ++k) {
        double gk = dLdr * dr[k] / r;
        dLdx[3 * b + k] += gk;
        dLdx[3 * a + k] -= gk;
      }
    }
  }
  free(H);
  free(G);
  return 0;
}

/* ---------------------------------------------------------------- MSD loss */

/* calculate_D (diffusion/diffusion.jl:861-877) on the O atoms of frames f = 0..2K-1 spaced
   msd_every*dt apart: D = (1/6K) sum_s MSD(s, s+K) / (K msd_every dt), i.e. the Einstein relation
   at lag K frames averaged over K origins; unwrapped displacement by summed minimum-image
   increments; D in nm^2/ps;
   L = w_msd 1e18 (1e-6 D - target)^2 (diffusion/diffusion.jl:897-911). dL/dframes analytic:
   the increments telescope, so dΔ_s/dx_{s+K} = I and dΔ_s/dx_s = -I. */
int or_msd(or_ctx* o, const double* frames, int32_t nf, double* Dout, double* loss, double* dLdf) {
  const int M = o->M;
  if (nf < 2 || nf % 2) return RS_E_CONFIG;
  int K = nf / 2;
  double lag = K * o->cfg.msd_every * o->cfg.dt;
  double* del = (double*)malloc(sizeof(double) * (size_t)K * M * 3);
  double Dsum = 0.0;
  for (int s = 0; s < K; ++s) {
    /* Σ_i |Δ_i|^2 in the library's order: index order within chunks of 256 molecules, then the
       chunk sums in order */
    double msd = 0.0, part = 0.0;
    for (int i = 0; i < M; ++i) {
      double acc[3] = {0, 0, 0};
      for (int f = s; f < s + K; ++f) {
        double d[3];
        mic3(o, frames + (size_t)f * 3 * M + 3 * i, frames + (size_t)(f + 1) * 3 * M + 3 * i, d);
        for (int k = 0; k < 3; ++k) acc[k] += d[k];
      }
      for (int k = 0; k < 3; ++k) del[((size_t)s * M + i) * 3 + k] = acc[k];
      part += acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2];
      if ((i + 1) % 256 == 0 || i == M - 1) {
        msd += part;
        part = 0.0;
      }
    }
    msd /= M;
    Dsum += msd / lag;
  }
  double Dv = Dsum / (6.0 * K);
  double diff = 1e-6 * Dv - o->cfg.msd_target;
  double L = o->cfg.w_msd * 1e18 * diff * diff;
  if (Dout) *Dout = Dv;
  if (loss) *loss = L;
  if (dLdf) {
    double dLdD = o->cfg.w_msd * 1e18 * 2.0 * diff * 1e-6;
    double c = dLdD * 2.0 / (6.0 * K * lag * M);
    memset(dLdf, 0, sizeof(double) * (size_t)nf * 3 * M);
    for (int s = 0; s < K; ++s)
      for (int i = 0; i < M; ++i)
        for (int k = 0; k < 3; ++k) {
          double d = del[((size_t)s * M + i) * 3 + k];
          dLdf[(size_t)(s + K) * 3 * M + 3 * i + k] += c * d;
          dLdf[(size_t)s * 3 * M + 3 * i + k] -= c * d;
        }
  }
  free(del);
  return 0;
}

/* ---------------------------------------------------------------- trajectory loop */

typedef struct {
  int n_loss, n_blocks, n_trunc, n_log;
  int msd_on, nf, first_frame;
} plan_t;

static int make_plan(const or_ctx* o, int n_steps, plan_t* p) {
  const rs_config* c = &o->cfg;
  p->n_loss = c->n_steps_loss > 0 ? c->n_steps_loss : n_steps;
  if (n_steps % p->n_loss) return RS_E_CONFIG;
  p->n_blocks = n_steps / p->n_loss;
  p->n_trunc = c->n_steps_trunc > 0 ? c->n_steps_trunc : p->n_loss;
  if (p->n_trunc > p->n_loss) return RS_E_CONFIG;
  p->n_log = c->n_steps_log;
  p->msd_on = (c->msd_window_steps > 0 && c->msd_every > 0 && c->w_msd != 0.0);
  p->nf = 0;
  p->first_frame = 0;
  if (p->msd_on) {
    if (c->msd_window_steps % c->msd_every || c->msd_window_steps > n_steps) return RS_E_CONFIG;
    p->nf = c->msd_window_steps / c->msd_every;
    if (p->nf < 2 || p->nf % 2) return RS_E_CONFIG;
    p->first_frame = n_steps - c->msd_window_steps + c->msd_every;
  }
  return 0;
}

static int frame_index(const plan_t* p, int t, int every) {
  if (!p->msd_on || t < p->first_frame) return -1;
  if ((t - p->first_frame) % every) return -1;
  return (t - p->first_frame) / every;
}

static void wrap_all(const or_ctx* o, double* x) {
  for (int a = 0; a < o->N; ++a)
    for (int k = 0; k < 3; ++k) x[3 * a + k] = wrap1(x[3 * a + k], o->L[k]);
}

static void add_snap(or_ctx* o, int step, const double* x, const double* v) {
  if (o->nsnaps == o->capsnaps) {
    o->capsnaps = o->capsnaps ? 2 * o->capsnaps : 16;
    o->snaps = (snap_t*)realloc(o->snaps, sizeof(snap_t) * o->capsnaps);
  }
  snap_t* s = &o->snaps[o->nsnaps++];
  s->step = step;
  s->x = (double*)malloc(sizeof(double) * o->N3);
  s->v = (double*)malloc(sizeof(double) * o->N3);
  memcpy(s->x, x, sizeof(double) * o->N3);
  memcpy(s->v, v, sizeof(double) * o->N3);
}

static snap_t* find_snap(or_ctx* o, int step) {
  for (int i = 0; i < o->nsnaps; ++i)
    if (o->snaps[i].step == step) return &o->snaps[i];
  return NULL;
}

/* New box edges (the per-box constants follow). */
static void set_box(or_ctx* o, const double* box) {
  for (int k = 0; k < 3; ++k) {
    o->cfg.box[k] = box[k];
    o->L[k] = box[k];
    o->invL[k] = 1.0 / box[k];
    o->qscale[k] = 4294967296.0 / box[k];
    o->unitf[k] = (float)(box[k] / 4294967296.0);
  }
  o->rc2q = (float)(o->rc2 * o->qscale[0] * o->qscale[0]);
}

/* Uniforms of the barostat move at step n: Philox4x64-10, counter (n, "BARS", 0, 0), key
   (seed, "barostat"), 53-bit mantissas. */
static void barostat_uniforms(uint64_t seed, int n, double u[2]) {
  uint64_t ctr[4] = {(uint64_t)(uint32_t)n, 0x42415253ULL, 0, 0}, key[2] = {seed, 0x6261726f73746174ULL}, out[4];
  or_philox4x64(ctr, key, out);
  u[0] = (double)(out[0] >> 11) * (1.0 / 9007199254740992.0);
  u[1] = (double)(out[1] >> 11) * (1.0 / 9007199254740992.0);
}

typedef struct {
  double volume_scale;
  int n_attempted, n_accepted;
} baro_t;

/* One Monte Carlo barostat move (apply_coupling!, water/water.jl:811-865) with Molly's
   scale_coords! for molecules: each water's centre of mass (H at the image nearest its O) scaled
   by (V'/V)^(1/3), the molecule moved rigidly, atoms wrapped into the new box; Metropolis on
   dW = dU + P dV - M kT ln(V'/V), P in bar. */
static int barostat_move(or_ctx* o, baro_t* bs, uint64_t seed, int n, double* x) {
  const int N3 = o->N3, M = o->M;
  double u[2], E0, E1;
  barostat_uniforms(seed, n, u);
  or_energy(o, x, &E0, NULL, NULL);
  const double V = o->L[0] * o->L[1] * o->L[2];
  const double dV = bs->volume_scale * (2.0 * u[0] - 1.0);
  const double vs = (V + dV) / V, ls = cbrt(vs);
  double old_box[3] = {o->L[0], o->L[1], o->L[2]}, new_box[3];
  for (int k = 0; k < 3; ++k) {
    new_box[k] = old_box[k] * ls;
    if (o->cfg.cutoff > 0.5 * new_box[k]) return RS_E_CONFIG;
  }
  double* xb = (double*)malloc(sizeof(double) * N3);
  memcpy(xb, x, sizeof(double) * N3);
  const double mO = o->cfg.mass_O, mH = o->cfg.mass_H, mt = mO + 2.0 * mH;
  for (int m = 0; m < M; ++m) {
    double* xm = x + 9 * m;
    double a[9];
    for (int k = 0; k < 3; ++k) {
      a[k] = xm[k];
      a[3 + k] = xm[k] + mic1(xm[3 + k] - xm[k], o->L[k], o->invL[k]);
      a[6 + k] = xm[k] + mic1(xm[6 + k] - xm[k], o->L[k], o->invL[k]);
    }
    for (int k = 0; k < 3; ++k) {
      const double com = (mO * a[k] + mH * a[3 + k] + mH * a[6 + k]) / mt;
      const double sh = (ls - 1.0) * com;
      for (int t = 0; t < 3; ++t) xm[3 * t + k] = wrap1(a[3 * t + k] + sh, new_box[k]);
    }
  }
  set_box(o, new_box);
  or_energy(o, x, &E1, NULL, NULL);
  const double kT = o->cfg.kB * o->cfg.temperature;
  const double P = o->cfg.barostat_pressure * 0.0602214076; /* bar -> kJ mol^-1 nm^-3 */
  const double dW = (E1 - E0) + P * dV - M * kT * log(vs);
  if (dW <= 0.0 || u[1] < exp(-dW / kT)) {
    bs->n_accepted++;
  } else {
    memcpy(x, xb, sizeof(double) * N3);
    set_box(o, old_box);
  }
  free(xb);
  bs->n_attempted++;
  if (bs->n_attempted >= 10) {
    if (bs->n_accepted < 0.25 * bs->n_attempted) {
      bs->volume_scale /= 1.1;
      bs->n_attempted = bs->n_accepted = 0;
    } else if (bs->n_accepted > 0.75 * bs->n_attempted) {
      bs->volume_scale = fmin(bs->volume_scale * 1.1, V * 0.3);
      bs->n_attempted = bs->n_accepted = 0;
    }
  }
  build_list(o, x);
  return 0;
}

/* Enthalpy-of-vaporization block loss at x (enthalpy_vaporization / loss_enth_vap,
   water/water.jl:738-754): dH = u_gas - U/M + kB T, l = w (dH - target)^2. Returns l; *cp_out =
   c' = 2 w (dH - target)/M, the factor of dl/dx = c' F and dl/dθ = -c' dU/dθ; dU_out (nullable)
   receives dU/dθ. */
static double dhvap_block(or_ctx* o, const double* x, double* F, double* dU_out, double* cp_out) {
  double U, dU[NTH];
  or_energy(o, x, &U, dU, F);
  const double w = o->cfg.w_dhvap;
  const double dH = o->cfg.dhvap_u_gas - U / o->M + o->cfg.kB * o->cfg.temperature;
  const double d = dH - o->cfg.dhvap_target;
  if (cp_out) *cp_out = 2.0 * w * d / o->M;
  if (dU_out) memcpy(dU_out, dU, sizeof(dU));
  return w * d * d;
}

/* forward_sim loop (water/water.jl:1029-1045), run_grads=false, with the loss of each block and
   the MSD frames. Optionally stores every state (for the stored-trajectory adjoint). With rw
   (19 accumulators) it is the forward loop of reweighting_sim (water/water.jl:1157-1192): at each
   block end rw[0] += L_b, rw[1+k] += dL_b/dθ_k, rw[7+k] += dU/dθ_k, rw[13+k] += L_b dU/dθ_k. */
static int forward_impl(or_ctx* o, const double* x0, const double* v0, uint64_t seed, int n_steps,
                        double* x_out, double* v_out, double* loss_out, double* traj_x, double* rw) {
  plan_t p;
  int rc = make_plan(o, n_steps, &p);
  if (rc) return rc;
  if (o->ext && o->ext_steps < n_steps) return RS_E_CONFIG;
  const int N3 = o->N3, M = o->M;
  const double dt = o->cfg.dt, e = exp(-o->cfg.gamma * dt), s = sqrt(1.0 - e * e);
  double* x = (double*)malloc(sizeof(double) * N3);
  double* v = (double*)malloc(sizeof(double) * N3);
  double* F = (double*)malloc(sizeof(double) * N3);
  double* xi = (double*)malloc(sizeof(double) * N3);
  memcpy(x, x0, sizeof(double) * N3);
  memcpy(v, v0, sizeof(double) * N3);
  wrap_all(o, x);
  free_records(o);
  if (p.msd_on) {
    o->nframes = p.nf;
    o->frames = (double*)malloc(sizeof(double) * (size_t)p.nf * 3 * M);
  }
  add_snap(o, 0, x, v);
  if (traj_x) memcpy(traj_x, x, sizeof(double) * N3);
  build_list(o, x);
  const int baro = o->cfg.barostat_every > 0;
  if (baro && p.msd_on) {
    free(x);
    free(v);
    free(F);
    free(xi);
    return RS_E_CONFIG;
  }
  baro_t bs = {0.01 * o->L[0] * o->L[1] * o->L[2], 0, 0};
  double lrdf = 0.0, ldh = 0.0;
  double* Fe = (double*)malloc(sizeof(double) * N3);
  if (rw) memset(rw, 0, sizeof(double) * 19);
  for (int n = 1; n <= n_steps; ++n) {
    eval_terms(o, x, NULL, NULL, F, NULL, NULL, NULL, NULL, NULL);
    or_noise(o, seed, n, xi);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N3; ++i) {
      double m = o->mass3[i];
      double vi = v[i] + (F[i] * dt) / m;
      double xx = x[i] + (vi * dt) / 2.0;
      vi = vi * e + xi[i] * s;
      xx = xx + (vi * dt) / 2.0;
      x[i] = wrap1(xx, o->L[i % 3]);
      v[i] = vi;
    }
    if (n % o->cfg.nl_every == 0) build_list(o, x);
    if (traj_x) memcpy(traj_x + (size_t)n * N3, x, sizeof(double) * N3);
    if ((p.n_log > 0 && n % p.n_log == 0) || n % p.n_loss == 0) add_snap(o, n, x, v);
    double lb = 0.0;
    if (n % p.n_loss == 0 && o->cfg.w_rdf != 0.0) {
      double l;
      or_rdf(o, x, NULL, NULL, &l, NULL);
      lrdf += l;
      lb += o->cfg.w_rdf * l;
    }
    if (n % p.n_loss == 0 && (o->cfg.w_dhvap != 0.0 || rw)) {
      double dU[NTH], cp = 0.0, l = 0.0;
      if (o->cfg.w_dhvap != 0.0) {
        l = dhvap_block(o, x, Fe, dU, &cp);
        ldh += l;
        lb += l;
      } else {
        or_energy(o, x, NULL, dU, Fe);
      }
      if (rw) {
        rw[0] += lb;
        for (int k = 0; k < NTH; ++k) {
          rw[1 + k] += -cp * dU[k];
          rw[7 + k] += dU[k];
          rw[13 + k] += lb * dU[k];
        }
      }
    }
    int f = frame_index(&p, n, o->cfg.msd_every);
    if (f >= 0)
      for (int i = 0; i < M; ++i)
        for (int k = 0; k < 3; ++k) o->frames[(size_t)f * 3 * M + 3 * i + k] = x[9 * i + k];
    if (baro && n % o->cfg.barostat_every == 0) {
      int brc = barostat_move(o, &bs, seed, n, x);
      if (brc) {
        free(x);
        free(v);
        free(F);
        free(xi);
        free(Fe);
        return brc;
      }
    }
  }
  double loss = o->cfg.w_rdf * lrdf / p.n_blocks + ldh / p.n_blocks;
  free(Fe);
  o->terms[0] = o->cfg.w_rdf * lrdf / p.n_blocks;
  o->terms[1] = ldh / p.n_blocks;
  o->terms[2] = o->terms[3] = 0.0;
  if (p.msd_on) {
    double lm, Dm;
    or_msd(o, o->frames, p.nf, &Dm, &lm, NULL);
    loss += lm;
    o->terms[2] = lm;
    o->terms[3] = Dm;
  }
  o->rec_steps = baro ? 0 : n_steps; /* a trajectory that moved the box has no reverse pass */
  o->rec_seed = seed;
  if (x_out) memcpy(x_out, x, sizeof(double) * N3);
  if (v_out) memcpy(v_out, v, sizeof(double) * N3);
  if (loss_out) *loss_out = loss;
  free(x);
  free(v);
  free(F);
  free(xi);
  return 0;
}

int or_forward(or_ctx* o, const double* x0, const double* v0, uint64_t seed, int32_t n_steps,
               double* x_out, double* v_out, double* loss_out) {
  return forward_impl(o, x0, v0, seed, n_steps, x_out, v_out, loss_out, NULL, NULL);
}

/* Gradient seed at state t: RDF and enthalpy-of-vaporization block losses (weights w/n_blocks;
   the latter also adds its direct term -c' dU/dθ into dth) and the MSD frame gradient. The two
   block-loss terms are added to terms[0] and terms[1]. */
static double seed_grad(or_ctx* o, const plan_t* p, int t, const double* x, const double* dLdf,
                        double* g, int with_rdf, double* dth) {
  const int N3 = o->N3, M = o->M;
  memset(g, 0, sizeof(double) * N3);
  double l = 0.0;
  if (with_rdf && o->cfg.w_rdf != 0.0) {
    double* gr = (double*)malloc(sizeof(double) * N3);
    or_rdf(o, x, NULL, NULL, &l, gr);
    double wgt = o->cfg.w_rdf / p->n_blocks;
    for (int i = 0; i < N3; ++i) g[i] += wgt * gr[i];
    l *= wgt;
    o->terms[0] += l;
    free(gr);
  }
  if (with_rdf && o->cfg.w_dhvap != 0.0) {
    double* Fe = (double*)malloc(sizeof(double) * N3);
    double dU[NTH], cp;
    const double wgt = 1.0 / p->n_blocks;
    const double le = wgt * dhvap_block(o, x, Fe, dU, &cp);
    l += le;
    o->terms[1] += le;
    for (int i = 0; i < N3; ++i) g[i] += wgt * cp * Fe[i];
    for (int k = 0; k < NTH; ++k) dth[k] -= wgt * cp * dU[k];
    free(Fe);
  }
  int f = frame_index(p, t, o->cfg.msd_every);
  if (f >= 0 && dLdf)
    for (int i = 0; i < M; ++i)
      for (int k = 0; k < 3; ++k) g[9 * i + k] += dLdf[(size_t)f * 3 * M + 3 * i + k];
  return l;
}

/* reverse_sim (water/water.jl:1435-1558): reverse integration with regenerated noise and the
   A, B, C, D accumulator recurrence; block losses reset the accumulators, MSD frame gradients
   inside a block are added to them (same initial terms, by linearity of the recurrence). */
int or_reverse(or_ctx* o, const double* xN, const double* vN, uint64_t seed, int32_t n_steps,
               double* loss_out, double* dLdth, double* x0_out, double* v0_out) {
  plan_t p;
  int rc = make_plan(o, n_steps, &p);
  if (rc) return rc;
  if (o->rec_steps != n_steps || o->rec_seed != seed) return RS_E_STATE;
  const int N3 = o->N3, M = o->M;
  const double dt = o->cfg.dt, dt2 = dt * dt, e = exp(-o->cfg.gamma * dt), s = sqrt(1.0 - e * e);
  double *x = malloc(sizeof(double) * N3), *v = malloc(sizeof(double) * N3);
  double *A = malloc(sizeof(double) * N3), *B = malloc(sizeof(double) * N3);
  double *C = calloc(N3, sizeof(double)), *Dd = malloc(sizeof(double) * N3);
  double *h = calloc(N3, sizeof(double)), *F = malloc(sizeof(double) * N3);
  double *g = malloc(sizeof(double) * N3), *xi = malloc(sizeof(double) * N3);
  double* w = malloc(sizeof(double) * N3);
  double* dLdf = NULL;
  double loss = 0.0;
  memcpy(x, xN, sizeof(double) * N3);
  memcpy(v, vN, sizeof(double) * N3);
  o->terms[0] = o->terms[1] = o->terms[2] = o->terms[3] = 0.0;
  double lmsd = 0.0; /* added after the block losses, in the library's order */
  if (p.msd_on) {
    double Dm;
    dLdf = (double*)malloc(sizeof(double) * (size_t)p.nf * 3 * M);
    or_msd(o, o->frames, p.nf, &Dm, &lmsd, dLdf);
    o->terms[2] = lmsd;
    o->terms[3] = Dm;
  }
  double th[NTH] = {0};
  loss += seed_grad(o, &p, n_steps, x, dLdf, g, 1, th);
  for (int i = 0; i < N3; ++i) {
    A[i] = g[i] * (1.0 + e);
    B[i] = g[i] * (1.0 + e);
    Dd[i] = g[i] * (1.0 + 2.0 * e + e * e) / (2.0 * e);
  }
  if (o->use_mlp) {
    if (!o->dW) o->dW = (double*)malloc(sizeof(double) * MLP_NW);
    memset(o->dW, 0, sizeof(double) * MLP_NW);
  }
  build_list(o, x);
  int since_build = 0, trunc_count = 0;
  for (int n = n_steps; n >= 1; --n) {
    trunc_count++;
    if (trunc_count <= p.n_trunc) {
      or_noise(o, seed, n, xi);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < N3; ++i) {
        double xx = x[i] - (v[i] * dt) / 2.0;
        double vi = (v[i] - xi[i] * s) / e;
        xx = xx - (vi * dt) / 2.0;
        x[i] = wrap1(xx, o->L[i % 3]);
        v[i] = vi;
        Dd[i] = e * Dd[i] + ((dt2 * h[i]) * (1.0 + e)) / 2.0;
        C[i] = C[i] + Dd[i];
        w[i] = A[i] * (dt2 / (2.0 * o->mass3[i]));
      }
      if (++since_build >= o->cfg.nl_every) {
        build_list(o, x);
        since_build = 0;
      }
      double vj[NTH];
      eval_terms(o, x, C, w, F, h, vj, NULL, o->dW, NULL);
      if (o->cfg.clip_norm > 0.0 && isfinite(o->cfg.clip_norm)) {
        double nn = 0.0;
        for (int i = 0; i < N3; ++i) nn += h[i] * h[i];
        nn = sqrt(nn);
        if (nn > o->cfg.clip_norm)
          for (int i = 0; i < N3; ++i) h[i] = h[i] * o->cfg.clip_norm / nn;
      }
#pragma omp parallel for schedule(static)
      for (int i = 0; i < N3; ++i) {
        v[i] = v[i] - (F[i] * dt) / o->mass3[i];
        B[i] = e * B[i] + dt2 * h[i];
        A[i] = A[i] + B[i];
      }
      for (int k = 0; k < NTH; ++k) th[k] += vj[k];
    }
    int t = n - 1;
    if (t == 0) break;
    int restored = 0;
    if (p.n_log > 0 && t % p.n_log == 0 && (trunc_count <= p.n_trunc || t % p.n_loss == 0)) {
      snap_t* sn = find_snap(o, t);
      if (sn) {
        memcpy(x, sn->x, sizeof(double) * N3);
        memcpy(v, sn->v, sizeof(double) * N3);
        restored = 1;
      }
    }
    if (t % p.n_loss == 0) {
      if (!restored && trunc_count > p.n_trunc) {
        snap_t* sn = find_snap(o, t);
        memcpy(x, sn->x, sizeof(double) * N3);
        memcpy(v, sn->v, sizeof(double) * N3);
      }
      trunc_count = 0;
      loss += seed_grad(o, &p, t, x, dLdf, g, 1, th);
      build_list(o, x);
      since_build = 0;
      for (int i = 0; i < N3; ++i) {
        A[i] = g[i] * (1.0 + e);
        B[i] = g[i] * (1.0 + e);
        C[i] = 0.0;
        Dd[i] = g[i] * (1.0 + 2.0 * e + e * e) / (2.0 * e);
        h[i] = 0.0;
      }
    } else if (frame_index(&p, t, o->cfg.msd_every) >= 0) {
      seed_grad(o, &p, t, x, dLdf, g, 0, th);
      for (int i = 0; i < N3; ++i) {
        A[i] += g[i] * (1.0 + e);
        B[i] += g[i] * (1.0 + e);
        Dd[i] += g[i] * (1.0 + 2.0 * e + e * e) / (2.0 * e);
      }
    }
  }
  loss += lmsd;
  if (loss_out) *loss_out = loss;
  if (dLdth) memcpy(dLdth, th, sizeof(th));
  if (x0_out) memcpy(x0_out, x, sizeof(double) * N3);
  if (v0_out) memcpy(v0_out, v, sizeof(double) * N3);
  free(x);
  free(v);
  free(A);
  free(B);
  free(C);
  free(Dd);
  free(h);
  free(F);
  free(g);
  free(xi);
  free(w);
  free(dLdf);
  return 0;
}

int or_value_and_grad(or_ctx* o, const double* x0, const double* v0, uint64_t seed,
                      int32_t n_steps, double* loss_out, double* dLdth) {
  const int N3 = o->N3;
  double *x = malloc(sizeof(double) * N3), *v = malloc(sizeof(double) * N3);
  int rc = or_forward(o, x0, v0, seed, n_steps, x, v, NULL);
  if (!rc) rc = or_reverse(o, x, v, seed, n_steps, loss_out, dLdth, NULL, NULL);
  free(x);
  free(v);
  return rc;
}

/* Independent check of the reverse pass: the textbook adjoint of the forward map over the STORED
   forward trajectory (no time reversal). For step n: a = (dt/2)(1+e) λx_n + e λv_n,
   λx_{n-1} = λx_n + (dF/dx)(x_{n-1})^T (dt a/m) + g_{n-1}, λv_{n-1} = a,
   dL/dθ += (dt a/m) · dF/dθ(x_{n-1}). Full gradient: requires a single loss block (no reset). */
int or_value_and_grad_stored(or_ctx* o, const double* x0, const double* v0, uint64_t seed,
                             int32_t n_steps, double* loss_out, double* dLdth) {
  plan_t p;
  int rc = make_plan(o, n_steps, &p);
  if (rc) return rc;
  if (p.n_blocks != 1 || p.n_trunc != p.n_loss) return RS_E_CONFIG;
  const int N3 = o->N3, M = o->M;
  const double dt = o->cfg.dt, e = exp(-o->cfg.gamma * dt);
  double* traj = (double*)malloc(sizeof(double) * (size_t)(n_steps + 1) * N3);
  double loss;
  rc = forward_impl(o, x0, v0, seed, n_steps, NULL, NULL, &loss, traj, NULL);
  if (rc) {
    free(traj);
    return rc;
  }
  double *lx = malloc(sizeof(double) * N3), *lv = calloc(N3, sizeof(double));
  double *g = malloc(sizeof(double) * N3), *F = malloc(sizeof(double) * N3);
  double *hh = malloc(sizeof(double) * N3), *cc = malloc(sizeof(double) * N3);
  double* dLdf = NULL;
  if (p.msd_on) {
    dLdf = (double*)malloc(sizeof(double) * (size_t)p.nf * 3 * M);
    or_msd(o, o->frames, p.nf, NULL, NULL, dLdf);
  }
  double th[NTH] = {0};
  double terms_fwd[4];
  memcpy(terms_fwd, o->terms, sizeof(terms_fwd));
  seed_grad(o, &p, n_steps, traj + (size_t)n_steps * N3, dLdf, lx, 1, th);
  memcpy(o->terms, terms_fwd, sizeof(terms_fwd)); /* the forward pass's terms stand */
  if (o->use_mlp) {
    if (!o->dW) o->dW = (double*)malloc(sizeof(double) * MLP_NW);
    memset(o->dW, 0, sizeof(double) * MLP_NW);
  }
  for (int n = n_steps; n >= 1; --n) {
    const double* xp = traj + (size_t)(n - 1) * N3;
    for (int i = 0; i < N3; ++i) {
      double a = (dt / 2.0) * (1.0 + e) * lx[i] + e * lv[i];
      lv[i] = a;
      /* pass c with c/m = dt a/m and w = dt a/m: c = dt*a, w = dt*a/m */
      cc[i] = dt * a;
      hh[i] = dt * a / o->mass3[i];
    }
    double vj[NTH];
    build_list(o, xp);
    eval_terms(o, xp, cc, hh, F, g, vj, NULL, o->dW, NULL);
    for (int i = 0; i < N3; ++i) lx[i] += g[i];
    for (int k = 0; k < NTH; ++k) th[k] += vj[k];
    int t = n - 1;
    if (t > 0 && frame_index(&p, t, o->cfg.msd_every) >= 0) {
      seed_grad(o, &p, t, traj + (size_t)t * N3, dLdf, g, 0, th);
      for (int i = 0; i < N3; ++i) lx[i] += g[i];
    }
  }
  if (loss_out) *loss_out = loss;
  if (dLdth) memcpy(dLdth, th, sizeof(th));
  free(traj);
  free(lx);
  free(lv);
  free(g);
  free(F);
  free(hh);
  free(cc);
  free(dLdf);
  return 0;
}

/* reweighting_sim (water/water.jl:1107-1210): forward loop with per-block loss, direct loss
   gradient and dU/dθ; dL/dθ = <dL_b/dθ> - β(<L_b dU_b/dθ> - <L_b><dU_b/dθ>) (:1201-1207). */
int or_reweight(or_ctx* o, const double* x0, const double* v0, uint64_t seed, int32_t n_steps,
                double* loss_out, double* dLdth, double* x_out, double* v_out) {
  plan_t p;
  int rc = make_plan(o, n_steps, &p);
  if (rc) return rc;
  if (p.msd_on) return RS_E_CONFIG;
  double rw[19];
  rc = forward_impl(o, x0, v0, seed, n_steps, x_out, v_out, NULL, NULL, rw);
  if (rc) return rc;
  o->rec_steps = 0; /* no records for a reverse pass */
  const double B = p.n_blocks, beta = 1.0 / (o->cfg.kB * o->cfg.temperature);
  const double loss = rw[0] / B;
  if (loss_out) *loss_out = loss;
  if (dLdth)
    for (int k = 0; k < NTH; ++k) dLdth[k] = rw[1 + k] / B - beta * (rw[13 + k] / B - loss * rw[7 + k] / B);
  return 0;
}

double or_list_radius(const or_ctx* o) { return o->rlist; }

/* Pair forces summed in FP64 (1) or as the library's exact fixed-point integers (0, the
   default). Finite differences of forces with tiny steps resolve below the fixed-point unit
   (2^-33 kJ/mol/nm per term), so those tests use FP64 sums; a separate test bounds the difference
   between the two. */
int or_set_float_sums(or_ctx* o, int on) {
  o->float_sums = on != 0;
  return 0;
}

/* the deterministic elementary functions, for their accuracy tests */
double or_log(double x) { return rs_log(x); }
void or_sincos2pi(double u, double* sn, double* cs) { rs_sincos2pi(u, sn, cs); }
double or_acos(double x) { return rs_acos(x); }
double or_rsqrt(double x) { return rs_rsqrt(x); }

/* (l_rdf, l_ev, l_msd, D in m^2/s) of the last forward / reverse call, as rs_loss_terms. */
int or_loss_terms(const or_ctx* o, double* t) {
  t[0] = o->terms[0];
  t[1] = o->terms[1];
  t[2] = o->terms[2];
  t[3] = o->terms[3] * 1e-6;
  return 0;
}

int or_get_box(const or_ctx* o, double* box) {
  for (int k = 0; k < 3; ++k) box[k] = o->L[k];
  return 0;
}

/* dL/dW of the last reverse pass (or_reverse / or_value_and_grad / or_value_and_grad_stored). */
int or_mlp_grad(const or_ctx* o, double* dW) {
  if (!o->use_mlp || !o->dW) return RS_E_STATE;
  memcpy(dW, o->dW, sizeof(double) * MLP_NW);
  return 0;
}
int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
