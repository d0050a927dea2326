// api.hip — the C ABI of librevsim.so (include/revsim.h): context, buffers and the host-side
// orchestration of the forward trajectory and of the reverse-time adjoint.
//
// forward_sim  (water/water.jl:977-1105, run_grads=false) -> rs_forward
// reverse_sim  (water/water.jl:1435-1558)                 -> rs_reverse
// All device work of a call is queued on the context's stream; the host synchronises once at the
// end of the call to copy results back and check the device flags.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.h"
#include "mlp.h"

namespace {

thread_local std::string g_last_error;

struct Plan {
  int n_loss, n_blocks, n_trunc, n_log;
  bool msd_on;
  int nf, first_frame, every;
  int frame_of(int t) const {
    if (!msd_on || t < first_frame || (t - first_frame) % every) return -1;
    return (t - first_frame) / every;
  }
  bool snap_at(int t) const { return t == 0 || (n_log > 0 && t % n_log == 0) || t % n_loss == 0; }
};

}  // namespace

struct rs_ctx {
  rs_config cfg;
  int device = 0;
  hipStream_t s = nullptr;
  Launch L;
  double theta[6];
  int M = 0, N = 0, N3 = 0;
  double rlist = 0.0;
  int nslots = 0;
  // records of the last forward trajectory
  bool rec_valid = false;
  int rec_steps = 0;
  uint64_t rec_seed = 0;
  std::vector<int> snap_steps;
  std::vector<double*> snap_x, snap_v;
  int frames_cap = 0, msd_del_cap = 0;
  double msd_loss = 0.0, msd_D = 0.0;  // MSD term and D (nm^2/ps) of the last forward trajectory
  int ext_steps = 0;
  int since_build = 0;
  double hext_extra = 0.0;  // added to the O-H extent after an extent violation
  int cell_cap = 0;         // allocated length of the per-cell arrays (>= ncell + 1)
  int cell_refine = 0;      // cells refined below r_list/2 after a home cell overflowed a build block
  // list-driven reruns (capacity, stale list, O-H extent, cell overflow) and list builds of the
  // last trajectory or evaluation call, and reruns since the context was created
  int64_t reruns_call = 0, builds_call = 0, reruns_total = 0;
  // loss terms of the last trajectory call: RDF, enthalpy of vaporization, MSD, and D (nm^2/ps)
  double terms[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t baro_attempted = 0, baro_accepted = 0;  // barostat moves of the last forward trajectory
  bool mlp_set = false;     // weights uploaded and tabulated
  std::string err;
  // timing: events recorded around launches on the context's stream
  bool timing = false;
  double ms[RS_PHASE_COUNT] = {0};
  int64_t cnt[RS_PHASE_COUNT] = {0};
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<int> mark_phase;
  std::vector<double> slot_host;
};

namespace {

int fail(rs_ctx* c, int code, const std::string& msg) {
  g_last_error = msg;
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(c, expr)                                                                    \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return fail((c), RS_E_HIP, std::string(#expr ": ") + hipGetErrorString(e_));          \
  } while (0)

const double H_EXTENT_FACTOR = 1.15;  // h_ext = 1.15 r0 + 0.002 nm (checked every step)

void set_params(rs_ctx* c) {
  SysParams& p = c->L.p;
  const rs_config& g = c->cfg;
  p.M = c->M;
  p.N = c->N;
  // list entries: 16-bit sorted indices wherever they fit (half the pair kernels' list stream and
  // the build's stores, at unchanged kernel times), else 32-bit
  p.list16 = c->M <= 65536 ? 1 : 0;
  for (int k = 0; k < 3; ++k) {
    p.L[k] = g.box[k];
    p.invL[k] = 1.0 / g.box[k];
    p.qscale[k] = 4294967296.0 / g.box[k];
    p.unitf[k] = (float)(g.box[k] / 4294967296.0);
  }
  const double rc = g.cutoff;
  p.rc2 = rc * rc;
  p.rc2f = (float)(rc * rc);
  p.rc2q = (float)(rc * rc * p.qscale[0] * p.qscale[0]);
  p.krf = (1.0 / (rc * rc * rc)) * ((g.eps_rf - 1.0) / (2.0 * g.eps_rf + 1.0));
  p.krff = (float)p.krf;
  p.crf = (1.0 / rc) * ((3.0 * g.eps_rf) / (2.0 * g.eps_rf + 1.0));
  p.kc = g.coulomb_k;
  p.sig = c->theta[0];
  p.eps = c->theta[1];
  p.qH = c->theta[2];
  p.qO = -2.0 * c->theta[2];
  p.kb = c->theta[3];
  p.r0 = c->theta[4];
  p.ka = c->theta[5];
  p.theta0 = g.theta0;
  p.kqq[0] = g.coulomb_k * p.qO * p.qO;
  p.kqq[1] = g.coulomb_k * p.qO * p.qH;
  p.kqq[2] = g.coulomb_k * p.qH * p.qH;
  for (int k = 0; k < 3; ++k) p.kqqf[k] = (float)p.kqq[k];
  p.sig2f = (float)(p.sig * p.sig);
  p.eps24f = (float)(24.0 * p.eps);
  p.mO = g.mass_O;
  p.mH = g.mass_H;
  p.hext = H_EXTENT_FACTOR * std::fabs(p.r0) + 0.002 + c->hext_extra;
  c->rlist = rc + 2.0 * p.hext + g.skin;
  p.rlist2 = c->rlist * c->rlist;
  p.skin_half2 = 0.25 * g.skin * g.skin;
  // shells: an O-O pair can come within r_c only if its distance at the build was <= r_c + skin
  // (each O moves less than skin/2), an O-H pair only if <= r_c + h_ext + skin
  const double shr[2] = {rc + g.skin, rc + p.hext + g.skin};
  for (int k = 0; k < 2; ++k) {
    p.sh2[k] = shr[k] * shr[k] * (1.0 + 1e-5);
    p.sh2q[k] = (float)(p.sh2[k] * p.qscale[0] * p.qscale[0]);
  }
  // The FP32 list screen errs by < 1e-6 relative in r^2 (quantized positions, two roundings per
  // component), so a 1e-5 band around r_list^2 leaves every decision to the FP64 rule's outcome.
  p.rlo2f = (float)(p.rlist2 * (1.0 - 1e-5));
  p.rhi2f = (float)(p.rlist2 * (1.0 + 1e-5));
  // quantized units share one length scale only in a cubic box; otherwise every candidate goes
  // to the FP64 rule
  const bool cubic = g.box[0] == g.box[1] && g.box[1] == g.box[2];
  p.rlo2q = cubic ? (float)(p.rlist2 * p.qscale[0] * p.qscale[0] * (1.0 - 1e-5)) : -1.0f;
  p.rhi2q = cubic ? (float)(p.rlist2 * p.qscale[0] * p.qscale[0] * (1.0 + 1e-5)) : INFINITY;
  // cells of at least r_list/2 with a 5^3 stencil where the box allows it, else r_list and 3^3,
  // else all pairs
  p.brute = 0;
  p.ncell = 1;
  for (p.cs = std::min(2 + c->cell_refine, 3); p.cs >= 1; --p.cs) {
    bool ok = true;
    for (int k = 0; k < 3; ++k) {
      p.nc[k] = (int)std::floor(g.box[k] * p.cs / c->rlist);
      if (p.nc[k] < 2 * p.cs + 1) ok = false;
    }
    if (ok) break;
  }
  if (p.cs < 1) {
    p.brute = 1;
    p.cs = 1;
    p.nc[0] = p.nc[1] = p.nc[2] = 1;
  } else {
    p.ncell = p.nc[0] * p.nc[1] * p.nc[2];
  }
}

int default_capc(const rs_ctx* c) {
  const SysParams& p = c->L.p;
  if (p.brute) return (p.M + p.CH - 1) / p.CH + 1;
  const double V = c->cfg.box[0] * c->cfg.box[1] * c->cfg.box[2];
  const double nexp = (p.M / V) * (4.0 / 3.0) * M_PI * c->rlist * c->rlist * c->rlist;
  return (int)std::ceil(1.3 * nexp / p.CH + 24.0);
}

template <class T>
int dalloc(rs_ctx* c, T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIPCHK(c, hipMalloc((void**)p, n * sizeof(T)));
  HIPCHK(c, hipMemsetAsync(*p, 0, n * sizeof(T), c->s));
  return 0;
}

int alloc_list(rs_ctx* c) {
  DevBufs& b = c->L.b;
  if (b.nbr) hipFree(b.nbr);
  if (b.nbr_tmp) hipFree(b.nbr_tmp);
  b.nbr = b.nbr_tmp = nullptr;
  const size_t lanes = (size_t)pair_blocks(c->L.p) * RS_BLOCK;
  // sized for 32-bit entries, the wider of the two widths a build may use
  int r = dalloc(c, (int**)&b.nbr, lanes * LIST_SHELLS * c->L.p.capc);
  return r ? r : dalloc(c, &b.nbr_tmp, lanes * c->L.p.capc);
}

int alloc_all(rs_ctx* c) {
  DevBufs& b = c->L.b;
  std::memset(&b, 0, sizeof(b));
  const size_t N3 = c->N3, M = c->M, N = c->N;
  const SysParams& p = c->L.p;
  const size_t lanes = (size_t)pair_blocks(p) * RS_BLOCK;
  int r = 0;
  double** st[] = {&b.x, &b.v, &b.F, &b.h, &b.A, &b.B, &b.C, &b.D, &b.grad};
  if ((r = dalloc(c, &b.vjb, 3 * M))) return r;
  for (auto q : st)
    if ((r = dalloc(c, q, N3))) return r;
  c->cell_cap = p.ncell + 1;
  if ((r = dalloc(c, &b.cell, M)) || (r = dalloc(c, &b.cell_count, p.ncell + 1)) ||
      (r = dalloc(c, &b.cell_start, p.ncell + 1)) || (r = dalloc(c, &b.cell_fill, p.ncell + 1)) ||
      (r = dalloc(c, &b.tmp, M)) || (r = dalloc(c, &b.order, M)) || (r = dalloc(c, &b.rank, M)))
    return r;
  if ((r = dalloc(c, &b.posq, 3 * N)) || (r = dalloc(c, &b.pos64, N)) || (r = dalloc(c, &b.posq_alt, 3 * N)) ||
      (r = dalloc(c, &b.pos64_alt, N)) || (r = dalloc(c, &b.pu, 5 * M)) ||
      (r = dalloc(c, &b.ud, N)) || (r = dalloc(c, &b.O64, M)) || (r = dalloc(c, &b.Oq, M)))
    return r;
  if ((r = dalloc(c, &b.ncnt, lanes)) || (r = dalloc(c, &b.flags, FLAG_COUNT)) ||
      (r = dalloc(c, &b.slots, (size_t)pair_blocks(p) * 8)) || (r = dalloc(c, &b.dth, 8)) ||
      (r = dalloc(c, &b.scal, 16)) || (r = dalloc(c, &b.wmax, 128)) || (r = dalloc(c, &b.hpart, 256)))
    return r;
  const int nb = std::max(1, c->cfg.rdf_bins);
  if ((r = dalloc(c, &b.hist, nb)) || (r = dalloc(c, &b.rdfG, nb)) || (r = dalloc(c, &b.target, nb)) ||
      (r = dalloc(c, &b.rdf_g, nb)))
    return r;
  if ((r = dalloc(c, &b.Fe, N3)) || (r = dalloc(c, &b.eslots, (size_t)pair_blocks(p) * 8)) ||
      (r = dalloc(c, &b.eout, 8)) || (r = dalloc(c, &b.rwacc, 32)))
    return r;
  if (c->cfg.barostat_every > 0 && (r = dalloc(c, &b.xbak, N3))) return r;
  if (c->cfg.use_mlp) {
    if ((r = dalloc(c, &b.mlp_w, MLP_NW)) || (r = dalloc(c, &b.mlp_wb, MLP_NW)) ||
        (r = dalloc(c, (double2**)&b.mlp_tab, MLP_NT)) ||
        (r = dalloc(c, &b.mlp_slot, (size_t)pair_blocks(p) * MLP_NG)) || (r = dalloc(c, &b.mlp_G, MLP_NG)) ||
        (r = dalloc(c, &b.mlp_part, (size_t)(MLP_NT / 64) * MLP_NW)) || (r = dalloc(c, &b.mlp_dW, MLP_NW)))
      return r;
  }
  return alloc_list(c);
}

// The per-cell arrays follow the cell grid, which a box change can enlarge.
int ensure_cells(rs_ctx* c) {
  const int need = c->L.p.ncell + 1;
  if (need <= c->cell_cap) return 0;
  DevBufs& b = c->L.b;
  hipFree(b.cell_count);
  hipFree(b.cell_start);
  hipFree(b.cell_fill);
  b.cell_count = b.cell_start = b.cell_fill = nullptr;
  int r;
  if ((r = dalloc(c, &b.cell_count, need)) || (r = dalloc(c, &b.cell_start, need)) ||
      (r = dalloc(c, &b.cell_fill, need)))
    return r;
  c->cell_cap = need;
  return 0;
}

int check_box(rs_ctx* c, const double* box) {
  for (int k = 0; k < 3; ++k)
    if (!(box[k] > 0.0) || !std::isfinite(box[k]) || c->cfg.cutoff > 0.5 * box[k])
      return fail(c, RS_E_CONFIG, "cutoff must be <= half of every box edge");
  if (c->cfg.precision == RS_PRECISION_FP32 && !(box[0] == box[1] && box[1] == box[2]))
    return fail(c, RS_E_CONFIG, "RS_PRECISION_FP32 needs a cubic box");
  return 0;
}

// New box edges: parameters, cell grid and list radius follow (the caller rebuilds the list).
int change_box(rs_ctx* c, const double* box) {
  int r = check_box(c, box);
  if (r) return r;
  for (int k = 0; k < 3; ++k) c->cfg.box[k] = box[k];
  set_params(c);
  return ensure_cells(c);
}

// Philox4x64-10 on the host (the stream the device uses for the noise, another key): the
// barostat's uniforms of step n are a pure function of (seed, n).
void philox_host(uint64_t ctr[4], uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
  for (int r = 0; r < 10; ++r) {
    const unsigned __int128 p0 = (unsigned __int128)M0 * ctr[0], p1 = (unsigned __int128)M1 * ctr[2];
    const uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
    const uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
    const uint64_t n0 = hi1 ^ ctr[1] ^ k0, n2 = hi0 ^ ctr[3] ^ k1;
    ctr[0] = n0;
    ctr[1] = lo1;
    ctr[2] = n2;
    ctr[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

void barostat_uniforms(uint64_t seed, int step, double u[2]) {
  uint64_t ctr[4] = {(uint64_t)(uint32_t)step, 0x42415253ULL, 0, 0};
  philox_host(ctr, seed, 0x6261726f73746174ULL);
  u[0] = (double)(ctr[0] >> 11) * (1.0 / 9007199254740992.0);
  u[1] = (double)(ctr[1] >> 11) * (1.0 / 9007199254740992.0);
}

// Adaptive state of one trajectory's barostat (MonteCarloBarostatNoUnits, water/water.jl:789-807:
// scale_factor 0.01, scale_increment 1.1, max_volume_frac 0.3, one iteration per call).
struct BarostatState {
  double volume_scale;
  int n_attempted = 0, n_accepted = 0;
  int64_t total_attempted = 0, total_accepted = 0;
};

double box_volume(const rs_ctx* c) { return c->cfg.box[0] * c->cfg.box[1] * c->cfg.box[2]; }
double kT_of(const rs_ctx* c);
void rebuild(rs_ctx* c, bool with_u, bool filed = false);

// One barostat move at step n (apply_coupling!, water/water.jl:811-865): U, a trial volume
// V + dV with dV = s (2u - 1), molecule centres scaled by (V'/V)^(1/3), the list rebuilt at the
// trial state (the reference's trial_find_neighbors = true form), U', and Metropolis on
// dW = dU + P dV - M kT ln(V'/V).
int barostat_move(rs_ctx* c, BarostatState& bs, uint64_t seed, int n) {
  DevBufs& b = c->L.b;
  const rs_config& g = c->cfg;
  double u[2], e0[8], e1[8];
  barostat_uniforms(seed, n, u);
  launch_energy(c->L);
  HIPCHK(c, hipMemcpyAsync(e0, b.eout, sizeof(e0), hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipMemcpyAsync(b.xbak, b.x, sizeof(double) * c->N3, hipMemcpyDeviceToDevice, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  const double V = box_volume(c);
  const double dV = bs.volume_scale * (2.0 * u[0] - 1.0);
  const double vs = (V + dV) / V, ls = std::cbrt(vs);
  const double old_box[3] = {g.box[0], g.box[1], g.box[2]};
  const double new_box[3] = {old_box[0] * ls, old_box[1] * ls, old_box[2] * ls};
  int r = check_box(c, new_box);
  if (r) return fail(c, RS_E_CONFIG, "the barostat shrank the box below twice the cutoff");
  launch_scale_molecules(c->L, ls, new_box);
  if ((r = change_box(c, new_box))) return r;
  rebuild(c, false);
  launch_energy(c->L);
  HIPCHK(c, hipMemcpyAsync(e1, b.eout, sizeof(e1), hipMemcpyDeviceToHost, c->s));
  HIPCHK(c, hipStreamSynchronize(c->s));
  const double kT = kT_of(c);
  const double P = g.barostat_pressure * 0.0602214076;  // bar -> kJ mol^-1 nm^-3
  const double dW = (e1[0] - e0[0]) + P * dV - c->M * kT * std::log(vs);
  const bool accept = dW <= 0.0 || u[1] < std::exp(-dW / kT);
  if (accept) {
    ++bs.n_accepted;
    ++bs.total_accepted;
  } else {
    HIPCHK(c, hipMemcpyAsync(b.x, b.xbak, sizeof(double) * c->N3, hipMemcpyDeviceToDevice, c->s));
    if ((r = change_box(c, old_box))) return r;
    rebuild(c, false);
  }
  ++bs.n_attempted;
  ++bs.total_attempted;
  if (bs.n_attempted >= 10) {
    if (bs.n_accepted < 0.25 * bs.n_attempted) {
      bs.volume_scale /= 1.1;
      bs.n_attempted = bs.n_accepted = 0;
    } else if (bs.n_accepted > 0.75 * bs.n_attempted) {
      bs.volume_scale = std::min(bs.volume_scale * 1.1, V * 0.3);
      bs.n_attempted = bs.n_accepted = 0;
    }
  }
  return 0;
}

void free_records(rs_ctx* c) {
  for (auto q : c->snap_x) hipFree(q);
  for (auto q : c->snap_v) hipFree(q);
  c->snap_x.clear();
  c->snap_v.clear();
  c->snap_steps.clear();
  c->rec_valid = false;
}

int make_plan(rs_ctx* c, int n_steps, Plan* p) {
  const rs_config& g = c->cfg;
  if (n_steps < 1) return fail(c, RS_E_CONFIG, "n_steps must be >= 1");
  p->n_loss = g.n_steps_loss > 0 ? g.n_steps_loss : n_steps;
  if (n_steps % p->n_loss) return fail(c, RS_E_CONFIG, "n_steps_loss must divide n_steps");
  p->n_blocks = n_steps / p->n_loss;
  p->n_trunc = g.n_steps_trunc > 0 ? g.n_steps_trunc : p->n_loss;
  if (p->n_trunc > p->n_loss) return fail(c, RS_E_CONFIG, "n_steps_trunc must not exceed n_steps_loss");
  p->n_log = g.n_steps_log;
  p->msd_on = g.msd_window_steps > 0 && g.msd_every > 0 && g.w_msd != 0.0;
  p->nf = 0;
  p->first_frame = 0;
  p->every = std::max(1, g.msd_every);
  if (p->msd_on) {
    if (g.msd_window_steps % g.msd_every || g.msd_window_steps > n_steps)
      return fail(c, RS_E_CONFIG, "msd window must be a multiple of msd_every and <= n_steps");
    p->nf = g.msd_window_steps / g.msd_every;
    if (p->nf < 2 || p->nf % 2 || p->nf > 128)
      return fail(c, RS_E_CONFIG, "msd window must hold an even number of frames in [2, 128]");
    p->first_frame = n_steps - g.msd_window_steps + g.msd_every;
  }
  return 0;
}

int ensure_frames(rs_ctx* c, int nf) {
  DevBufs& b = c->L.b;
  if (nf <= c->frames_cap) return 0;
  c->frames_cap = 0;  // until every buffer below is allocated again
  if (b.frames) hipFree(b.frames);
  if (b.msd_grad) hipFree(b.msd_grad);
  if (b.msd_del) hipFree(b.msd_del);
  if (b.msd_sq) hipFree(b.msd_sq);
  if (b.msd_part) hipFree(b.msd_part);
  int r;
  if ((r = dalloc(c, &b.frames, (size_t)nf * 3 * c->M)) ||
      (r = dalloc(c, &b.msd_grad, (size_t)nf * 3 * c->M)) ||
      (r = dalloc(c, &b.msd_del, (size_t)(nf / 2 + 1) * 3 * c->M)) ||
      (r = dalloc(c, &b.msd_sq, (size_t)(nf / 2 + 1) * c->M)) ||
      (r = dalloc(c, &b.msd_part, (size_t)(nf / 2 + 1) * ((c->M + 255) / 256))))
    return r;
  c->frames_cap = nf;
  return 0;
}

// Phase timer: tbegin records an event before a launch, tend after it; the elapsed times are
// collected after the stream has been synchronised (collect_timing).
hipEvent_t next_event(rs_ctx* c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    hipEventCreate(&e);
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}
void tbegin(rs_ctx* c, int phase) {
  if (!c->timing) return;
  hipEventRecord(next_event(c), c->s);
  c->mark_phase.push_back(phase);
}
void tend(rs_ctx* c) {
  if (!c->timing) return;
  hipEventRecord(next_event(c), c->s);
}
void collect_timing(rs_ctx* c) {
  if (!c->timing) return;
  for (size_t k = 0; k < c->mark_phase.size(); ++k) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, c->ev_pool[2 * k], c->ev_pool[2 * k + 1]) == hipSuccess) {
      c->ms[c->mark_phase[k]] += t;
      c->cnt[c->mark_phase[k]] += 1;
    }
  }
  c->mark_phase.clear();
  c->ev_used = 0;
}

void rebuild(rs_ctx* c, bool with_u, bool filed) {
  // the per-cell arrays must cover the cell grid; every change of the grid reallocates them
  // (ensure_cells), and cells_fit() refuses a call that would launch past them
  if (c->L.p.ncell + 1 > c->cell_cap) return;
  ++c->builds_call;
  tbegin(c, RS_PHASE_LIST);
  launch_list_build(c->L, with_u, filed);
  tend(c);
  c->since_build = 0;
}

// Host-side check before any launch that files molecules into cells or builds a list.
int cells_fit(rs_ctx* c) {
  if (c->L.p.ncell + 1 > c->cell_cap)
    return fail(c, RS_E_STATE, "internal: the per-cell arrays are shorter than the cell grid");
  return 0;
}

// Start of a call: clear the status flags, and the cell counts (an earlier call that ended
// between an integrator step that filed molecules and its list build leaves them counted).
int reset_flags(rs_ctx* c) {
  HIPCHK(c, hipMemsetAsync(c->L.b.flags, 0, sizeof(int) * FLAG_COUNT, c->s));
  HIPCHK(c, hipMemsetAsync(c->L.b.cell_count, 0, sizeof(int) * (size_t)c->L.p.ncell, c->s));
  return 0;
}

// Synchronise and translate the device flags into a status. *retry is set when the list was
// too small or too short-sighted for this call; the caller widens it (grow_or_widen) and reruns.
int finish(rs_ctx* c, int* retry) {
  HIPCHK(c, hipStreamSynchronize(c->s));
  collect_timing(c);
  int f[FLAG_COUNT];
  HIPCHK(c, hipMemcpy(f, c->L.b.flags, sizeof(f), hipMemcpyDeviceToHost));
  if (f[FLAG_NONFINITE]) return fail(c, RS_E_NONFINITE, "non-finite positions, velocities or forces");
  if (f[FLAG_MLP_RANGE]) return fail(c, RS_E_NONFINITE, "an O-O pair came closer than the MLP table's range");
  if (f[FLAG_FIXED_RANGE])
    return fail(c, RS_E_OVERFLOW, "an FP64 pair-force term exceeded the fixed-point sums' range (2^17 kJ/mol/nm)");
  *retry = (f[FLAG_LIST_OVERFLOW] ? 1 : 0) | (f[FLAG_STALE] ? 2 : 0) | (f[FLAG_HEXTENT] ? 4 : 0) |
           (f[FLAG_CELL_OVERFLOW] ? 8 : 0);
  return 0;
}

// After an overflow the per-lane capacity doubles; after a stale list the skin grows 1.5x (and the
// O-H extent grows by 0.01 nm after an extent violation); after a home cell held more molecules
// than a build block stages, the cells shrink (r_list/3, a 7^3 stencil), and where they cannot
// the call fails at once. The list radius and cell grid follow; the set of interacting pairs does
// not depend on them. Every rerun is counted (rs_stats).
int grow_or_widen(rs_ctx* c, int why, int attempt) {
  ++c->reruns_call;
  ++c->reruns_total;
  if (attempt >= 6)
    return fail(c, (why & 1) ? RS_E_OVERFLOW : RS_E_STALE,
                (why & 1) ? "neighbour list capacity exceeded after retries"
                          : "neighbour list still stale after widening the skin");
  const int capc = c->L.p.capc;
  if (why & 8) {
    const int cs = c->L.p.cs;
    if (c->L.p.brute || cs >= 3)
      return fail(c, RS_E_OVERFLOW, "a cell of the list grid holds more than 256 molecules");
    c->cell_refine = cs + 1 - 2;
    set_params(c);
    if (c->L.p.cs <= cs)
      return fail(c, RS_E_OVERFLOW,
                  "a cell of the list grid holds more than 256 molecules and the box is too small for finer cells");
  }
  if (why & 2) c->cfg.skin = 1.5 * c->cfg.skin + 0.005;
  if (why & 4) c->hext_extra += 0.01;
  if (why & 6) set_params(c);  // a list radius beyond a third of the box falls back to all pairs
  int r = ensure_cells(c);
  if (r) return r;
  c->L.p.capc = std::max(capc, default_capc(c));
  if (why & 1) c->L.p.capc = 2 * capc;
  return alloc_list(c);
}

double kT_of(const rs_ctx* c) { return c->cfg.kB * c->cfg.temperature; }

// Forward trajectory (forward_sim, water/water.jl:1029-1045). With `reweight` it is the forward
// loop of reweighting_sim (water/water.jl:1157-1192) instead: no records for a reverse pass, and at
// every block end the block loss, its direct θ-gradient and dU/dθ go into the reweighting
// accumulators.
int forward_impl(rs_ctx* c, const double* x0, const double* v0, uint64_t seed, int n_steps,
                 double* x_out, double* v_out, double* loss_out, bool reweight = false) {
  Plan P;
  int r = make_plan(c, n_steps, &P);
  if (r) return r;
  if (c->cfg.use_mlp && !c->mlp_set) return fail(c, RS_E_STATE, "use_mlp is set but rs_set_mlp was not called");
  if (c->ext_steps && c->ext_steps < n_steps)
    return fail(c, RS_E_CONFIG, "external noise holds fewer steps than requested");
  if ((r = cells_fit(c))) return r;
  DevBufs& b = c->L.b;
  const rs_config& g = c->cfg;
  const double dt = g.dt, e = std::exp(-g.gamma * dt), s = std::sqrt(1.0 - e * e), kT = kT_of(c);
  free_records(c);
  if (reweight && P.msd_on)
    return fail(c, RS_E_CONFIG, "the reweighting estimator takes block losses only: set w_msd = 0");
  if (reweight && c->cfg.use_mlp)
    return fail(c, RS_E_CONFIG, "the reweighting estimator needs the classical force field (use_mlp = 0)");
  const bool baro = g.barostat_every > 0;
  if (baro && P.msd_on) return fail(c, RS_E_CONFIG, "the MSD loss needs a constant box (barostat_every = 0)");
  if (baro && c->cfg.use_mlp) return fail(c, RS_E_CONFIG, "the barostat needs the classical force field (use_mlp = 0)");
  const double box0[3] = {g.box[0], g.box[1], g.box[2]};
  BarostatState bs;
  if (P.msd_on && (r = ensure_frames(c, P.nf))) return r;
  const double RT = kT;
  const bool dh = g.w_dhvap != 0.0;
  for (int t = 0; t <= n_steps && !reweight && !baro; ++t)
    if (P.snap_at(t)) {
      double *sx, *sv;
      if ((r = dalloc(c, &sx, c->N3)) || (r = dalloc(c, &sv, c->N3))) return r;
      c->snap_steps.push_back(t);
      c->snap_x.push_back(sx);
      c->snap_v.push_back(sv);
    }
  for (int attempt = 0;; ++attempt) {
    if (baro) {
      // every attempt starts from the call's box with a fresh barostat (water/water.jl:1133-1139)
      if ((r = change_box(c, box0))) return r;
      bs = BarostatState();
      bs.volume_scale = 0.01 * box_volume(c);
    }
    HIPCHK(c, hipMemcpyAsync(b.x, x0, sizeof(double) * c->N3, hipMemcpyHostToDevice, c->s));
    HIPCHK(c, hipMemcpyAsync(b.v, v0, sizeof(double) * c->N3, hipMemcpyHostToDevice, c->s));
    if ((r = reset_flags(c))) return r;
    HIPCHK(c, hipMemsetAsync(b.scal, 0, sizeof(double) * 16, c->s));
    if (reweight) HIPCHK(c, hipMemsetAsync(b.rwacc, 0, sizeof(double) * 32, c->s));
    launch_wrap_state(c->L);
    rebuild(c, false);
    launch_bonded(c->L, false, nullptr, 0, 0, nullptr, 0, 0);
    size_t snap = 0;
    auto take_snap = [&](int t) {
      if (snap < c->snap_steps.size() && c->snap_steps[snap] == t) {
        hipMemcpyAsync(c->snap_x[snap], b.x, sizeof(double) * c->N3, hipMemcpyDeviceToDevice, c->s);
        hipMemcpyAsync(c->snap_v[snap], b.v, sizeof(double) * c->N3, hipMemcpyDeviceToDevice, c->s);
        ++snap;
      }
    };
    take_snap(0);
    for (int n = 1; n <= n_steps; ++n) {
      // one kernel per step: the pair forces at x_{n-1}, then step n and the bonded forces at x_n
      tbegin(c, RS_PHASE_FWD_PAIR);
      const bool build = n % g.nl_every == 0;
      launch_pair_step(c->L, n, seed, dt, e, s, kT, build);
      std::swap(b.posq, b.posq_alt);
      std::swap(b.pos64, b.pos64_alt);
      tend(c);
      if (build) rebuild(c, false, true);
      take_snap(n);
      if (n % P.n_loss == 0 && g.w_rdf != 0.0) {
        tbegin(c, RS_PHASE_LOSS);
        launch_rdf(c->L, g.rdf_bins, g.rdf_rmax, g.rdf_width, g.rdf_window, g.w_rdf / P.n_blocks,
                   false, b.scal + 0);
        tend(c);
      }
      if (n % P.n_loss == 0 && (dh || reweight)) {
        // enthalpy of vaporization of the block (the state after the step, as the RDF)
        tbegin(c, RS_PHASE_LOSS);
        launch_energy(c->L);
        launch_dhvap(c->L, g.w_dhvap, g.dhvap_target, g.dhvap_u_gas, RT, dh ? 1.0 / P.n_blocks : 0.0,
                     false, false);
        if (reweight) launch_reweight_acc(c->L, g.w_rdf, dh);
        tend(c);
      }
      int f = P.frame_of(n);
      if (f >= 0) launch_frame_copy(c->L, f);
      if (baro && n % g.barostat_every == 0 && (r = barostat_move(c, bs, seed, n))) return r;
    }
    if (P.msd_on) {
      tbegin(c, RS_PHASE_LOSS);
      launch_msd(c->L, P.nf, (P.nf / 2) * g.msd_every * dt, g.w_msd, g.msd_target, b.scal + 0);
      tend(c);
    }
    int why = 0;
    if ((r = finish(c, &why))) return r;
    if (why) {
      if ((r = grow_or_widen(c, why, attempt))) return r;
      continue;
    }
    break;
  }
  double sc[16];
  HIPCHK(c, hipMemcpy(sc, b.scal, sizeof(sc), hipMemcpyDeviceToHost));
  c->msd_loss = P.msd_on ? sc[2] : 0.0;
  c->msd_D = P.msd_on ? sc[3] : 0.0;
  c->terms[0] = sc[8];
  c->terms[1] = sc[9];
  c->terms[2] = c->msd_loss;
  c->terms[3] = c->msd_D;
  if (x_out) HIPCHK(c, hipMemcpy(x_out, b.x, sizeof(double) * c->N3, hipMemcpyDeviceToHost));
  if (v_out) HIPCHK(c, hipMemcpy(v_out, b.v, sizeof(double) * c->N3, hipMemcpyDeviceToHost));
  if (loss_out) *loss_out = sc[0];
  c->rec_valid = !reweight && !baro;
  c->baro_attempted = baro ? bs.total_attempted : 0;
  c->baro_accepted = baro ? bs.total_accepted : 0;
  c->rec_steps = n_steps;
  c->rec_seed = seed;
  return 0;
}

// Gradient seed at state t into b.grad: RDF block loss (weight w_rdf / n_blocks) plus the MSD frame
// gradient; then reset or add into the accumulators.
void seed_at(rs_ctx* c, const Plan& P, int t, bool with_rdf, bool reset, double e) {
  const rs_config& g = c->cfg;
  DevBufs& b = c->L.b;
  tbegin(c, RS_PHASE_LOSS);
  launch_zero(c->L, b.grad, c->N3);
  if (with_rdf && g.w_rdf != 0.0)
    launch_rdf(c->L, g.rdf_bins, g.rdf_rmax, g.rdf_width, g.rdf_window, g.w_rdf / P.n_blocks, true,
               b.scal + 0);
  if (with_rdf && g.w_dhvap != 0.0) {
    // enthalpy-of-vaporization block loss: dl/dx = c' F into the seed, -c' dU/dθ into dL/dθ
    launch_energy(c->L);
    launch_dhvap(c->L, g.w_dhvap, g.dhvap_target, g.dhvap_u_gas, kT_of(c), 1.0 / P.n_blocks, true, true);
  }
  int f = P.frame_of(t);
  if (f >= 0) launch_msd_inject(c->L, f);
  launch_seed(c->L, reset, e);
  tend(c);
}

int restore(rs_ctx* c, int t) {
  for (size_t k = 0; k < c->snap_steps.size(); ++k)
    if (c->snap_steps[k] == t) {
      hipMemcpyAsync(c->L.b.x, c->snap_x[k], sizeof(double) * c->N3, hipMemcpyDeviceToDevice, c->s);
      hipMemcpyAsync(c->L.b.v, c->snap_v[k], sizeof(double) * c->N3, hipMemcpyDeviceToDevice, c->s);
      return 1;
    }
  return 0;
}

int reverse_impl(rs_ctx* c, const double* xN, const double* vN, uint64_t seed, int n_steps,
                 double* loss_out, double* dth_out, double* x0_out, double* v0_out) {
  if (!c->rec_valid || c->rec_steps != n_steps || c->rec_seed != seed)
    return fail(c, RS_E_STATE, "rs_reverse needs the rs_forward of the same seed and n_steps first");
  Plan P;
  int r = make_plan(c, n_steps, &P);
  if (r) return r;
  if ((r = cells_fit(c))) return r;
  DevBufs& b = c->L.b;
  const rs_config& g = c->cfg;
  const double dt = g.dt, e = std::exp(-g.gamma * dt), s = std::sqrt(1.0 - e * e), kT = kT_of(c);
  const double cs[2] = {1.0 / g.mass_O, 1.0 / g.mass_H};
  const double ws[2] = {dt * dt / (2.0 * g.mass_O), dt * dt / (2.0 * g.mass_H)};
  const bool clip = g.clip_norm > 0.0 && std::isfinite(g.clip_norm);
  const int nslots = pair_blocks(c->L.p);
  if (c->cfg.use_mlp && !c->mlp_set) return fail(c, RS_E_STATE, "use_mlp is set but rs_set_mlp was not called");
  for (int attempt = 0;; ++attempt) {
    if (c->L.use_mlp)
      HIPCHK(c, hipMemsetAsync(b.mlp_slot, 0, sizeof(double) * (size_t)pair_blocks(c->L.p) * MLP_NG, c->s));
    if (xN && vN) {
      HIPCHK(c, hipMemcpyAsync(b.x, xN, sizeof(double) * c->N3, hipMemcpyHostToDevice, c->s));
      HIPCHK(c, hipMemcpyAsync(b.v, vN, sizeof(double) * c->N3, hipMemcpyHostToDevice, c->s));
    } else if (!restore(c, n_steps)) {
      // no host state given (rs_value_and_grad): the forward pass's final snapshot, in HBM
      return fail(c, RS_E_STATE, "internal: no snapshot of the final forward state");
    }
    if ((r = reset_flags(c))) return r;
    HIPCHK(c, hipMemsetAsync(b.dth, 0, sizeof(double) * 8, c->s));
    HIPCHK(c, hipMemsetAsync(b.slots, 0, sizeof(double) * 8 * (size_t)nslots, c->s));
    HIPCHK(c, hipMemsetAsync(b.scal, 0, sizeof(double) * 16, c->s));
    HIPCHK(c, hipMemsetAsync(b.wmax, 0, sizeof(double) * 128, c->s));
    rebuild(c, false);
    seed_at(c, P, n_steps, true, true, e);
    int trunc_count = 0;
    bool pending = false;
    for (int n = n_steps; n >= 1; --n) {
      ++trunc_count;
      if (trunc_count <= P.n_trunc) {
        tbegin(c, RS_PHASE_REV_STEP);
        const bool build = c->since_build + 1 >= g.nl_every;
        launch_step_rev(c->L, pending, true, n, seed, dt, e, s, kT, nslots, pending && clip, build);
        tend(c);
        pending = false;
        ++c->since_build;
        if (build) rebuild(c, true, true);
        tbegin(c, RS_PHASE_REV_PAIR);
        c->L.parity = n & 1;
        launch_pair(c->L, true, b.A, ws[0], ws[1]);
        tend(c);
        if (clip) launch_hnorm(c->L, g.clip_norm);
        pending = true;
      }
      const int t = n - 1;
      if (t == 0) break;
      const bool blk = t % P.n_loss == 0;
      const bool logr = P.n_log > 0 && t % P.n_log == 0 && (trunc_count <= P.n_trunc || blk);
      const bool frame = P.frame_of(t) >= 0;
      if (!(blk || logr || frame)) continue;
      if (pending) {
        tbegin(c, RS_PHASE_REV_STEP);
        launch_step_rev(c->L, true, false, n, seed, dt, e, s, kT, nslots, clip);
        tend(c);
        pending = false;
      }
      bool restored = false;
      if (logr) restored = restore(c, t);
      if (blk) {
        if (!restored && trunc_count > P.n_trunc) restored = restore(c, t);
        trunc_count = 0;
        rebuild(c, false);
        seed_at(c, P, t, true, true, e);
      } else if (frame) {
        seed_at(c, P, t, false, false, e);
      }
    }
    if (pending) launch_step_rev(c->L, true, false, 1, seed, dt, e, s, kT, nslots, clip);
    launch_vjp_reduce(c->L);  // dL/dθ = Σ over pair blocks of their per-trajectory slots
    int why = 0;
    if ((r = finish(c, &why))) return r;
    if (why) {
      if ((r = grow_or_widen(c, why, attempt))) return r;
      continue;
    }
    break;
  }
  double sc[16], th[8];
  HIPCHK(c, hipMemcpy(sc, b.scal, sizeof(sc), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(th, b.dth, sizeof(th), hipMemcpyDeviceToHost));
  if (loss_out) *loss_out = sc[0] + c->msd_loss;
  // the reverse pass's own block losses (water/water.jl:1533-1546); the MSD term and D come from
  // the forward pass's frames
  c->terms[0] = sc[8];
  c->terms[1] = sc[9];
  c->terms[2] = c->msd_loss;
  c->terms[3] = c->msd_D;
  if (dth_out) std::memcpy(dth_out, th, sizeof(double) * 6);
  if (x0_out) HIPCHK(c, hipMemcpy(x0_out, b.x, sizeof(double) * c->N3, hipMemcpyDeviceToHost));
  if (v0_out) HIPCHK(c, hipMemcpy(v0_out, b.v, sizeof(double) * c->N3, hipMemcpyDeviceToHost));
  return 0;
}

// Single evaluation helpers: upload x, build the list (retrying on overflow), run `body`.
template <class Body>
int single_eval(rs_ctx* c, const double* x, bool with_u, Body body) {
  DevBufs& b = c->L.b;
  int r = cells_fit(c);
  if (r) return r;
  for (int attempt = 0;; ++attempt) {
    HIPCHK(c, hipMemcpyAsync(b.x, x, sizeof(double) * c->N3, hipMemcpyHostToDevice, c->s));
    if ((r = reset_flags(c))) return r;
    rebuild(c, with_u);
    body();
    HIPCHK(c, hipStreamSynchronize(c->s));
    collect_timing(c);
    int f[FLAG_COUNT];
    HIPCHK(c, hipMemcpy(f, b.flags, sizeof(f), hipMemcpyDeviceToHost));
    if (f[FLAG_NONFINITE]) return fail(c, RS_E_NONFINITE, "non-finite forces");
    if (f[FLAG_FIXED_RANGE])
      return fail(c, RS_E_OVERFLOW, "an FP64 pair-force term exceeded the fixed-point sums' range (2^17 kJ/mol/nm)");
    // a single evaluation builds its list at x: only capacity, the O-H extent and a cell's
    // occupancy can fail
    const int why = (f[FLAG_LIST_OVERFLOW] ? 1 : 0) | (f[FLAG_HEXTENT] ? 4 : 0) | (f[FLAG_CELL_OVERFLOW] ? 8 : 0);
    if (why) {
      if ((r = grow_or_widen(c, why, attempt))) return r;
      continue;
    }
    return 0;
  }
}

// Plain 16-byte vector copy: reads and writes exactly n * 16 bytes (grid-stride loop).
__global__ void k_copy16(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

int check_ctx(rs_ctx* c) {
  if (!c) return fail(nullptr, RS_E_CONFIG, "null context");
  HIPCHK(c, hipSetDevice(c->device));
  return 0;
}

// ---- scratch of rs_msd_noalloc: a file-scope device array (each GPU holds its own copy from the
// library's load until the process ends). Host code never uses the array's name as an address: at
// set-up each device's copy is looked up once (hipGetSymbolAddress with that device current) and
// that pointer is what kernels and copies get.
constexpr size_t NA_KM = (size_t)1 << 20;         // max origins x molecules of one call
constexpr size_t NA_PART = NA_KM / 256 + 64;      // chunk sums: K ceil(M/256) <= K M/256 + K
constexpr size_t NA_DOUBLES = 3 * NA_KM + NA_KM + NA_PART + 16;  // del, |del|^2, chunk sums, scalars
static __device__ __align__(256) double scratch[NA_DOUBLES];
constexpr int NA_MAX_DEVICES = 64;
double* g_scratch[NA_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

// The device's address of `scratch`; the context's device must be current.
int scratch_ptr(rs_ctx* c, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (c->device < 0 || c->device >= NA_MAX_DEVICES)
    return fail(c, RS_E_CONFIG, "rs_msd_noalloc supports device indices below 64");
  if (!g_scratch[c->device]) {
    double* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(c, RS_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[c->device] = ptr;
  }
  *out = g_scratch[c->device];
  return 0;
}

// check_ctx for the calls that run device work: also clears the per-call counters of rs_stats.
int begin_call(rs_ctx* c) {
  int r = check_ctx(c);
  if (r) return r;
  c->reruns_call = 0;
  c->builds_call = 0;
  return 0;
}

}  // namespace

// ============================================================================ C ABI

extern "C" {

int rs_version(void) { return RS_ABI_VERSION; }
int32_t rs_config_size(void) { return (int32_t)sizeof(rs_config); }

void rs_default_config(rs_config* g) {
  std::memset(g, 0, sizeof(*g));
  g->abi_version = RS_ABI_VERSION;
  g->n_molecules = 216;
  g->box[0] = g->box[1] = g->box[2] = 1.862;
  g->cutoff = 0.8;
  g->skin = 0.06;
  g->nl_every = 10;
  g->precision = RS_PRECISION_FP32;
  g->dt = 0.001;
  g->gamma = 1.0;
  g->temperature = 300.0;
  g->eps_rf = 78.3;
  g->coulomb_k = 138.935457644;
  g->kB = 0.00831446262;
  g->mass_O = 15.99943;
  g->mass_H = 1.007947;
  g->theta0 = 1.82421813418;
  g->rdf_bins = 200;
  g->rdf_rmax = 0.8;
  g->rdf_width = 0.0005;
  g->rdf_window = 0.04;
  g->w_rdf = 1.0;
  g->msd_target = 2.3e-9;
  g->clip_norm = INFINITY;
  g->w_dhvap = 0.0;
  g->dhvap_target = 44.12 + 2.8;
  g->dhvap_u_gas = 3.69;
}

const char* rs_last_error(const rs_ctx* c) { return c ? c->err.c_str() : g_last_error.c_str(); }

int rs_create(const rs_config* cfg, int32_t device, rs_ctx** out) {
  if (!out) return fail(nullptr, RS_E_CONFIG, "null output pointer");
  *out = nullptr;
  if (!cfg || cfg->abi_version != RS_ABI_VERSION)
    return fail(nullptr, RS_E_CONFIG, "abi_version mismatch");
  const rs_config& g = *cfg;
  if (g.n_molecules < 2) return fail(nullptr, RS_E_CONFIG, "n_molecules must be >= 2");
  for (int k = 0; k < 3; ++k)
    if (!(g.box[k] > 0.0) || g.cutoff > 0.5 * g.box[k])
      return fail(nullptr, RS_E_CONFIG, "cutoff must be positive and <= half of every box edge");
  if (!(g.cutoff > 0.0) || !(g.dt > 0.0) || !(g.gamma > 0.0) || g.nl_every < 1 || !(g.skin >= 0.0))
    return fail(nullptr, RS_E_CONFIG, "cutoff, dt, gamma must be > 0, nl_every >= 1, skin >= 0");
  if (g.precision != RS_PRECISION_FP32 && g.precision != RS_PRECISION_FP64)
    return fail(nullptr, RS_E_CONFIG, "precision must be RS_PRECISION_FP64 or RS_PRECISION_FP32");
  if (g.precision == RS_PRECISION_FP32 && !(g.box[0] == g.box[1] && g.box[1] == g.box[2]))
    return fail(nullptr, RS_E_CONFIG, "RS_PRECISION_FP32 needs a cubic box (pair arithmetic in units of L/2^32)");
  if (g.rdf_bins < 1 || g.rdf_bins > 1024 || !(g.rdf_rmax > 0.0) || !(g.rdf_width > 0.0) ||
      !(g.rdf_window > 0.0))
    return fail(nullptr, RS_E_CONFIG, "rdf_bins in [1,1024] and positive rdf_rmax/width/window required");
  if (g.use_mlp != 0 && g.use_mlp != 1) return fail(nullptr, RS_E_CONFIG, "use_mlp must be 0 or 1");
  if (g.barostat_every < 0 || !std::isfinite(g.barostat_pressure))
    return fail(nullptr, RS_E_CONFIG, "barostat_every must be >= 0 and barostat_pressure finite");
  if (g.use_mlp && g.w_dhvap != 0.0)
    return fail(nullptr, RS_E_CONFIG, "the enthalpy-of-vaporization loss needs the classical force field (use_mlp = 0)");
  if (!std::isfinite(g.w_dhvap) || !std::isfinite(g.dhvap_target) || !std::isfinite(g.dhvap_u_gas))
    return fail(nullptr, RS_E_CONFIG, "w_dhvap, dhvap_target and dhvap_u_gas must be finite");
  if (g.use_mlp && (g.cutoff <= MLP_RLO + 0.1))
    return fail(nullptr, RS_E_CONFIG, "the MLP term needs a cutoff above 0.25 nm");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
    return fail(nullptr, RS_E_NODEVICE, "no HIP device with that index");
  rs_ctx* c = new rs_ctx();
  c->cfg = g;
  c->device = device;
  c->M = g.n_molecules;
  c->N = 3 * c->M;
  c->N3 = 9 * c->M;
  const double th0[6] = {0.315, 0.636, 0.417, 462750.4, 0.09572, 836.8};
  std::memcpy(c->theta, th0, sizeof(th0));
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->s) != hipSuccess) {
    delete c;
    return fail(nullptr, RS_E_HIP, "hipSetDevice/hipStreamCreate failed");
  }
  c->L.s = c->s;
  c->L.precision = g.precision;
  c->L.use_mlp = 0;
  c->L.mlp_rc = g.cutoff;
  c->L.p.CH = g.pair_chunks > 0 ? g.pair_chunks : 16;  // lanes per molecule in the pair kernels
  if (c->L.p.CH != 8 && c->L.p.CH != 16) {
    rs_destroy(c);
    return fail(nullptr, RS_E_CONFIG, "pair_chunks must be 8 or 16");
  }
  set_params(c);
  if (g.rdf_rmax + g.rdf_window > c->rlist) {
    rs_destroy(c);
    return fail(nullptr, RS_E_CONFIG, "rdf_rmax + rdf_window must not exceed the list radius");
  }
  c->L.p.capc = default_capc(c);
  double* sp = nullptr;
  if (scratch_ptr(c, &sp)) {  // set-up of rs_msd_noalloc's scratch on this device
    std::string m = c->err;
    rs_destroy(c);
    return fail(nullptr, RS_E_HIP, m);
  }
  if (alloc_all(c)) {
    std::string m = c->err;
    rs_destroy(c);
    return fail(nullptr, RS_E_HIP, m);
  }
  hipStreamSynchronize(c->s);
  *out = c;
  return 0;
}

int rs_destroy(rs_ctx* c) {
  if (!c) return 0;
  hipSetDevice(c->device);
  if (c->s) hipStreamSynchronize(c->s);
  free_records(c);
  DevBufs& b = c->L.b;
  void* ptrs[] = {b.x, b.v, b.F, b.h, b.A, b.B, b.C, b.D, b.grad, b.cell, b.cell_count,
                  b.cell_start, b.cell_fill, b.tmp, b.order, b.rank, b.posq, b.pos64, b.posq_alt, b.pos64_alt, b.pu,
                  b.ud, b.O64, b.Oq, b.nbr, b.nbr_tmp, b.ncnt, b.flags, b.slots, b.vjb, b.dth, b.scal, b.wmax, b.hist, b.rdfG,
                  b.target, b.rdf_g, b.frames, b.msd_del, b.msd_sq, b.msd_part, b.msd_grad, b.ext_noise,
                  b.hpart, b.mlp_w, b.mlp_wb, b.mlp_tab, b.mlp_slot, b.mlp_G, b.mlp_part, b.mlp_dW,
                  b.Fe, b.eslots, b.eout, b.rwacc, b.xbak};
  for (void* q : ptrs)
    if (q) hipFree(q);
  for (hipEvent_t e : c->ev_pool) hipEventDestroy(e);
  if (c->s) hipStreamDestroy(c->s);
  delete c;
  return 0;
}

int rs_set_theta(rs_ctx* c, const double* th) {
  int r = check_ctx(c);
  if (r) return r;
  if (!th) return fail(c, RS_E_CONFIG, "null theta");
  for (int k = 0; k < 6; ++k)
    if (!std::isfinite(th[k])) return fail(c, RS_E_CONFIG, "non-finite theta");
  if (!(th[0] > 0.0) || !(th[2] != 0.0) || !(th[4] > 0.0))
    return fail(c, RS_E_CONFIG, "sigma_O and r0 must be > 0 and q_H != 0");
  std::memcpy(c->theta, th, sizeof(double) * 6);
  const int capc = c->L.p.capc;
  set_params(c);  // r0 sets the O-H extent, hence the list radius and the cell grid
  c->L.p.capc = capc;
  c->rec_valid = false;
  if ((r = ensure_cells(c))) return r;
  if (c->cfg.rdf_rmax + c->cfg.rdf_window > c->rlist)
    return fail(c, RS_E_CONFIG, "rdf_rmax + rdf_window must not exceed the list radius");
  c->rec_valid = false;
  return 0;
}

int rs_set_reverse_options(rs_ctx* c, int32_t n_trunc, double clip) {
  int r = check_ctx(c);
  if (r) return r;
  if (n_trunc < 0 || std::isnan(clip)) return fail(c, RS_E_CONFIG, "n_steps_trunc must be >= 0 and clip_norm not NaN");
  c->cfg.n_steps_trunc = n_trunc;
  c->cfg.clip_norm = clip;
  return 0;
}

int rs_set_target_rdf(rs_ctx* c, const double* t, int32_t n) {
  int r = check_ctx(c);
  if (r) return r;
  if (!t || n != c->cfg.rdf_bins) return fail(c, RS_E_CONFIG, "target must hold rdf_bins values");
  HIPCHK(c, hipMemcpy(c->L.b.target, t, sizeof(double) * n, hipMemcpyHostToDevice));
  return 0;
}

int rs_set_external_noise(rs_ctx* c, const double* xi, int32_t n_steps) {
  int r = check_ctx(c);
  if (r) return r;
  if (c->L.b.ext_noise) hipFree(c->L.b.ext_noise);
  c->L.b.ext_noise = nullptr;
  c->ext_steps = 0;
  c->rec_valid = false;
  if (!xi) return 0;
  if (n_steps < 1) return fail(c, RS_E_CONFIG, "n_steps must be >= 1");
  size_t n = (size_t)n_steps * c->N3;
  HIPCHK(c, hipMalloc((void**)&c->L.b.ext_noise, n * sizeof(double)));
  HIPCHK(c, hipMemcpy(c->L.b.ext_noise, xi, n * sizeof(double), hipMemcpyHostToDevice));
  c->ext_steps = n_steps;
  return 0;
}

int rs_forward(rs_ctx* c, const double* x0, const double* v0, uint64_t seed, int32_t n_steps,
               double* x_out, double* v_out, double* loss_out) {
  int r = begin_call(c);
  if (r) return r;
  if (!x0 || !v0) return fail(c, RS_E_CONFIG, "null x0/v0");
  return forward_impl(c, x0, v0, seed, n_steps, x_out, v_out, loss_out);
}

int rs_reverse(rs_ctx* c, const double* xN, const double* vN, uint64_t seed, int32_t n_steps,
               double* loss_out, double* dth, double* x0_out, double* v0_out) {
  int r = begin_call(c);
  if (r) return r;
  if (!xN || !vN) return fail(c, RS_E_CONFIG, "null xN/vN");
  return reverse_impl(c, xN, vN, seed, n_steps, loss_out, dth, x0_out, v0_out);
}

int rs_value_and_grad(rs_ctx* c, const double* x0, const double* v0, uint64_t seed,
                      int32_t n_steps, double* loss_out, double* dth, float* dL_dmlp) {
  int r = begin_call(c);
  if (r) return r;
  if (!x0 || !v0) return fail(c, RS_E_CONFIG, "null x0/v0");
  // the reverse pass starts from the forward pass's final state where it lies, in HBM
  if ((r = forward_impl(c, x0, v0, seed, n_steps, nullptr, nullptr, nullptr))) return r;
  if ((r = reverse_impl(c, nullptr, nullptr, seed, n_steps, loss_out, dth, nullptr, nullptr))) return r;
  if (dL_dmlp) {
    if (!c->L.use_mlp) return fail(c, RS_E_STATE, "dL_dmlp requested without the MLP term");
    std::vector<double> d(MLP_NW);
    if ((r = rs_mlp_grad(c, d.data(), MLP_NW))) return r;
    for (int k = 0; k < MLP_NW; ++k) dL_dmlp[k] = (float)d[k];
  }
  return 0;
}

int rs_reweight(rs_ctx* c, const double* x0, const double* v0, uint64_t seed, int32_t n_steps,
                double* loss_out, double* dth_out, double* x_out, double* v_out) {
  int r = begin_call(c);
  if (r) return r;
  if (!x0 || !v0) return fail(c, RS_E_CONFIG, "null argument");
  if ((r = forward_impl(c, x0, v0, seed, n_steps, x_out, v_out, nullptr, true))) return r;
  Plan P;
  if ((r = make_plan(c, n_steps, &P))) return r;
  double acc[19];
  HIPCHK(c, hipMemcpy(acc, c->L.b.rwacc, sizeof(acc), hipMemcpyDeviceToHost));
  // water/water.jl:1201-1207: dl_dp/B - β (l_dE_dp/B - loss dE_dp/B), loss = Σ L_b / B
  const double B = P.n_blocks, beta = 1.0 / kT_of(c);
  const double loss = acc[0] / B;
  if (loss_out) *loss_out = loss;
  if (dth_out)
    for (int k = 0; k < 6; ++k)
      dth_out[k] = acc[1 + k] / B - beta * (acc[13 + k] / B - loss * acc[7 + k] / B);
  return 0;
}

int rs_energy(rs_ctx* c, const double* x, double* U, double* dU, double* F) {
  int r = begin_call(c);
  if (r) return r;
  if (!x) return fail(c, RS_E_CONFIG, "null argument");
  if (c->cfg.use_mlp) return fail(c, RS_E_CONFIG, "rs_energy covers the classical force field only (use_mlp = 0)");
  if ((r = single_eval(c, x, false, [&] { launch_energy(c->L); }))) return r;
  double e[8];
  HIPCHK(c, hipMemcpy(e, c->L.b.eout, sizeof(e), hipMemcpyDeviceToHost));
  if (U) *U = e[0];
  if (dU) std::memcpy(dU, e + 1, sizeof(double) * 6);
  if (F) HIPCHK(c, hipMemcpy(F, c->L.b.Fe, sizeof(double) * c->N3, hipMemcpyDeviceToHost));
  return 0;
}

int rs_set_mlp(rs_ctx* c, const float* w, int64_t n) {
  int r = check_ctx(c);
  if (r) return r;
  if (!c->cfg.use_mlp) return fail(c, RS_E_CONFIG, "the context was created with use_mlp = 0");
  if (!w || n != MLP_NW) return fail(c, RS_E_CONFIG, "weights must hold 5313 values");
  std::vector<double> wd(MLP_NW);
  std::vector<unsigned short> wb(MLP_NW);
  for (int k = 0; k < MLP_NW; ++k) {
    if (!std::isfinite(w[k])) return fail(c, RS_E_CONFIG, "non-finite MLP weight");
    uint32_t u;
    std::memcpy(&u, &w[k], 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest bfloat16, ties to even
    u &= 0xFFFF0000u;
    float f;
    std::memcpy(&f, &u, 4);
    wd[k] = (double)f;
    wb[k] = (unsigned short)(u >> 16);
  }
  HIPCHK(c, hipMemcpy(c->L.b.mlp_w, wd.data(), sizeof(double) * MLP_NW, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->L.b.mlp_wb, wb.data(), sizeof(unsigned short) * MLP_NW, hipMemcpyHostToDevice));
  tbegin(c, RS_PHASE_MLP);
  launch_mlp_tabulate(c->L, c->cfg.cutoff);
  tend(c);
  HIPCHK(c, hipStreamSynchronize(c->s));
  collect_timing(c);
  c->L.use_mlp = 1;
  c->mlp_set = true;
  c->rec_valid = false;
  return 0;
}

int rs_mlp_grad(rs_ctx* c, double* dW, int64_t n) {
  int r = check_ctx(c);
  if (r) return r;
  if (!c->L.use_mlp) return fail(c, RS_E_STATE, "no MLP weights set");
  if (!dW || n != MLP_NW) return fail(c, RS_E_CONFIG, "dL_dW must hold 5313 values");
  tbegin(c, RS_PHASE_MLP);
  launch_mlp_grad(c->L, c->cfg.cutoff);
  tend(c);
  HIPCHK(c, hipStreamSynchronize(c->s));
  collect_timing(c);
  HIPCHK(c, hipMemcpy(dW, c->L.b.mlp_dW, sizeof(double) * MLP_NW, hipMemcpyDeviceToHost));
  return 0;
}

int rs_mlp_table(rs_ctx* c, double* f, double* df, int32_t capacity, int32_t* n_nodes, double* r_lo) {
  int r = check_ctx(c);
  if (r) return r;
  if (!c->L.use_mlp) return fail(c, RS_E_STATE, "no MLP weights set");
  if (n_nodes) *n_nodes = MLP_NT;
  if (r_lo) *r_lo = MLP_RLO;
  if (capacity < MLP_NT) return capacity == 0 ? 0 : fail(c, RS_E_CONFIG, "capacity below the node count");
  if (c->L.precision == RS_PRECISION_FP32) {
    std::vector<float2> t(MLP_NT);
    HIPCHK(c, hipMemcpy(t.data(), c->L.b.mlp_tab, sizeof(float2) * MLP_NT, hipMemcpyDeviceToHost));
    for (int k = 0; k < MLP_NT; ++k) {
      if (f) f[k] = t[k].x;
      if (df) df[k] = t[k].y;
    }
  } else {
    std::vector<double2> t(MLP_NT);
    HIPCHK(c, hipMemcpy(t.data(), c->L.b.mlp_tab, sizeof(double2) * MLP_NT, hipMemcpyDeviceToHost));
    for (int k = 0; k < MLP_NT; ++k) {
      if (f) f[k] = t[k].x;
      if (df) df[k] = t[k].y;
    }
  }
  return 0;
}

int rs_forces(rs_ctx* c, const double* x, double* F) {
  int r = begin_call(c);
  if (r) return r;
  if (!x || !F) return fail(c, RS_E_CONFIG, "null argument");
  if (c->cfg.use_mlp && !c->mlp_set) return fail(c, RS_E_STATE, "use_mlp is set but rs_set_mlp was not called");
  if ((r = single_eval(c, x, false, [&] {
         launch_bonded(c->L, false, nullptr, 0, 0, nullptr, 0, 0);
         launch_pair(c->L, false, nullptr, 0, 0);
       })))
    return r;
  HIPCHK(c, hipMemcpy(F, c->L.b.F, sizeof(double) * c->N3, hipMemcpyDeviceToHost));
  return 0;
}

int rs_adjoint_terms(rs_ctx* c, const double* x, const double* cc, const double* w, double* F,
                     double* h, double* vjp6) {
  int r = begin_call(c);
  if (r) return r;
  if (!x || !cc || !w) return fail(c, RS_E_CONFIG, "null argument");
  DevBufs& b = c->L.b;
  HIPCHK(c, hipMemcpy(b.C, cc, sizeof(double) * c->N3, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(b.A, w, sizeof(double) * c->N3, hipMemcpyHostToDevice));
  const double cs[2] = {1.0 / c->cfg.mass_O, 1.0 / c->cfg.mass_H};
  if (c->cfg.use_mlp && !c->mlp_set) return fail(c, RS_E_STATE, "use_mlp is set but rs_set_mlp was not called");
  if ((r = single_eval(c, x, true, [&] {
         c->L.parity = 0;
         hipMemsetAsync(b.wmax, 0, sizeof(double) * 128, c->s);
         hipMemsetAsync(b.slots, 0, sizeof(double) * 8 * (size_t)pair_blocks(c->L.p), c->s);
         if (c->L.use_mlp)
           hipMemsetAsync(b.mlp_slot, 0, sizeof(double) * (size_t)pair_blocks(c->L.p) * MLP_NG, c->s);
         launch_bonded(c->L, true, b.C, cs[0], cs[1], b.A, 1.0, 1.0);
         launch_pair(c->L, true, b.A, 1.0, 1.0);
       })))
    return r;
  if (F) HIPCHK(c, hipMemcpy(F, b.F, sizeof(double) * c->N3, hipMemcpyDeviceToHost));
  if (h) HIPCHK(c, hipMemcpy(h, b.h, sizeof(double) * c->N3, hipMemcpyDeviceToHost));
  if (vjp6) {
    const int ns = pair_blocks(c->L.p);
    c->slot_host.resize((size_t)ns * 8);
    HIPCHK(c, hipMemcpy(c->slot_host.data(), b.slots, sizeof(double) * ns * 8, hipMemcpyDeviceToHost));
    for (int k = 0; k < 6; ++k) {
      double s = 0.0;
      for (int q = 0; q < ns; ++q) s += c->slot_host[(size_t)q * 8 + k];
      vjp6[k] = s;
    }
  }
  return 0;
}

int rs_noise(rs_ctx* c, uint64_t seed, int32_t step, double* xi) {
  int r = check_ctx(c);
  if (r) return r;
  if (!xi || step < 1) return fail(c, RS_E_CONFIG, "step must be >= 1 and xi non-null");
  if (c->ext_steps && step > c->ext_steps) return fail(c, RS_E_CONFIG, "step beyond external noise");
  launch_noise(c->L, step, seed, kT_of(c), c->L.b.grad);
  HIPCHK(c, hipStreamSynchronize(c->s));
  HIPCHK(c, hipMemcpy(xi, c->L.b.grad, sizeof(double) * c->N3, hipMemcpyDeviceToHost));
  return 0;
}

int rs_rdf(rs_ctx* c, const double* x, double* g, int64_t* hist, double* loss, double* dLdx) {
  int r = begin_call(c);
  if (r) return r;
  if (!x) return fail(c, RS_E_CONFIG, "null x");
  const rs_config& cf = c->cfg;
  DevBufs& b = c->L.b;
  if ((r = single_eval(c, x, false, [&] {
         launch_zero(c->L, b.grad, c->N3);
         launch_rdf(c->L, cf.rdf_bins, cf.rdf_rmax, cf.rdf_width, cf.rdf_window, 1.0, dLdx != nullptr,
                    nullptr);
       })))
    return r;
  const int nb = cf.rdf_bins;
  if (g) HIPCHK(c, hipMemcpy(g, b.rdf_g, sizeof(double) * nb, hipMemcpyDeviceToHost));
  if (hist) HIPCHK(c, hipMemcpy(hist, b.hist, sizeof(int64_t) * nb, hipMemcpyDeviceToHost));
  if (loss) HIPCHK(c, hipMemcpy(loss, b.scal + 1, sizeof(double), hipMemcpyDeviceToHost));
  if (dLdx) HIPCHK(c, hipMemcpy(dLdx, b.grad, sizeof(double) * c->N3, hipMemcpyDeviceToHost));
  return 0;
}

int rs_msd(rs_ctx* c, const double* frames, int32_t nf, double* D, double* loss, double* dLdf) {
  int r = check_ctx(c);
  if (r) return r;
  if (!frames || nf < 2 || nf % 2 || nf / 2 > 64)
    return fail(c, RS_E_CONFIG, "frames must hold an even number of frames in [2, 128]");
  if (c->cfg.msd_every < 1) return fail(c, RS_E_CONFIG, "msd_every must be >= 1");
  if ((r = ensure_frames(c, nf))) return r;
  DevBufs& b = c->L.b;
  const size_t n = (size_t)nf * 3 * c->M;
  HIPCHK(c, hipMemcpy(b.frames, frames, sizeof(double) * n, hipMemcpyHostToDevice));
  launch_msd(c->L, nf, (nf / 2) * c->cfg.msd_every * c->cfg.dt, c->cfg.w_msd, c->cfg.msd_target, nullptr);
  HIPCHK(c, hipStreamSynchronize(c->s));
  double sc[4];
  HIPCHK(c, hipMemcpy(sc, b.scal, sizeof(sc), hipMemcpyDeviceToHost));
  if (D) *D = sc[3];
  if (loss) *loss = sc[2];
  if (dLdf) HIPCHK(c, hipMemcpy(dLdf, b.msd_grad, sizeof(double) * n, hipMemcpyDeviceToHost));
  return 0;
}

int rs_msd_noalloc(rs_ctx* c, const double* frames, int32_t nf, double* D, double* loss, double* dLdf) {
  int r = check_ctx(c);
  if (r) return r;
  if (!frames || !dLdf || nf < 2 || nf % 2 || nf / 2 > 64)
    return fail(c, RS_E_CONFIG, "frames and dL_dframes must be device arrays of an even number of frames in [2, 128]");
  if (c->cfg.msd_every < 1) return fail(c, RS_E_CONFIG, "msd_every must be >= 1");
  if ((size_t)(nf / 2) * (size_t)c->M > NA_KM)
    return fail(c, RS_E_CONFIG, "rs_msd_noalloc: n_frames/2 * n_molecules exceeds 1,048,576 (its static scratch)");
  double* s = nullptr;
  if ((r = scratch_ptr(c, &s))) return r;
  double *del = s, *sq = s + 3 * NA_KM, *part = sq + NA_KM, *scal = part + NA_PART;
  // no buffer needs clearing: every element the kernels read was written by them in this call
  launch_msd_bufs(c->L, nf, (nf / 2) * c->cfg.msd_every * c->cfg.dt, c->cfg.w_msd, c->cfg.msd_target, frames,
                  del, sq, part, scal, dLdf, nullptr);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->s));
  double sc[4];
  HIPCHK(c, hipMemcpy(sc, scal, sizeof(sc), hipMemcpyDeviceToHost));
  if (D) *D = sc[3];
  if (loss) *loss = sc[2];
  return 0;
}

int rs_molecule_pairs(rs_ctx* c, const double* x, int32_t* pairs, int64_t capacity, int64_t* n_pairs) {
  int r = begin_call(c);
  if (r) return r;
  if (!x || !n_pairs || (capacity > 0 && !pairs)) return fail(c, RS_E_CONFIG, "null argument");
  if ((r = single_eval(c, x, false, [] {}))) return r;
  const SysParams& p = c->L.p;
  const size_t lanes = (size_t)pair_blocks(p) * RS_BLOCK;
  const size_t entries = lanes * LIST_SHELLS * p.capc;
  std::vector<int> nbr(entries), order(c->M);
  std::vector<unsigned short> nbr16(p.list16 ? entries : 0);
  std::vector<int4> cnt(lanes);
  if (p.list16) {
    HIPCHK(c, hipMemcpy(nbr16.data(), c->L.b.nbr, sizeof(unsigned short) * entries, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < entries; ++k) nbr[k] = nbr16[k];
  } else {
    HIPCHK(c, hipMemcpy(nbr.data(), c->L.b.nbr, sizeof(int) * entries, hipMemcpyDeviceToHost));
  }
  HIPCHK(c, hipMemcpy(cnt.data(), c->L.b.ncnt, sizeof(int4) * cnt.size(), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(order.data(), c->L.b.order, sizeof(int) * c->M, hipMemcpyDeviceToHost));
  std::vector<std::pair<int, int>> out;
  for (int i = 0; i < c->M; ++i)
    for (int ch = 0; ch < p.CH; ++ch) {
      const size_t gt = (size_t)i * p.CH + ch;
      const int n[LIST_SHELLS] = {cnt[gt].x, cnt[gt].y, cnt[gt].z};
      for (int s = 0; s < LIST_SHELLS; ++s)
        for (int k = 0; k < n[s]; ++k) {
          int j = nbr[list_base(p, gt, s) + (size_t)k * RS_WAVE];
          int a = order[i], bb = order[j];
          if (a < bb) out.emplace_back(a, bb);
        }
    }
  std::sort(out.begin(), out.end());
  *n_pairs = (int64_t)out.size();
  const int64_t nw = std::min<int64_t>(capacity, (int64_t)out.size());
  for (int64_t k = 0; k < nw; ++k) {
    pairs[2 * k] = out[k].first;
    pairs[2 * k + 1] = out[k].second;
  }
  if ((int64_t)out.size() > capacity) return fail(c, RS_E_OVERFLOW, "pairs capacity too small");
  return 0;
}

double rs_list_radius(const rs_ctx* c) { return c ? c->rlist : 0.0; }

int rs_loss_terms(const rs_ctx* c, double* t) {
  if (!c || !t) return fail(nullptr, RS_E_CONFIG, "null argument");
  t[0] = c->terms[0];
  t[1] = c->terms[1];
  t[2] = c->terms[2];
  t[3] = c->terms[3] * 1e-6;  // nm^2/ps -> m^2/s
  return 0;
}

int rs_stats(const rs_ctx* c, int64_t* st, int32_t n) {
  if (!c || (!st && n > 0) || n < 0 || n > RS_STAT_COUNT)
    return fail(nullptr, RS_E_CONFIG, "rs_stats: null context or n outside [0, RS_STAT_COUNT]");
  const int64_t v[RS_STAT_COUNT] = {c->reruns_call, c->builds_call, c->reruns_total};
  for (int k = 0; k < n; ++k) st[k] = v[k];
  return 0;
}

int rs_get_box(const rs_ctx* c, double* box) {
  if (!c || !box) return fail(nullptr, RS_E_CONFIG, "null argument");
  for (int k = 0; k < 3; ++k) box[k] = c->cfg.box[k];
  return 0;
}

int rs_set_box(rs_ctx* c, const double* box) {
  int r = check_ctx(c);
  if (r) return r;
  if (!box) return fail(c, RS_E_CONFIG, "null argument");
  if ((r = change_box(c, box))) return r;
  c->L.p.capc = std::max(c->L.p.capc, default_capc(c));
  if ((r = alloc_list(c))) return r;
  c->rec_valid = false;
  return 0;
}

int rs_timings(rs_ctx* c, double* ms, int32_t n) {
  int r = check_ctx(c);
  if (r) return r;
  for (int k = 0; k < n && k < RS_PHASE_COUNT; ++k) {
    ms[k] = c->ms[k];
    c->ms[k] = 0.0;
    c->cnt[k] = 0;
  }
  return 0;
}

int rs_timing_counts(rs_ctx* c, int64_t* counts, int32_t n) {
  int r = check_ctx(c);
  if (r) return r;
  for (int k = 0; k < n && k < RS_PHASE_COUNT; ++k) counts[k] = c->cnt[k];
  return 0;
}

int rs_enable_timing(rs_ctx* c, int32_t on) {
  int r = check_ctx(c);
  if (r) return r;
  c->timing = on != 0;
  return 0;
}

int rs_copy_calibration(rs_ctx* c, int64_t bytes, int32_t reps) {
  int r = check_ctx(c);
  if (r) return r;
  if (bytes <= 0 || bytes % 16 || bytes > ((int64_t)1 << 34) || reps < 1 || reps > 100)
    return fail(c, RS_E_CONFIG, "rs_copy_calibration: bytes must be a positive multiple of 16 up to 16 GiB, reps 1..100");
  uint4 *a = nullptr, *b = nullptr;
  HIPCHK(c, hipMalloc((void**)&a, (size_t)bytes));
  if (hipMalloc((void**)&b, (size_t)bytes) != hipSuccess) {
    hipFree(a);
    return fail(c, RS_E_HIP, "rs_copy_calibration: hipMalloc failed");
  }
  const size_t n = (size_t)bytes / 16;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 16384);
  hipError_t e = hipMemsetAsync(a, 0, (size_t)bytes, c->s);
  for (int k = 0; k < reps && e == hipSuccess; ++k) {
    k_copy16<<<blocks, 256, 0, c->s>>>(a, b, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->s);
  hipFree(a);
  hipFree(b);
  if (e != hipSuccess) return fail(c, RS_E_HIP, std::string("rs_copy_calibration: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
