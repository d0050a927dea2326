// energy.hip — potential energy U(x), its θ-gradient dU/dθ and the forces -dU/dx in FP64, the
// enthalpy-of-vaporization block loss, and the accumulators of the ensemble-reweighting estimator.
//
// U is what the reference differentiates in enthalpy_vaporization (potential_energy,
// water/water.jl:738-746) and in reweighting_sim (gradient(potential_energy, sys, ...),
// water/water.jl:1183-1191):
//   reaction field  k q_a q_b (1/r + k_rf r^2 - c_rf), c_rf = (1/r_c) 3ε_rf/(2ε_rf + 1), so the
//                   pair energy is zero at r_c and its -d/dr is the force of water/water.jl:903
//   Lennard-Jones   4ε((σ/r)^12 - (σ/r)^6) on O-O pairs within r_c (truncated, as the force)
//   bonds           k_b/2 (r - r0)^2;  angle  k_a/2 (θ - θ0)^2
// dU/dθ for θ = (σ_O, ε_O, q_H, k_b, r0, k_a): (24ε/σ)(2s^12 - s^6), 4(s^12 - s^6), 2 U_RF/q_H
// (every charge product is q_H^2 times a constant), (r - r0)^2/2, -k_b (r - r0), (θ - θ0)^2/2.
//
// Evaluated once per loss block, so FP64 in both precisions (the pair set is the FP64 rule's).
// Owner-computes over the full molecule list as in the pair kernel: CH lanes own a molecule, each
// pair's energy is halved on each side, the CH partials are combined by a fixed shuffle tree, and
// the per-block sums are reduced in block order: deterministic.
#include "internal.h"
#include "bonded.h"

namespace {

constexpr int NE = 7;  // U, dU/dθ[6]

template <int CH, bool L16>
__global__ __launch_bounds__(RS_BLOCK) void k_energy(SysParams p,
                                                      const typename ListEntry<L16>::type* __restrict__ nbr,
                                                      const int4* __restrict__ ncnt,
                                                      const int* __restrict__ order,
                                                      const double* __restrict__ x, double* Fe,
                                                      double* slots, int* flags) {
#pragma clang fp contract(off)
  const int gtid = blockIdx.x * RS_BLOCK + threadIdx.x;
  const int i = gtid / CH, c = gtid % CH;
  const bool act = i < p.M;
  const int mi = act ? order[i] : 0;
  const double* xi = x + 9 * (size_t)mi;
  double f[9], urf = 0.0, ulj = 0.0, dsig = 0.0;
#pragma unroll
  for (int t = 0; t < 9; ++t) f[t] = 0.0;
  if (act) {
    const int4 nc4 = ncnt[gtid];
    for (int s = 0; s < LIST_SHELLS; ++s) {
      const int n = s == 0 ? nc4.x : (s == 1 ? nc4.y : nc4.z);
      const typename ListEntry<L16>::type* L = nbr + list_base(p, gtid, s);
      const double sig2 = p.sig * p.sig;
      for (int k = 0; k < n; ++k) {
        const double* xj = x + 9 * (size_t)order[L[(size_t)k * RS_WAVE]];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            double d[3];
#pragma unroll
            for (int t = 0; t < 3; ++t) d[t] = mic1(xj[3 * b + t] - xi[3 * a + t], p.L[t], p.invL[t]);
            const double r2 = r2_of(d[0], d[1], d[2]);
            if (!(r2 <= p.rc2)) continue;
            const double r = sqrt(r2), ir2 = 1.0 / r2;
            const double cp = p.kqq[(a > 0) + (b > 0)];
            urf += cp * (1.0 / r + p.krf * r2 - p.crf);
            double fr = cp * (ir2 / r - 2.0 * p.krf);  // f/r
            if (a == 0 && b == 0 && p.eps != 0.0) {
              const double s2 = sig2 * ir2, s6 = s2 * s2 * s2, s12 = s6 * s6;
              ulj += 4.0 * p.eps * (s12 - s6);
              dsig += (24.0 * p.eps / p.sig) * (2.0 * s12 - s6);
              fr += 24.0 * p.eps * ir2 * (2.0 * s12 - s6);
            }
#pragma unroll
            for (int t = 0; t < 3; ++t) f[3 * a + t] -= fr * d[t];
          }
        }
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 9; ++t) f[t] = chunk_sum<CH>(f[t]);
  urf = chunk_sum<CH>(urf);
  ulj = chunk_sum<CH>(ulj);
  dsig = chunk_sum<CH>(dsig);

  __shared__ double part[RS_BLOCK / CH][NE];
  if (c == 0) {
    double e[NE] = {0, 0, 0, 0, 0, 0, 0};
    if (act) {
      // each pair was seen from both molecules: half of its energy here
      e[0] = 0.5 * (urf + ulj);
      e[1] = 0.5 * dsig;
      e[2] = p.eps != 0.0 ? 0.5 * ulj / p.eps : 0.0;
      e[3] = urf / p.qH;  // 2 (U_RF/2) / q_H
      // intramolecular terms
      double xm[9], Fb[9], hb[9], vj[3];
#pragma unroll
      for (int t = 0; t < 9; ++t) xm[t] = xi[t];
      bonded_mol<false>(p, xm, nullptr, nullptr, Fb, hb, vj, flags, make_double4(xm[0], xm[1], xm[2], 0.0));
      double ub = 0.0;
      double dv[2][3];
      for (int hh = 1; hh <= 2; ++hh) {
        for (int t = 0; t < 3; ++t) dv[hh - 1][t] = mic1(xm[3 * hh + t] - xm[t], p.L[t], p.invL[t]);
        const double r = sqrt(r2_of(dv[hh - 1][0], dv[hh - 1][1], dv[hh - 1][2]));
        const double disp = r - p.r0;
        ub += 0.5 * p.kb * disp * disp;
        e[4] += 0.5 * disp * disp;
        e[5] += -p.kb * disp;
      }
      const double ra = sqrt(r2_of(dv[0][0], dv[0][1], dv[0][2]));
      const double rb = sqrt(r2_of(dv[1][0], dv[1][1], dv[1][2]));
      double cs = (dv[0][0] * dv[1][0] + dv[0][1] * dv[1][1] + dv[0][2] * dv[1][2]) / (ra * rb);
      cs = cs > 1.0 ? 1.0 : (cs < -1.0 ? -1.0 : cs);
      const double da = acos(cs) - p.theta0;
      ub += 0.5 * p.ka * da * da;
      e[6] = 0.5 * da * da;
      e[0] += ub;
#pragma unroll
      for (int t = 0; t < 9; ++t) Fe[9 * (size_t)mi + t] = f[t] + Fb[t];
    }
#pragma unroll
    for (int t = 0; t < NE; ++t) part[threadIdx.x / CH][t] = e[t];
  }
  __syncthreads();
  if (threadIdx.x < NE) {
    double s = 0.0;
    for (int q = 0; q < RS_BLOCK / CH; ++q) s += part[q][threadIdx.x];
    slots[(size_t)blockIdx.x * 8 + threadIdx.x] = s;
  }
}

// out[k] = Σ_blocks slots[blk*8 + k] in block order (one block per component).
__global__ __launch_bounds__(RS_BLOCK) void k_energy_reduce(const double* __restrict__ slots,
                                                             int nslots, double* out) {
  __shared__ double red[RS_BLOCK];
  const int k = blockIdx.x;
  double acc = 0.0;
  for (int q = threadIdx.x; q < nslots; q += RS_BLOCK) acc += slots[(size_t)q * 8 + k];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = RS_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

// Enthalpy of vaporization of the block (water/water.jl:738-754) from e = (U, dU/dθ):
// dH = u_gas - U/M + RT, l = w (dH - target)^2, c' = dl/d(-U) = 2 w (dH - target)/M. scal[6] = l,
// scal[7] = c'. With wgt != 0 (a block of the averaged trajectory loss): scal[0] and scal[9] (the
// trajectory's enthalpy term) += wgt l and, if direct, dth[k] -= wgt c' dU/dθ_k (the loss's own
// θ-dependence through U).
__global__ void k_dhvap(int M, const double* e, double w, double target, double ugas, double RT,
                        double wgt, int direct, double* scal, double* dth) {
  if (threadIdx.x != 0) return;
  const double dH = ugas - e[0] / M + RT;
  const double l = w * (dH - target) * (dH - target);
  const double cp = 2.0 * w * (dH - target) / M;
  scal[6] = l;
  scal[7] = cp;
  if (wgt != 0.0) {
    scal[0] += wgt * l;
    scal[9] += wgt * l;
    if (direct)
      for (int k = 0; k < 6; ++k) dth[k] -= wgt * cp * e[1 + k];
  }
}

// grad += wgt c' Fe: dl/dx = c' F for the enthalpy-of-vaporization loss.
__global__ void k_dhvap_seed(int n3, const double* __restrict__ Fe, const double* scal, double wgt,
                             double* grad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n3) grad[t] += (wgt * scal[7]) * Fe[t];
}

// Reweighting accumulators of one block: acc[0] += L_b, acc[1+k] += dL_b/dθ_k (direct),
// acc[7+k] += dU/dθ_k, acc[13+k] += L_b dU/dθ_k. L_b = w_rdf L_rdf (scal[1]) + l_dH (scal[6]).
__global__ void k_reweight_acc(const double* scal, const double* e, double w_rdf, int dh, double* acc) {
  if (threadIdx.x != 0) return;
  const double L = (w_rdf != 0.0 ? w_rdf * scal[1] : 0.0) + (dh ? scal[6] : 0.0);
  acc[0] += L;
  for (int k = 0; k < 6; ++k) {
    acc[1 + k] += dh ? -scal[7] * e[1 + k] : 0.0;
    acc[7 + k] += e[1 + k];
    acc[13 + k] += L * e[1 + k];
  }
}

}  // namespace

void launch_energy(const Launch& L) {
  const SysParams& p = L.p;
  const int g = pair_blocks(p);
  const unsigned short* n16 = (const unsigned short*)L.b.nbr;
  const int* n32 = (const int*)L.b.nbr;
  if (p.CH == 8 && p.list16)
    k_energy<8, true><<<g, RS_BLOCK, 0, L.s>>>(p, n16, L.b.ncnt, L.b.order, L.b.x, L.b.Fe, L.b.eslots, L.b.flags);
  else if (p.CH == 8)
    k_energy<8, false><<<g, RS_BLOCK, 0, L.s>>>(p, n32, L.b.ncnt, L.b.order, L.b.x, L.b.Fe, L.b.eslots, L.b.flags);
  else if (p.list16)
    k_energy<16, true><<<g, RS_BLOCK, 0, L.s>>>(p, n16, L.b.ncnt, L.b.order, L.b.x, L.b.Fe, L.b.eslots, L.b.flags);
  else
    k_energy<16, false><<<g, RS_BLOCK, 0, L.s>>>(p, n32, L.b.ncnt, L.b.order, L.b.x, L.b.Fe, L.b.eslots, L.b.flags);
  k_energy_reduce<<<NE, RS_BLOCK, 0, L.s>>>(L.b.eslots, g, L.b.eout);
}

void launch_dhvap(const Launch& L, double w, double target, double ugas, double RT, double wgt,
                  bool direct, bool seed) {
  k_dhvap<<<1, 64, 0, L.s>>>(L.p.M, L.b.eout, w, target, ugas, RT, wgt, direct ? 1 : 0, L.b.scal, L.b.dth);
  if (seed) {
    const int n3 = 3 * L.p.N;
    k_dhvap_seed<<<(n3 + RS_BLOCK - 1) / RS_BLOCK, RS_BLOCK, 0, L.s>>>(n3, L.b.Fe, L.b.scal, wgt, L.b.grad);
  }
}

void launch_reweight_acc(const Launch& L, double w_rdf, bool dh) {
  k_reweight_acc<<<1, 64, 0, L.s>>>(L.b.scal, L.b.eout, w_rdf, dh ? 1 : 0, L.b.rwacc);
}

// ------------------------------------------------------------------ Monte Carlo barostat move

namespace {

// Molly's scale_coords! for a system with molecules (called by the reference's barostat,
// water/water.jl:831): each water's centre of mass, with its H atoms taken at the image nearest
// their O in the old box, is scaled by l; the molecule moves rigidly with it and its atoms are
// wrapped into the new box (edges Ln).
__global__ void k_scale_molecules(SysParams p, double* x, double l, double Ln0, double Ln1, double Ln2) {
#pragma clang fp contract(off)
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= p.M) return;
  double* xm = x + 9 * (size_t)m;
  const double Ln[3] = {Ln0, Ln1, Ln2};
  double a[9];
  for (int k = 0; k < 3; ++k) {
    a[k] = xm[k];
    a[3 + k] = xm[k] + mic1(xm[3 + k] - xm[k], p.L[k], p.invL[k]);
    a[6 + k] = xm[k] + mic1(xm[6 + k] - xm[k], p.L[k], p.invL[k]);
  }
  const double mt = p.mO + 2.0 * p.mH;
  for (int k = 0; k < 3; ++k) {
    const double com = (p.mO * a[k] + p.mH * a[3 + k] + p.mH * a[6 + k]) / mt;
    const double sh = (l - 1.0) * com;
    for (int t = 0; t < 3; ++t) xm[3 * t + k] = wrap1(a[3 * t + k] + sh, Ln[k]);
  }
}

}  // namespace

void launch_scale_molecules(const Launch& L, double l, const double* Lnew) {
  k_scale_molecules<<<(L.p.M + RS_BLOCK - 1) / RS_BLOCK, RS_BLOCK, 0, L.s>>>(L.p, L.b.x, l, Lnew[0], Lnew[1], Lnew[2]);
}
