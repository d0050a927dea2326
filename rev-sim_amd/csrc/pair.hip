// pair.hip — the fused force / adjoint kernel: reaction-field Coulomb + O-O Lennard-Jones over the
// molecule list, plus the intramolecular harmonic bonds and angle, in one pass.
//
// Forward variant (calc_F!, water/water.jl:877-975): F(x).
// Adjoint variant (calc_F_and_grads!, water/water.jl:1212-1433): F(x), the Hessian-vector product
//   h = (dF/dx)^T u with u = accum_C / m, and the θ-VJP sum_dof w_dof dF_dof/dθ with
//   w = accum_A dt^2/(2m) (water/water.jl:1517-1519), θ = (σ_O, ε_O, q_H, k_b, r0, k_a).
//
// Work decomposition: CH consecutive lanes own one molecule i (sorted order); lane c walks
// sub-list c of i's full molecule list, evaluates the 3x3 atom pairs of each listed molecule j and
// accumulates only i's side. The CH partial sums are combined by xor-shuffles in a fixed order, so
// the result is deterministic without atomics. Lane c == 0 then adds the bonded terms (FP64) and
// writes the molecule's outputs in the caller's atom order.
//
// Pair arithmetic: FP32 on quantized positions (the int32 difference of 2^32-periodic positions is
// the minimum image by construction) or FP64 (parity mode). Per pair of atoms a (own), b:
//   f/r  = k q_a q_b (1/r^3 - 2 k_rf)              [+ 24ε/r^2 (2 s^12 - s^6) for O-O, s = σ/r]
//   F_a -= (f/r) d, d = x_b - x_a
//   h_a += K (u_a - u_b), K = f' r̂r̂ᵀ + (f/r)(I - r̂r̂ᵀ), i.e. (f/r) Δu + ((f' - f/r)/r^2)(d.Δu) d
//          with (f' - f/r)/r^2 = -3 k q_a q_b / r^5   [- 24ε/r^4 (28 s^12 - 8 s^6) for O-O]
//   q-VJP: dF/dq_H = 2 F_RF / q_H; ε-VJP: dF/dε = F_LJ/ε; σ-VJP: dF_a/dσ = -(144ε/(σ r^2))(4s^12-s^6) d
#include "internal.h"

namespace {

struct PairArgs {
  SysParams p;
  const int* __restrict__ nbr;
  const int* __restrict__ ncnt;
  const int* __restrict__ order;
  const void* __restrict__ pos;
  const void* __restrict__ u;
  const double* __restrict__ x;
  const double* __restrict__ csrc;
  const double* __restrict__ wsrc;
  double cscale[2], wscale[2];
  const double4* __restrict__ O64;
  double* F;
  double* h;
  double* slots;
  int* flags;
};

struct P32 {
  typedef int4 Pos;
  typedef float4 U;
  typedef float R;
  static __device__ __forceinline__ void diff(const SysParams& p, const int4& a, const int4& b,
                                              float& dx, float& dy, float& dz) {
    dx = qdiff((uint32_t)a.x, (uint32_t)b.x, p.unitf[0]);
    dy = qdiff((uint32_t)a.y, (uint32_t)b.y, p.unitf[1]);
    dz = qdiff((uint32_t)a.z, (uint32_t)b.z, p.unitf[2]);
  }
  static __device__ __forceinline__ float r2(float dx, float dy, float dz) { return r2f_of(dx, dy, dz); }
  static __device__ __forceinline__ float rc2(const SysParams& p) { return p.rc2f; }
  static __device__ __forceinline__ float rinv(float r2) { return __builtin_amdgcn_rsqf(r2); }
};

struct P64 {
  typedef double4 Pos;
  typedef double4 U;
  typedef double R;
  static __device__ __forceinline__ void diff(const SysParams& p, const double4& a, const double4& b,
                                              double& dx, double& dy, double& dz) {
    dx = mic1(b.x - a.x, p.L[0], p.invL[0]);
    dy = mic1(b.y - a.y, p.L[1], p.invL[1]);
    dz = mic1(b.z - a.z, p.L[2], p.invL[2]);
  }
  static __device__ __forceinline__ double r2(double dx, double dy, double dz) { return r2_of(dx, dy, dz); }
  static __device__ __forceinline__ double rc2(const SysParams& p) { return p.rc2; }
  static __device__ __forceinline__ double rinv(double r2) { return 1.0 / sqrt(r2); }
};

// ------------------------------------------------------------------ bonded terms (FP64)

struct dd {
  double v, d;
};
__device__ __forceinline__ dd mk(double v, double d) { return dd{v, d}; }
__device__ __forceinline__ dd operator+(dd a, dd b) { return mk(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ dd operator-(dd a, dd b) { return mk(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ dd operator-(dd a) { return mk(-a.v, -a.d); }
__device__ __forceinline__ dd operator*(dd a, dd b) { return mk(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ dd operator/(dd a, dd b) {
  return mk(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v));
}
__device__ __forceinline__ dd dsqrt(dd a) {
  double s = sqrt(a.v);
  return mk(s, a.d / (2.0 * s));
}
__device__ __forceinline__ dd dacos(dd a) { return mk(acos(a.v), -a.d / sqrt(1.0 - a.v * a.v)); }
__device__ __forceinline__ void dcross(const dd a[3], const dd b[3], dd c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ dd ddot(const dd a[3], const dd b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Bonds O-H1, O-H2 and angle H1-O-H2 of one molecule (water/water.jl:915-968, 1279-1422).
// The angle's HVP is the tangent part of its force evaluated on dual numbers seeded with u:
// (dF/dx) u = (dF/dx)^T u because dF/dx is minus the (symmetric) Hessian of the energy.
template <bool ADJ>
__device__ void bonded(const SysParams& p, const double* xm, const double* um, const double* wm,
                       double F[9], double h[9], double vj[6], int* flags) {
  double O[3] = {xm[0], xm[1], xm[2]};
  for (int hh = 1; hh <= 2; ++hh) {
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = mic1(xm[3 * hh + k] - O[k], p.L[k], p.invL[k]);
    double r = sqrt(r2_of(d[0], d[1], d[2]));
    if (r > p.hext) atomicOr(&flags[FLAG_HEXTENT], 1);
    double n[3] = {d[0] / r, d[1] / r, d[2] / r};
    double disp = r - p.r0;
    double Fv = -p.kb * disp;
    for (int k = 0; k < 3; ++k) {
      F[k] -= Fv * n[k];
      F[3 * hh + k] += Fv * n[k];
    }
    if (ADJ) {
      double du[3], dot = 0.0;
      for (int k = 0; k < 3; ++k) {
        du[k] = um[k] - um[3 * hh + k];
        dot += n[k] * du[k];
      }
      double Fr = Fv / r, Fp = -p.kb;
      for (int k = 0; k < 3; ++k) {
        double hv = Fr * du[k] + (Fp - Fr) * dot * n[k];
        h[k] += hv;
        h[3 * hh + k] -= hv;
        double dw = wm[k] - wm[3 * hh + k];
        vj[3] += disp * n[k] * dw;
        vj[4] += -p.kb * n[k] * dw;
      }
    }
  }
  // angle, i = H1, j = O (centre), k = H2
  dd a[3], b[3];
  for (int k = 0; k < 3; ++k) {
    a[k] = mk(mic1(xm[3 + k] - O[k], p.L[k], p.invL[k]), ADJ ? um[3 + k] - um[k] : 0.0);
    b[k] = mk(mic1(xm[6 + k] - O[k], p.L[k], p.invL[k]), ADJ ? um[6 + k] - um[k] : 0.0);
  }
  dd cr[3];
  dcross(a, b, cr);
  if (cr[0].v == 0.0 && cr[1].v == 0.0 && cr[2].v == 0.0) return;
  dd rji = dsqrt(ddot(a, a)), rjk = dsqrt(ddot(b, b));
  dd p1u[3], p3u[3], mb[3];
  dcross(a, cr, p1u);
  for (int k = 0; k < 3; ++k) mb[k] = -b[k];
  dcross(mb, cr, p3u);
  dd p1n = dsqrt(ddot(p1u, p1u)), p3n = dsqrt(ddot(p3u, p3u));
  dd cs = ddot(a, b) / (rji * rjk);
  dd th = dacos(cs);
  dd disp = mk(p.theta0, 0.0) - th;
  dd Fa = mk(p.ka, 0.0) * disp;
  for (int k = 0; k < 3; ++k) {
    dd p1d = p1u[k] / p1n / rji, p3d = p3u[k] / p3n / rjk;
    dd gi = Fa * p1d, gk = Fa * p3d, gj = -(gi + gk);
    F[3 + k] += gi.v;
    F[6 + k] += gk.v;
    F[k] += gj.v;
    if (ADJ) {
      h[3 + k] += gi.d;
      h[6 + k] += gk.d;
      h[k] += gj.d;
      vj[5] += disp.v * (p1d.v * wm[3 + k] + p3d.v * wm[6 + k] - (p1d.v + p3d.v) * wm[k]);
    }
  }
}

// ------------------------------------------------------------------ the pair kernel

template <class P, int CH, bool ADJ>
__global__ __launch_bounds__(RS_BLOCK) void k_pair(PairArgs A) {
  typedef typename P::R R;
  typedef typename P::Pos Pos;
  typedef typename P::U U;
  const SysParams& p = A.p;
  const int gtid = blockIdx.x * RS_BLOCK + threadIdx.x;
  const int i = gtid / CH, c = gtid % CH;
  const bool act = i < p.M;
  const Pos* __restrict__ pos = (const Pos*)A.pos;
  const U* __restrict__ uu = (const U*)A.u;

  Pos own[3];
  R uo[3][3];
  R F[3][3], hh[3][3], Flj[3], Fs[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    own[a] = act ? pos[3 * i + a] : pos[0];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      F[a][k] = 0;
      hh[a][k] = 0;
    }
    Flj[a] = 0;
    Fs[a] = 0;
    if (ADJ) {
      U t = act ? uu[3 * i + a] : uu[0];
      uo[a][0] = t.x;
      uo[a][1] = t.y;
      uo[a][2] = t.z;
    }
  }
  const R rc2 = P::rc2(p);
  const R krf2 = (R)(2.0 * p.krf);
  const R kq[3] = {(R)p.kqq[0], (R)p.kqq[1], (R)p.kqq[2]};
  const R sig2 = (R)(p.sig * p.sig), eps24 = (R)(24.0 * p.eps);
  const R epss = (R)(144.0 * p.eps / p.sig);

  const int n = act ? A.ncnt[gtid] : 0;
  const int* __restrict__ L = A.nbr + (size_t)(gtid / RS_WAVE) * p.capc * RS_WAVE + (gtid & (RS_WAVE - 1));
  for (int k = 0; k < n; ++k) {
    const int j = L[(size_t)k * RS_WAVE];
    Pos pj[3];
    R uj[3][3];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      pj[b] = pos[3 * j + b];
      if (ADJ) {
        U t = uu[3 * j + b];
        uj[b][0] = t.x;
        uj[b][1] = t.y;
        uj[b][2] = t.z;
      }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        R d[3];
        P::diff(p, own[a], pj[b], d[0], d[1], d[2]);
        const R r2 = P::r2(d[0], d[1], d[2]);
        const bool in = (r2 <= rc2) && (r2 > (R)0);
        const R ri = P::rinv(in ? r2 : (R)1);
        const R ri2 = ri * ri, ri3 = ri2 * ri;
        const R q = kq[(a > 0) + (b > 0)];
        const R fr = in ? q * (ri3 - krf2) : (R)0;
#pragma unroll
        for (int t = 0; t < 3; ++t) F[a][t] -= fr * d[t];
        R du[3], dot = 0, g2 = 0;
        if (ADJ) {
#pragma unroll
          for (int t = 0; t < 3; ++t) {
            du[t] = uo[a][t] - uj[b][t];
            dot += d[t] * du[t];
          }
          g2 = in ? (R)(-3) * q * ri3 * ri2 : (R)0;
#pragma unroll
          for (int t = 0; t < 3; ++t) hh[a][t] += fr * du[t] + (g2 * dot) * d[t];
        }
        if (a == 0 && b == 0) {
          const R s2 = sig2 * ri2, s6 = s2 * s2 * s2, s12 = s6 * s6;
          const R frl = in ? eps24 * ri2 * ((R)2 * s12 - s6) : (R)0;
#pragma unroll
          for (int t = 0; t < 3; ++t) F[0][t] -= frl * d[t];
          if (ADJ) {
            const R g2l = in ? -eps24 * ri2 * ri2 * ((R)28 * s12 - (R)8 * s6) : (R)0;
            const R fs = in ? epss * ri2 * ((R)4 * s12 - s6) : (R)0;
#pragma unroll
            for (int t = 0; t < 3; ++t) {
              hh[0][t] += frl * du[t] + (g2l * dot) * d[t];
              Flj[t] -= frl * d[t];
              Fs[t] -= fs * d[t];
            }
          }
        }
      }
    }
  }

  // combine the CH lanes of the molecule (fixed xor tree, FP64)
  double Fd[9], hd[9], Fljd[3], Fsd[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      Fd[3 * a + t] = chunk_sum<CH>((double)F[a][t]);
      if (ADJ) hd[3 * a + t] = chunk_sum<CH>((double)hh[a][t]);
    }
    if (ADJ) {
      Fljd[a] = chunk_sum<CH>((double)Flj[a]);
      Fsd[a] = chunk_sum<CH>((double)Fs[a]);
    }
  }

  __shared__ double vjs[RS_BLOCK / CH][6];
  if (c == 0) {
    double vj[6] = {0, 0, 0, 0, 0, 0};
    if (act) {
      const int m = A.order[i];
      const double* xm = A.x + 9 * (size_t)m;
      double um[9], wm[9];
      if (ADJ) {
#pragma unroll
        for (int t = 0; t < 9; ++t) {
          const int ty = t < 3 ? 0 : 1;
          um[t] = A.csrc[9 * (size_t)m + t] * A.cscale[ty];
          wm[t] = A.wsrc[9 * (size_t)m + t] * A.wscale[ty];
        }
        // pair-part VJPs on this molecule's atoms: F_RF = F_pair - F_LJ (LJ acts on O only)
        double sq = 0.0, se = 0.0, ss = 0.0;
#pragma unroll
        for (int t = 0; t < 9; ++t) sq += wm[t] * (Fd[t] - (t < 3 ? Fljd[t] : 0.0));
#pragma unroll
        for (int t = 0; t < 3; ++t) {
          se += wm[t] * Fljd[t];
          ss += wm[t] * Fsd[t];
        }
        vj[0] = ss;
        vj[1] = p.eps != 0.0 ? se / p.eps : 0.0;
        vj[2] = 2.0 * sq / p.qH;
      }
      bonded<ADJ>(p, xm, um, wm, Fd, hd, vj, A.flags);
      // staleness of the list: O displacement since the build
      const double4 ob = A.O64[i];
      double dx = mic1(xm[0] - ob.x, p.L[0], p.invL[0]);
      double dy = mic1(xm[1] - ob.y, p.L[1], p.invL[1]);
      double dz = mic1(xm[2] - ob.z, p.L[2], p.invL[2]);
      if (r2_of(dx, dy, dz) > p.skin_half2) atomicOr(&A.flags[FLAG_STALE], 1);
      bool fin = true;
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        A.F[9 * (size_t)m + t] = Fd[t];
        fin = fin && isfinite(Fd[t]);
        if (ADJ) A.h[9 * (size_t)m + t] = hd[t];
      }
      if (!fin) atomicOr(&A.flags[FLAG_NONFINITE], 1);
    }
    if (ADJ) {
#pragma unroll
      for (int t = 0; t < 6; ++t) vjs[threadIdx.x / CH][t] = vj[t];
    }
  }
  if (ADJ) {
    __syncthreads();
    if (threadIdx.x < 6) {
      double s = 0.0;
      for (int q = 0; q < RS_BLOCK / CH; ++q) s += vjs[q][threadIdx.x];
      A.slots[(size_t)blockIdx.x * 8 + threadIdx.x] = s;
    }
  }
}

template <class P, bool ADJ>
void dispatch(const Launch& L, const PairArgs& a) {
  int g = pair_blocks(L.p);
  switch (L.p.CH) {
    case 8: k_pair<P, 8, ADJ><<<g, RS_BLOCK, 0, L.s>>>(a); break;
    default: k_pair<P, 16, ADJ><<<g, RS_BLOCK, 0, L.s>>>(a); break;
  }
}

}  // namespace

void launch_pair(const Launch& L, bool adjoint, const double* csrc, double cscale_O, double cscale_H,
                 const double* wsrc, double wscale_O, double wscale_H) {
  PairArgs a;
  a.p = L.p;
  a.nbr = L.b.nbr;
  a.ncnt = L.b.ncnt;
  a.order = L.b.order;
  a.x = L.b.x;
  a.csrc = csrc;
  a.wsrc = wsrc;
  a.cscale[0] = cscale_O;
  a.cscale[1] = cscale_H;
  a.wscale[0] = wscale_O;
  a.wscale[1] = wscale_H;
  a.O64 = L.b.O64;
  a.F = L.b.F;
  a.h = L.b.h;
  a.slots = L.b.slots;
  a.flags = L.b.flags;
  if (L.precision == RS_PRECISION_FP32) {
    a.pos = L.b.posq;
    a.u = L.b.uf;
    if (adjoint) dispatch<P32, true>(L, a);
    else dispatch<P32, false>(L, a);
  } else {
    a.pos = L.b.pos64;
    a.u = L.b.ud;
    if (adjoint) dispatch<P64, true>(L, a);
    else dispatch<P64, false>(L, a);
  }
}
