// bonded.h — intramolecular terms of one flexible water, FP64: harmonic O-H bonds and the H-O-H
// angle (calc_F! water/water.jl:915-968; calc_F_and_grads! :1279-1422).
//
// Forces, and in the adjoint variant the Hessian-vector product (dF/dx)^T u and the θ-VJP terms
// for (k_bond, r0, k_angle). The angle's HVP is the tangent part of its force evaluated on dual
// numbers seeded with u: (dF/dx) u = (dF/dx)^T u because dF/dx is minus the symmetric Hessian of
// the energy. Also checks the O-H extent assumed by the molecule-list radius and the O
// displacement since the last list build.
#pragma once

#include "common.h"

namespace bonded_detail {

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

}  // namespace bonded_detail

// xm: the molecule's 9 coordinates (O, H1, H2); um = accum_C/m and wm = accum_A dt^2/(2m) for
// the adjoint variant. F and h are overwritten; vj[0..2] receive the (k_bond, r0, k_angle) VJPs.
template <bool ADJ>
__device__ void bonded_mol(const SysParams& p, const double* xm, const double* um, const double* wm,
                           double F[9], double h[9], double vj[3], int* flags, double4 o_build) {
  using namespace bonded_detail;
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    F[t] = 0.0;
    h[t] = 0.0;
  }
  vj[0] = vj[1] = vj[2] = 0.0;
  const double O[3] = {xm[0], xm[1], xm[2]};
  {
    // staleness of the molecule list: O displacement since the build
    double dx = mic1(O[0] - o_build.x, p.L[0], p.invL[0]);
    double dy = mic1(O[1] - o_build.y, p.L[1], p.invL[1]);
    double dz = mic1(O[2] - o_build.z, p.L[2], p.invL[2]);
    if (r2_of(dx, dy, dz) > p.skin_half2) atomicOr(&flags[FLAG_STALE], 1);
  }
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
        vj[0] += disp * n[k] * dw;
        vj[1] += -p.kb * n[k] * dw;
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
      vj[2] += disp.v * (p1d.v * wm[3 + k] + p3d.v * wm[6 + k] - (p1d.v + p3d.v) * wm[k]);
    }
  }
}
