// bonded.h — intramolecular terms of one flexible water, FP64: harmonic O-H bonds and the H-O-H
// angle (calc_F! water/water.jl:915-968; calc_F_and_grads! :1279-1422).
//
// Forces, and in the adjoint variant the Hessian-vector product (dF/dx)^T u and the θ-VJP terms
// for (k_bond, r0, k_angle). The angle's HVP is the tangent (directional derivative) of its force
// along u: (dF/dx) u = (dF/dx)^T u because dF/dx is minus the symmetric Hessian of the energy. Also checks the O-H extent assumed by the molecule-list radius and the O
// displacement since the last list build.
#pragma once

#include "common.h"

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double c[3]) {
#pragma clang fp contract(off)
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// xm: the molecule's 9 coordinates (O, H1, H2); um = accum_C/m and wm = accum_A dt^2/(2m) for
// the adjoint variant. F and h are overwritten; vj[0..2] receive the (k_bond, r0, k_angle) VJPs.
//
// The angle force is the reference's: θ = acos(a.b / (|a||b|)) with a = H1 - O, b = H2 - O, and
// F_H1 = k_a (θ0 - θ) p1, F_H2 = k_a (θ0 - θ) p3, F_O = -(F_H1 + F_H2), p1 = (a x c)/(|a x c||a|),
// p3 = (c x b)/(|c x b||b|), c = a x b. Since a and b are perpendicular to c, |a x c| = |a||c| and
// |c x b| = |b||c|, so p1 = (a x c)/(|a|^2 |c|) and p3 = (c x b)/(|b|^2 |c|); and
// dθ/d(cos θ) = -1/sin θ = -|a||b|/|c|. Written so, the value needs two divisions (1/(|a|^2|b|^2)
// and 1/|c|) and two square roots, and its tangent along (u_H1 - u_O, u_H2 - u_O) needs none
// more: FP64 divisions are long dependent sequences, and these kernels run few waves per SIMD.
//
// Evaluated without contraction, as written, with the library's own acos (common.h): the CPU
// oracle evaluates the forces by the same sequence, so they agree bit for bit.
template <bool ADJ>
__device__ void bonded_mol(const SysParams& p, const double* xm, const double* um, const double* wm,
                           double F[9], double h[9], double vj[3], int* flags, double4 o_build) {
#pragma clang fp contract(off)
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
  double a[3], b[3];  // H1 - O, H2 - O (minimum image)
  for (int k = 0; k < 3; ++k) {
    a[k] = mic1(xm[3 + k] - O[k], p.L[k], p.invL[k]);
    b[k] = mic1(xm[6 + k] - O[k], p.L[k], p.invL[k]);
  }
  for (int hh = 1; hh <= 2; ++hh) {
    const double* d = hh == 1 ? a : b;
    const double r = sqrt(r2_of(d[0], d[1], d[2]));
    if (r > p.hext) atomicOr(&flags[FLAG_HEXTENT], 1);
    const double ir = 1.0 / r;
    const double n[3] = {d[0] * ir, d[1] * ir, d[2] * ir};
    const double disp = r - p.r0;
    const double Fv = -p.kb * disp;
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
      const double Fr = Fv * ir, Fp = -p.kb;
      for (int k = 0; k < 3; ++k) {
        const double hv = Fr * du[k] + (Fp - Fr) * dot * n[k];
        h[k] += hv;
        h[3 * hh + k] -= hv;
        const double dw = wm[k] - wm[3 * hh + k];
        vj[0] += disp * n[k] * dw;
        vj[1] += -p.kb * n[k] * dw;
      }
    }
  }
  // angle, i = H1, j = O (centre), k = H2; primed quantities are tangents along
  // (a', b') = (u_H1 - u_O, u_H2 - u_O)
  double ad[3] = {0, 0, 0}, bd[3] = {0, 0, 0};
  if (ADJ)
    for (int k = 0; k < 3; ++k) {
      ad[k] = um[3 + k] - um[k];
      bd[k] = um[6 + k] - um[k];
    }
  double c[3], cd[3];
  cross3(a, b, c);
  if (c[0] == 0.0 && c[1] == 0.0 && c[2] == 0.0) return;
  const double aa = dot3(a, a), bb = dot3(b, b), ab = dot3(a, b), cc = dot3(c, c);
  const double P = aa * bb;
  const double iP = 1.0 / P;           // 1/(|a|^2 |b|^2)
  const double rab = sqrt(P);          // |a||b|
  const double irab = rab * iP;        // 1/(|a||b|)
  const double rc = sqrt(cc), irc = 1.0 / rc;
  const double cs = ab * irab;
  const double th = rs_acos(cs);
  const double disp = p.theta0 - th;
  const double Fa = p.ka * disp;
  const double s1 = bb * iP * irc, s3 = aa * iP * irc;  // 1/(|a|^2|c|), 1/(|b|^2|c|)
  double p1u[3], p3u[3];
  cross3(a, c, p1u);
  cross3(c, b, p3u);
  const double c1 = Fa * s1, c3 = Fa * s3;
  double c1d = 0.0, c3d = 0.0, p1ud[3] = {0, 0, 0}, p3ud[3] = {0, 0, 0};
  if (ADJ) {
    double t1[3], t2[3];
    cross3(ad, b, t1);
    cross3(a, bd, t2);
    for (int k = 0; k < 3; ++k) cd[k] = t1[k] + t2[k];
    const double aad = 2.0 * dot3(a, ad), bbd = 2.0 * dot3(b, bd);
    const double abd = dot3(ad, b) + dot3(a, bd), ccd = 2.0 * dot3(c, cd);
    const double Pd = aad * bb + aa * bbd;
    const double iPd = -Pd * iP * iP;
    const double rabd = 0.5 * Pd * irab;
    const double irabd = rabd * iP + rab * iPd;
    const double csd = abd * irab + ab * irabd;
    const double thd = -csd * rab * irc;  // dθ = -d(cos θ) / sin θ, sin θ = |c|/(|a||b|)
    const double rcd = 0.5 * ccd * irc;
    const double ircd = -rcd * irc * irc;
    const double s1d = (bbd * iP + bb * iPd) * irc + bb * iP * ircd;
    const double s3d = (aad * iP + aa * iPd) * irc + aa * iP * ircd;
    const double Fad = -p.ka * thd;
    c1d = Fad * s1 + Fa * s1d;
    c3d = Fad * s3 + Fa * s3d;
    cross3(ad, c, t1);
    cross3(a, cd, t2);
    for (int k = 0; k < 3; ++k) p1ud[k] = t1[k] + t2[k];
    cross3(cd, b, t1);
    cross3(c, bd, t2);
    for (int k = 0; k < 3; ++k) p3ud[k] = t1[k] + t2[k];
  }
  for (int k = 0; k < 3; ++k) {
    const double gi = c1 * p1u[k], gk = c3 * p3u[k];
    F[3 + k] += gi;
    F[6 + k] += gk;
    F[k] -= gi + gk;
    if (ADJ) {
      const double gid = c1d * p1u[k] + c1 * p1ud[k], gkd = c3d * p3u[k] + c3 * p3ud[k];
      h[3 + k] += gid;
      h[6 + k] += gkd;
      h[k] -= gid + gkd;
      const double p1d = s1 * p1u[k], p3d = s3 * p3u[k];
      vj[2] += disp * (p1d * wm[3 + k] + p3d * wm[6 + k] - (p1d + p3d) * wm[k]);
    }
  }
}
