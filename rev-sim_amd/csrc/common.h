// common.h — device-side definitions shared by the librevsim kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/revsim.h"

#define RS_WAVE 64
#define RS_BLOCK 256

// Error/diagnostic flag slots (device int32 array, checked by the host at the end of a call).
enum {
  FLAG_LIST_OVERFLOW = 0,  // a molecule's sub-list exceeded its capacity
  FLAG_STALE = 1,          // an O moved more than skin/2 since the last list build
  FLAG_NONFINITE = 2,      // NaN/Inf in positions, velocities or forces
  FLAG_HEXTENT = 3,        // an O-H distance exceeded the extent assumed by the list radius
  FLAG_MLP_RANGE = 4,      // an O-O pair closer than the MLP table's lower end
  FLAG_CELL_OVERFLOW = 5,  // a home cell of the list grid held more molecules than a build block stages
  FLAG_FIXED_RANGE = 6,    // an FP64 pair-force term beyond the fixed-point sums' exact range (2^17 kJ/mol/nm)
  FLAG_COUNT = 8
};

// Everything a kernel needs about the system, passed by value.
struct SysParams {
  int M, N;
  double L[3], invL[3], qscale[3];
  float unitf[3];
  double rc2, krf, kc;
  double crf;      // reaction-field energy shift (1/r_c) 3ε_rf/(2ε_rf + 1): pair energy zero at r_c
  float rc2f, krff;
  float rc2q;      // FP32 pair rule threshold in quantized units: (float)(rc*rc*qscale*qscale)
  // θ-derived
  double sig, eps, qH, qO, kb, r0, ka, theta0;
  double kqq[3];   // k*qO*qO, k*qO*qH, k*qH*qH
  float kqqf[3];
  float sig2f, eps24f;
  double mO, mH;
  // list
  double rlist2, skin_half2, hext;
  // distance shells of the list (O-O distance at the build): shell 0 up to r_c + skin may hold
  // all nine atom pairs, shell 1 up to r_c + h_ext + skin no O-O pair, shell 2 (up to r_list) only
  // H-H pairs. Upper bounds of r^2 with a 1e-5 relative margin, FP64 and in quantized units.
  double sh2[2];
  float sh2q[2];
  int nc[3], ncell, brute;
  int cs;          // cell stencil half-width: cells are >= r_list / cs wide, (2 cs + 1)^3 stencil
  float rlo2f, rhi2f;  // FP32 screen of the list rule: accept below, reject above, else FP64
  float rlo2q, rhi2q;  // the same band in quantized units (cubic boxes)
  int CH, capc;
  // molecule list entries: 16-bit words (at most 65,536 molecules) or 32-bit ints
  int list16;
};

// ----------------------------------------------------------------------------------------------
// Pair membership. The two rules below are part of the library's definition of "which pairs
// interact" and are evaluated exactly as written (no contraction) so that the CPU oracle,
// compiled with -ffp-contract=off, makes identical decisions.
//   FP64: dx = x_b - x_a; dx -= L*rint(dx*(1/L)); r2 = fma(dz,dz,fma(dy,dy,dx*dx))
//   FP32: dxi = (int32)(q_b - q_a) on uint32 positions q = llrint(x*2^32/L) (cubic box);
//         dxq = (float)dxi; r2q = fmaf(dzq,dzq,fmaf(dyq,dyq,dxq*dxq)) <= rc2q, in quantized units
//         (one length unit = L/2^32), rc2q = (float)(rc*rc*qscale*qscale). The pair kernels work
//         in these units throughout and fold L/2^32 into their constants. (float)(-dxi) is
//         -(float)dxi, so both sides of a pair decide alike.
//   The RDF bins and the list build's FP32 screen use physical FP32 differences, qdiff() below.
// ----------------------------------------------------------------------------------------------
__device__ __forceinline__ double mic1(double d, double L, double invL) {
#pragma clang fp contract(off)
  return d - L * rint(d * invL);
}

// a + b evaluated as written: instructions made in this function carry no contraction flag, so the
// compiler cannot fuse the add with a multiply that feeds it (a plain a + b in a kernel compiled
// with contraction may become an fma, which rounds differently).
__device__ __forceinline__ double add_nc(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}

__device__ __forceinline__ double r2_of(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  return fma(dz, dz, fma(dy, dy, dx * dx));
}

__device__ __forceinline__ float r2f_of(float dx, float dy, float dz) {
#pragma clang fp contract(off)
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

__device__ __forceinline__ float qdiff(uint32_t qa, uint32_t qb, float unit) {
#pragma clang fp contract(off)
  return (float)(int32_t)(qb - qa) * unit;
}

__device__ __forceinline__ float qdiffq(uint32_t qa, uint32_t qb) { return (float)(int32_t)(qb - qa); }

__device__ __forceinline__ uint32_t quant1(double x, double s) {
#pragma clang fp contract(off)
  return (uint32_t)(int64_t)rint(x * s);
}

__device__ __forceinline__ double wrap1(double x, double L) {
#pragma clang fp contract(off)
  return x - floor(x / L) * L;
}

// Cell of a molecule from its O position (cells of side L/nc along each axis).
__device__ __forceinline__ int cell_index(const SysParams& p, double x, double y, double z) {
  int c[3];
  double w[3] = {x, y, z};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    int ck = (int)(wrap1(w[k], p.L[k]) * p.invL[k] * p.nc[k]);
    c[k] = ck < 0 ? 0 : (ck >= p.nc[k] ? p.nc[k] - 1 : ck);
  }
  return (c[2] * p.nc[1] + c[1]) * p.nc[0] + c[0];
}

// First stage of a list build: molecule m's cell, counted in the per-cell totals (which the
// build's scan clears again).
__device__ __forceinline__ void file_molecule(const SysParams& p, int* cell, int* count, int m,
                                              double x, double y, double z) {
  const int c = p.brute ? 0 : cell_index(p, x, y, z);
  cell[m] = c;
  atomicAdd(&count[c], 1);
}

// ----------------------------------------------------------------------------------------------
// exp(t) for t <= 0: Cody-Waite reduction t = k ln2 + f with fdlibm's two-part ln2, Taylor
// polynomial in f by fused multiply-adds, exact scaling by 2^k. Used for the RDF smoothing
// weights so that histogram contributions are reproducible bit for bit on the host.
// ----------------------------------------------------------------------------------------------
__device__ __forceinline__ double rs_exp(double t) {
#pragma clang fp contract(off)
  double k = rint(t * 1.4426950408889634);
  double f = fma(-k, 6.93147180369123816490e-01, t);
  f = fma(-k, 1.90821492927058770002e-10, f);
  double p = 1.0 / 6227020800.0;
  p = fma(p, f, 1.0 / 479001600.0);
  p = fma(p, f, 1.0 / 39916800.0);
  p = fma(p, f, 1.0 / 3628800.0);
  p = fma(p, f, 1.0 / 362880.0);
  p = fma(p, f, 1.0 / 40320.0);
  p = fma(p, f, 1.0 / 5040.0);
  p = fma(p, f, 1.0 / 720.0);
  p = fma(p, f, 1.0 / 120.0);
  p = fma(p, f, 1.0 / 24.0);
  p = fma(p, f, 1.0 / 6.0);
  p = fma(p, f, 1.0 / 2.0);
  p = fma(p, f, 1.0);
  p = fma(p, f, 1.0);
  return ldexp(p, (int)k);
}

// ----------------------------------------------------------------------------------------------
// log, sin/cos of 2πu and acos from IEEE operations only (+, -, *, /, sqrt, fma, rint): fixed
// polynomials evaluated by fused multiply-adds in a fixed order, so the CPU oracle, which
// evaluates the same sequences, gets the same bits. They feed the Langevin noise (Box-Muller)
// and the angle force, i.e. the trajectory itself.
//   log(x), x > 0 normal: x = m 2^k with m in (1/√2, √2], log m = 2 atanh(s), s = (m-1)/(m+1),
//     |s| <= 0.1716: the atanh series to s^25, k ln2 with fdlibm's two-part ln2.
//   sincos(2πu): 4u = q + r with q = rint(4u), |r| <= 1/2 (both exact), θ = r π/2: Taylor series
//     of sin θ to θ^19 and cos θ to θ^20 (|θ| <= π/4), then the quadrant q.
//   acos(x): π/2 - asin x for |x| <= 1/2, 2 asin √((1-x)/2) above, π - 2 asin √((1+x)/2) below;
//     asin y = y Σ_n C(2n,n)/(4^n (2n+1)) y^(2n), to n = 25 (|y| <= 1/2).
// Accuracy: a few units in the last place (tests compare them with the C library's).
// ----------------------------------------------------------------------------------------------
__device__ __forceinline__ double rs_log(double x) {
#pragma clang fp contract(off)
  const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
  int k = (int)((bits >> 52) & 0x7ffULL) - 1023;
  double m = __longlong_as_double((long long)((bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
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

__device__ __forceinline__ void rs_sincos2pi(double u, double* sn, double* cs) {
#pragma clang fp contract(off)
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

__device__ __forceinline__ double rs_asin_series(double y) {
#pragma clang fp contract(off)
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

__device__ __forceinline__ double rs_acos(double x) {
#pragma clang fp contract(off)
  const double pio2_hi = 1.57079632679489655800e+00, pio2_lo = 6.12323399573676603587e-17;
  if (x > 0.5) return 2.0 * rs_asin_series(sqrt(0.5 * (1.0 - x)));
  if (x < -0.5) return 2.0 * pio2_hi - (2.0 * rs_asin_series(sqrt(0.5 * (1.0 + x))) - 2.0 * pio2_lo);
  return pio2_hi - (rs_asin_series(x) - pio2_lo);
}

// rs_rsqrt(x) = 1/√x for positive normal x from integer and IEEE operations only: the seed
// C - (bits >> 1) halves and negates the exponent (C = 0x5FE6EB50C7B537A9, Lomont's and
// Robertson's 64-bit constant for this seed; relative error below 3.5 %), then four Newton steps
// y ← y (3/2 - (x/2) y²), each about squaring the relative error (3.5e-2 → 1.8e-3 → 4.8e-6 →
// 3.5e-11 → below rounding): within 2 units in the last place. The FP64 pair terms use it for
// 1/r: 16 full-rate instructions against 29 for 1/sqrt(x) (two correctly rounded operations with
// their range scaling and two quarter-rate transcendentals). x = 0 gives NaN (y² overflows to
// infinity at the fourth step, and 0 · ∞ is NaN).
__device__ __forceinline__ double rs_rsqrt(double x) {
#pragma clang fp contract(off)
  const double hx = 0.5 * x;
  double y = __longlong_as_double(0x5FE6EB50C7B537A9LL - (__double_as_longlong(x) >> 1));
#pragma unroll
  for (int k = 0; k < 4; ++k) y = y * fma(-hx, y * y, 1.5);
  return y;
}

// ----------------------------------------------------------------------------------------------
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11), the generator numpy exposes as
// np.random.Philox. Counter (step, atom, 0, 0), key (seed, 0); four 64-bit outputs give the three
// normals of one atom by Box-Muller on 53-bit uniforms (rs_log, rs_sincos2pi above).
// ----------------------------------------------------------------------------------------------
__device__ __forceinline__ void philox4x64(uint64_t c[4], uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint64_t hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
    uint64_t hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
    uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

__device__ __forceinline__ void normals3(uint64_t seed, int step, int atom, double n[3]) {
  uint64_t c[4] = {(uint64_t)(uint32_t)step, (uint64_t)atom, 0, 0};
  philox4x64(c, seed, 0);
  const double two53 = 1.0 / 9007199254740992.0;
  double u1 = (double)((c[0] >> 11) + 1) * two53, u2 = (double)(c[1] >> 11) * two53;
  double u3 = (double)((c[2] >> 11) + 1) * two53, u4 = (double)(c[3] >> 11) * two53;
  const double rad1 = sqrt(-2.0 * rs_log(u1)), rad2 = sqrt(-2.0 * rs_log(u3));
  double s2, c2, s4, c4;
  rs_sincos2pi(u2, &s2, &c2);
  rs_sincos2pi(u4, &s4, &c4);
  n[0] = rad1 * c2;
  n[1] = rad1 * s2;
  n[2] = rad2 * c4;
}

// Word offset of atom a (0 = O) in a molecule's 20-word adjoint pair record: H1 at 0, H2 at 6,
// O at 12, each as 3 position words then 3 accum_C/m words, so the H-H shell reads 3 aligned
// 16-byte words and the others 5.
__device__ __forceinline__ int pair_rec(int a) { return a == 0 ? 12 : 6 * (a - 1); }
__device__ __forceinline__ void write_pair_rec(float4* pu, int mol, int a, const int q[3], const double u[3]) {
  float* r = (float*)(pu + 5 * (size_t)mol) + pair_rec(a);
  for (int k = 0; k < 3; ++k) {
    r[k] = __int_as_float(q[k]);
    r[3 + k] = (float)u[k];
  }
}

// Three consecutive 32-bit values (one atom of the packed FP32 pair inputs: 9 values per
// molecule) as a 4-vector; the fourth component is unused.
__device__ __forceinline__ int4 ld3i(const int* p) { return make_int4(p[0], p[1], p[2], 0); }
__device__ __forceinline__ float4 ld3f(const float* p) { return make_float4(p[0], p[1], p[2], 0.f); }

// The molecule list, ELL by wave and shell: lane gtid = i*CH + c holds sub-list c of molecule i in
// each of the three distance shells; entry k of shell s is at
//   nbr[list_base(p, gtid, s) + k*64]
// and ncnt[gtid] holds the three sub-list lengths (x, y, z; w unused).
constexpr int LIST_SHELLS = 3;
// Entry type of the molecule list: sorted molecule indices, 16-bit where they fit (L16), which
// halves the list stream of the pair kernels and the list build's stores.
template <bool L16> struct ListEntry { typedef int type; };
template <> struct ListEntry<true> { typedef unsigned short type; };
__device__ __host__ __forceinline__ size_t list_base(const SysParams& p, size_t gtid, int s) {
  return (((gtid / RS_WAVE) * LIST_SHELLS + s) * (size_t)p.capc) * RS_WAVE + (gtid & (RS_WAVE - 1));
}

// XCD-aware block order. The dispatcher deals consecutive workgroups round-robin over MI355X's 8
// XCDs, each with its own L2; remapping workgroup b to logical block (b mod 8)*(nb/8) + b/8 gives
// each XCD one contiguous eighth of the logical blocks, so blocks over neighbouring molecules (the
// cell-sorted order) share an L2. A bijection when nb is a multiple of 8; otherwise the identity.
constexpr int RS_XCDS = 8;
__device__ __forceinline__ int xcd_block(int b, int nb) {
  return nb % RS_XCDS == 0 ? (b % RS_XCDS) * (nb / RS_XCDS) + b / RS_XCDS : b;
}

// Wave-level reductions over lanes that share a molecule (CH consecutive lanes).
template <int CH>
__device__ __forceinline__ float chunk_sum(float v) {
#pragma unroll
  for (int o = CH / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, RS_WAVE);
  return v;
}
template <int CH>
__device__ __forceinline__ double chunk_sum(double v) {
#pragma unroll
  for (int o = CH / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, RS_WAVE);
  return v;
}
// Sum over the CH = 8 or 16 lanes of a molecule (consecutive, aligned) by DPP row
// permutations, one VALU instruction per step and no LDS round trip: xor 1 and xor 2 inside
// quads, then the mirrored half-row (lane i with 7 - i) and, for 16 lanes, the mirrored row
// (i with 15 - i). Each lane adds its partner's partial to its own; a + b = b + a exactly, so
// every lane of the group ends with the same sum, in a fixed order (deterministic).
// 64-bit values (FP64, integer sums) move as two 32-bit words; their adds are exact (integers) or
// commutative (FP64) as well.
template <int CTRL, class T>
__device__ __forceinline__ T dpp_move(T v) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
  } else {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(T, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
  }
}
template <int CH, class T>
__device__ __forceinline__ T chunk_sum_dpp(T v) {
  static_assert(CH == 8 || CH == 16, "DPP lane sums cover 8 or 16 lanes");
  v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_move<0x141>(v);  // row_half_mirror
  if constexpr (CH == 16) v += dpp_move<0x140>(v);  // row_mirror
  return v;
}
template <int CH>
__device__ __forceinline__ unsigned long long chunk_sum(unsigned long long v) {
#pragma unroll
  for (int o = CH / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, RS_WAVE);
  return v;
}
