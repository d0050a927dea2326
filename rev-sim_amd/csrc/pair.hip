// pair.hip — the fused pair force / adjoint kernel: reaction-field Coulomb + O-O Lennard-Jones over
// the molecule list, forces, Hessian-vector product and θ-VJP in one pass.
//
// Forward variant (calc_F!, water/water.jl:877-975): F(x). The intramolecular terms (FP64, see
// bonded.h) seed F (and h and the bonded θ-VJP in the adjoint variant) at the same positions:
// evaluated by the preceding integrator kernel in the adjoint, by the previous launch's step
// epilogue in the forward loop, whose kernel ends with the Langevin step (STEP variant). This
// kernel adds the pair terms.
// Adjoint variant (calc_F_and_grads!, water/water.jl:1212-1433): F(x), the Hessian-vector product
//   h = (dF/dx)^T u with u = accum_C / m, and the θ-VJP sum_dof w_dof dF_dof/dθ with
//   w = accum_A dt^2/(2m) (water/water.jl:1517-1519), θ = (σ_O, ε_O, q_H, k_b, r0, k_a).
//
// Work decomposition: CH consecutive lanes own one molecule i (sorted order); lane c walks
// sub-list c of i's full molecule list, evaluates the atom pairs of each listed molecule j and
// accumulates only i's side. The list's three distance shells are walked in turn, each in its own
// loop with the atom pairs that can be within r_c there: all nine (shell 0, with the O-O
// Lennard-Jones and MLP terms), the eight without O-O (shell 1), the four H-H (shell 2). A pair
// left out would contribute exact zeros, so the sums are the same as over all nine; within a loop,
// atom pairs outside r_c are branched over. The CH partial sums are combined by DPP row steps in
// a fixed order (chunk_sum_dpp, common.h), so the result is deterministic without atomics. Lane
// c == 0 then adds them to the molecule's outputs in the caller's atom order.
//
// Pair arithmetic: FP32 on quantized positions (the int32 difference of 2^32-periodic positions is
// the minimum image by construction) or FP64 (parity mode). The FP32 path computes in the quantized
// length unit lu = L/2^32: d is (float) of the int32 difference, with no scaling, and lu is folded
// into the per-type constants of PairArgs, so F comes out in physical units and h in units of
// 1/lu (rescaled once per molecule in FP64). The FP64 path uses the same code with lu = 1 nm.
// Per pair of atoms a (own), b:
//   f/r  = k q_a q_b (1/r^3 - 2 k_rf)              [+ 24ε/r^2 (2 s^12 - s^6) for O-O, s = σ/r]
//   F_a -= (f/r) d, d = x_b - x_a
//   h_a += K (u_a - u_b), K = f' r̂r̂ᵀ + (f/r)(I - r̂r̂ᵀ), i.e. (f/r) Δu + ((f' - f/r)/r^2)(d.Δu) d
//          with (f' - f/r)/r^2 = -3 k q_a q_b / r^5   [- 24ε/r^4 (28 s^12 - 8 s^6) for O-O]
//   q-VJP: dF/dq_H = 2 F_RF / q_H; ε-VJP: dF/dε = F_LJ/ε; σ-VJP: dF_a/dσ = -(144ε/(σ r^2))(4s^12-s^6) d
#include "internal.h"
#include "bonded.h"
#include "integrator.h"
#include "mlp.h"

namespace {

struct PairArgs {
  SysParams p;
  const void* __restrict__ nbr;  // ListEntry<L16> entries
  const int4* __restrict__ ncnt;
  const int* __restrict__ order;
  const void* __restrict__ pos;
  const void* __restrict__ u;
  const float4* __restrict__ pu;  // FP32 adjoint: per-molecule records (write_pair_rec)
  const double* __restrict__ wsrc;
  double wscale[2];
  const double* __restrict__ vjb;
  double* F;
  double* h;
  double* slots;
  int* flags;
  // MLP pair term on O-O pairs (mlp.h)
  int mlp;
  const void* __restrict__ mtab;
  double mrlo, mh, minvh;
  double* mslot;                    // per-block table gradient (adjoint)
  const double* __restrict__ wmax;  // this step's 64 max|w_O| slots (uint64 bits of doubles)
  // constants in the kernel's length unit lu (L/2^32 in FP32, 1 nm in FP64), per pair type
  // t = (O-O, O-H, H-H):
  //   kA = k q q / lu^2, kB = 2 k q q k_rf lu       f/r·lu      = kA/r'^3 - kB
  //   k3A = -3 k q q / lu^2                          ((f'-f/r)/r^2)·lu^3 = k3A / r'^5
  //   sig2A = σ^2/lu^2, eps24A = 24ε/lu, epssA = 144ε/(σ lu), rc2 in lu^2, lu, hscale = 1/lu
  double kA[3], kB[3], k3A[3], sig2A, eps24A, epssA, rc2, lu, hscale;
};

// Fixed-point scale of the per-block table-gradient sums: 2^(50 - e) with max|w_O| < 2^e, so one
// block's sum (≤ 2^11 pair contributions bounded by max|w_O|) stays below 2^61, and each
// contribution (|w_O . r̂| times a Hermite weight ≤ 1) below sqrt(3) 2^50 < 2^51.
// Called by every lane of a wave (uniform branch): the wave reduces the 64 slots.
__device__ __forceinline__ double mlp_scale(const double* wmax) {
  double m = wmax[threadIdx.x & 63];
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (!(m > 0.0)) return 1125899906842624.0;  // 2^50
  int e;
  frexp(m, &e);
  return ldexp(1.0, 50 - e);
}

// fix_rn(v) = round-to-nearest-even of v as a 64-bit integer for |v| < 2^51: adding 1.5 * 2^52
// leaves the rounded integer in the low mantissa bits (the exponent stays 52), three
// instructions against the general conversion's dozen.
__device__ __forceinline__ unsigned long long fix_rn(double v) {
  return (unsigned long long)(__double_as_longlong(v + 6755399441055744.0) - 0x4338000000000000LL);
}

// FP64 pair forces are summed as 64-bit integers in units of 2^-33 kJ/mol/nm, so a molecule's
// force does not depend on the order its pair terms arrive in (lanes, list order, list layout) and
// equals, bit for bit, what the oracle's sum in any order gives. A term c = (f/r) d_t (one IEEE
// product) enters as round(c 2^33). The kernel scales f/r by 2^33 once per pair (exact), so
// (f/r 2^33) d_t is c 2^33 with c's rounding; adding 1.5 2^52 (not fused) rounds it to an integer
// and leaves that integer plus the bias 0x4338000000000000 in the bit pattern (exact for
// |c 2^33| < 2^51; the kernel flags larger terms), so a lane adds the bit patterns and removes the
// bias once per term at the end.
constexpr double FX_SCALE = 8589934592.0;        // 2^33
constexpr double FX_INV = 1.0 / 8589934592.0;
// Range of a term: the bias trick is exact for |c 2^33| < 2^51. Each term is bounded by the pair's
// |f| = |f/r| r (|d_t| <= r), checked as |f/r 2^33| <= 2^50 / r: one multiply by 1/r, which
// the pair has, instead of forming r; the factor 2 covers 1/r's rounding. Out-of-range pairs
// (f/r and 1/r zeroed) pass as 0 <= 0; NaN fails.
constexpr double FX_CMAX2 = 1125899906842624.0;  // 2^50
constexpr unsigned long long FX_BIAS = 0x4338000000000000ULL;
__device__ __forceinline__ unsigned long long fx_term(double ftot2, double d) {
  return (unsigned long long)__double_as_longlong(add_nc(ftot2 * d, 6755399441055744.0));
}

struct P32 {
  typedef int4 Pos;
  typedef float4 U;
  typedef float R;
  // packed inputs: 9 ints / 9 floats per molecule
  static __device__ __forceinline__ Pos ld_pos(const void* p, int slot) { return ld3i((const int*)p + 3 * slot); }
  static __device__ __forceinline__ U ld_u(const void* p, int slot) { return ld3f((const float*)p + 3 * slot); }
  static __device__ __forceinline__ void diff(const SysParams& p, const int4& a, const int4& b,
                                              float& dx, float& dy, float& dz) {
    dx = qdiffq((uint32_t)a.x, (uint32_t)b.x);
    dy = qdiffq((uint32_t)a.y, (uint32_t)b.y);
    dz = qdiffq((uint32_t)a.z, (uint32_t)b.z);
  }
  static __device__ __forceinline__ float r2(float dx, float dy, float dz) { return r2f_of(dx, dy, dz); }
  static __device__ __forceinline__ float rinv(float r2) { return __builtin_amdgcn_rsqf(r2); }
  static __device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
};

struct P64 {
  typedef double4 Pos;
  typedef double4 U;
  typedef double R;
  static __device__ __forceinline__ Pos ld_pos(const void* p, int slot) { return ((const double4*)p)[slot]; }
  static __device__ __forceinline__ U ld_u(const void* p, int slot) { return ((const double4*)p)[slot]; }
  // mic1 with the multiply fused: the sorted positions lie in [0, L], so rint(d/L) is -1, 0 or 1,
  // L k is exact and fma(-L, k, d) rounds d - L k once, as mic1's separate multiply and subtract do
  static __device__ __forceinline__ double mic_f(double d, double L, double invL) {
    return __builtin_fma(-L, rint(d * invL), d);
  }
  static __device__ __forceinline__ void diff(const SysParams& p, const double4& a, const double4& b,
                                              double& dx, double& dy, double& dz) {
    dx = mic_f(add_nc(b.x, -a.x), p.L[0], p.invL[0]);
    dy = mic_f(add_nc(b.y, -a.y), p.L[1], p.invL[1]);
    dz = mic_f(add_nc(b.z, -a.z), p.L[2], p.invL[2]);
  }
  static __device__ __forceinline__ double r2(double dx, double dy, double dz) { return r2_of(dx, dy, dz); }
  static __device__ __forceinline__ double rinv(double r2) { return rs_rsqrt(r2); }
  static __device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }
};

// ------------------------------------------------------------------ the pair kernel

// Occupancy targets (waves per SIMD), the same for 16 and 8 lanes per molecule. FP32: the forward
// loop gathers one entry ahead: without the MLP term in 64 VGPRs, 8 waves per SIMD, so its 8,192
// waves at 98k atoms run in one round (2 % faster than 7 waves); with it in 72 VGPRs, 7 waves (at 8
// it spills inside the loops). The adjoint holds 100 VGPRs without the MLP term and 128 with
// it, no spill, 4 waves (at 5 waves the plain one spills 20 B per lane outside the loops: 0.3 %
// faster, 19 % more HBM traffic). FP64: the forward needs ~90 VGPRs (5 waves); the adjoint without the MLP term fits 168
// VGPRs, 3 waves (7 % faster than 2 waves), with it 210 VGPRs, 2 waves (at 3 it spills ~150 B).
// Both precisions run 16 lanes per molecule by default (at 8: FP32 1 % and FP64 2 % slower).
template <class P, int CH, bool ADJ, bool FMLP>
constexpr int pair_waves() {
  return sizeof(typename P::R) == 4 ? (ADJ ? 4 : (FMLP ? 7 : 8)) : (ADJ ? (FMLP ? 2 : 3) : 4);
}

// Whether the pair loop gathers the next entry's inputs one iteration ahead.
template <class P, int CH, bool ADJ>
constexpr bool pair_prefetch() { return sizeof(typename P::R) == 4 && !ADJ; }

// FMLP: the MLP term as a compile-time switch (the plain kernels then hold none of its state).
// STEP (forward only): the kernel ends with the integrator step of its molecules (St), so the
// forward loop runs one kernel per step. The new sorted positions go to the other buffer
// (St.posq_next / St.pos64_next): other blocks may still be reading this step's.
template <class P, int CH, bool ADJ, bool FMLP, bool STEP, bool L16>
__global__ __launch_bounds__(RS_BLOCK) __attribute__((amdgpu_waves_per_eu(pair_waves<P, CH, ADJ, FMLP>())))
void k_pair(PairArgs A, StepArgs St) {
  typedef typename ListEntry<L16>::type LT;
  typedef typename P::R R;
  typedef typename P::Pos Pos;
  typedef typename P::U U;
  const SysParams& p = A.p;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int gtid = blk * RS_BLOCK + threadIdx.x;
  const int i = gtid / CH, c = gtid % CH;
  const bool act = i < p.M;

  // the FP32 adjoint reads each molecule's positions and u as one record (write_pair_rec)
  constexpr bool REC = ADJ && sizeof(R) == 4;
  // molecule j's inputs, atoms b0..2 (b0 = 1: the H atoms only)
  auto gather = [&](int j, int b0, Pos* pp, R (*uu)[3]) {
    if constexpr (REC) {
      const float4* r = A.pu + 5 * (size_t)j;
      const float4 v0 = r[0], v1 = r[1], v2 = r[2];
      pp[1] = make_int4(__float_as_int(v0.x), __float_as_int(v0.y), __float_as_int(v0.z), 0);
      uu[1][0] = v0.w;
      uu[1][1] = v1.x;
      uu[1][2] = v1.y;
      pp[2] = make_int4(__float_as_int(v1.z), __float_as_int(v1.w), __float_as_int(v2.x), 0);
      uu[2][0] = v2.y;
      uu[2][1] = v2.z;
      uu[2][2] = v2.w;
      if (b0 == 0) {
        const float4 v3 = r[3], v4 = r[4];
        pp[0] = make_int4(__float_as_int(v3.x), __float_as_int(v3.y), __float_as_int(v3.z), 0);
        uu[0][0] = v3.w;
        uu[0][1] = v4.x;
        uu[0][2] = v4.y;
      }
    } else {
#pragma unroll
      for (int b = b0; b < 3; ++b) {
        pp[b] = P::ld_pos(A.pos, 3 * j + b);
        if (ADJ) {
          U t = P::ld_u(A.u, 3 * j + b);
          uu[b][0] = t.x;
          uu[b][1] = t.y;
          uu[b][2] = t.z;
        }
      }
    }
  };

  // STEP: the step's inputs that do not depend on this kernel's results (molecule index, x, v,
  // the F seed, the noise) are read and generated by the step waves before their pair loops,
  // while the other waves keep the SIMDs busy, and wait in LDS for the epilogue
  constexpr int NM = RS_BLOCK / CH, MPW = RS_WAVE / 3;  // molecules per block, per step wave
  struct StepIn {
    double x[3], v[3], f[3], xi[3];
  };
  __shared__ StepIn stin[STEP ? 3 * NM : 1];
  __shared__ int smol[STEP ? NM : 1];
  // a thread's place in the step: wave sw, lane sl, molecule sml of the block (sorted index sim)
  // and atom sa; sal: the lane holds an atom. The epilogue recomputes it from an opaque copy of
  // the thread index rather than keeping it live across the pair loops (in 64 VGPRs it spilled).
  struct StepLane {
    int sw, sl, sml, sa, sim;
    bool slv, sal;
  };
  auto step_lane = [&](int tid) {
    StepLane q;
    q.sw = tid / RS_WAVE;
    q.sl = tid % RS_WAVE;
    q.slv = q.sl < 3 * MPW;
    q.sml = q.sw * MPW + (q.slv ? q.sl / 3 : 0);
    q.sa = q.slv ? q.sl % 3 : 0;
    q.sim = blk * NM + q.sml;
    q.sal = q.slv && q.sml < NM && q.sim < p.M;
    return q;
  };
  if constexpr (STEP) {
    const StepLane SL = step_lane(threadIdx.x);
    const int sml = SL.sml, sa = SL.sa, sim = SL.sim;
    const bool sal = SL.sal;
    if (sal) {
      const int m = A.order[sim];
      const int atom = 3 * m + sa;
      StepIn& q = stin[3 * sml + sa];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        q.x[k] = St.b.x[3 * (size_t)atom + k];
        q.v[k] = St.b.v[3 * (size_t)atom + k];
        q.f[k] = A.F[3 * (size_t)atom + k];
      }
      atom_noise(St.b, p, St.step, St.seed, atom, sa == 0 ? p.mO : p.mH, St.kT, q.xi);
      if (sa == 0) smol[sml] = m;
    }
  }

  Pos own[3];
  R uo[3][3];
  R F[3][3], hh[3][3], Flj[3], Fs[3], Fml[3];
  // FP64 path: the pair forces as fixed-point integer sums (fx_term), F unused
  constexpr bool FX = sizeof(R) == 8;
  unsigned long long Fq[3][3];
  gather(act ? i : 0, 0, own, uo);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      F[a][k] = 0;
      hh[a][k] = 0;
      Fq[a][k] = 0;
    }
    Flj[a] = 0;
    Fs[a] = 0;
    Fml[a] = 0;
  }
  const R rc2 = (R)A.rc2;
  const R kA[3] = {(R)A.kA[0], (R)A.kA[1], (R)A.kA[2]};
  const R kB[3] = {(R)A.kB[0], (R)A.kB[1], (R)A.kB[2]};
  const R k3A[3] = {(R)A.k3A[0], (R)A.k3A[1], (R)A.k3A[2]};
  const R sig2 = (R)A.sig2A, eps24 = (R)A.eps24A, epss = (R)A.epssA;
  const R lu = (R)A.lu;

  // MLP: the (f, f') table in LDS (FP32 path; 4 KB), own O's w and the table-gradient
  // accumulator in LDS (adjoint)
  __shared__ long long lg[ADJ ? MLP_NG : 1];
  __shared__ float2 stab[sizeof(R) == 4 ? MLP_NT : 1];
  R wO[3] = {0, 0, 0};
  double gscale = 0.0;
  const bool mlp = FMLP;
  if (mlp) {
    if (sizeof(R) == 4)
      for (int q = threadIdx.x; q < MLP_NT; q += RS_BLOCK) stab[q] = ((const float2*)A.mtab)[q];
    if (ADJ) {
      for (int q = threadIdx.x; q < MLP_NG; q += RS_BLOCK) lg[q] = 0;
      gscale = mlp_scale(A.wmax);
      if (act) {
        const size_t mo = 9 * (size_t)A.order[i];
        wO[0] = (R)(A.wsrc[mo] * A.wscale[0]);
        wO[1] = (R)(A.wsrc[mo + 1] * A.wscale[0]);
        wO[2] = (R)(A.wsrc[mo + 2] * A.wscale[0]);
      }
    }
    __syncthreads();
  }
  const R mrlo = (R)A.mrlo, mh = (R)A.mh, minvh = (R)A.minvh;

  const int4 ncnt = act ? A.ncnt[gtid] : make_int4(0, 0, 0, 0);
  bool fxbad = false;  // FP64 adjoint: a fixed-point term out of range (flagged after the loops)
  // fixed-point terms each own atom added (FP64 path), whose bias is removed after the loops
  unsigned nt[3] = {0u, 0u, 0u};
  auto shell = [&](auto tag) {
    constexpr int S = decltype(tag)::value;
    constexpr int b0 = S == 2 ? 1 : 0;  // first atom of j that can interact
    const int n = S == 0 ? ncnt.x : (S == 1 ? ncnt.y : ncnt.z);
    const LT* __restrict__ L = (const LT*)A.nbr + list_base(p, gtid, S);
    // Software pipeline: the next entry's index (and, where registers allow, its inputs) is
    // loaded while this entry is evaluated; past the end the last entry is read again.
    constexpr bool PF = pair_prefetch<P, CH, ADJ>();
    const int j0 = n > 0 ? L[0] : 0;
    Pos pn[3];
    R un[3][3];
    int jn = j0;
    if constexpr (PF) {
      gather(j0, b0, pn, un);
      jn = n > 1 ? L[RS_WAVE] : j0;
    }
    for (int k = 0; k < n; ++k) {
      Pos pj[3];
      R uj[3][3];
      if constexpr (PF) {
#pragma unroll
        for (int b = b0; b < 3; ++b) {
          pj[b] = pn[b];
          if (ADJ)
            for (int t = 0; t < 3; ++t) uj[b][t] = un[b][t];
        }
        gather(jn, b0, pn, un);
        jn = L[(size_t)min(k + 2, n - 1) * RS_WAVE];
      } else {
        gather(jn, b0, pj, uj);
        jn = L[(size_t)min(k + 1, n - 1) * RS_WAVE];
      }
#pragma unroll
      for (int a = b0; a < 3; ++a) {
#pragma unroll
        for (int b = b0; b < 3; ++b) {
          if (S == 1 && a == 0 && b == 0) continue;
          R d[3];
          P::diff(p, own[a], pj[b], d[0], d[1], d[2]);
          const R r2 = P::r2(d[0], d[1], d[2]);
          // An atom pair outside the cutoff contributes nothing: the lane branches over it, and a
          // wave whose lanes all miss it skips it (the branch is taken for whole waves often
          // enough, most of all in the outer shells, that this beats evaluating it to zeros).
          // r2 == 0 (coincident atoms) within the cutoff gives non-finite forces, which are flagged.
          if (!(r2 <= rc2)) continue;
          if constexpr (FX) ++nt[a];
          const R ri = P::rinv(r2);
          const R ri2 = ri * ri, ri3 = ri2 * ri;
          const int ty = (a > 0) + (b > 0);
          const R fr = P::fma_(kA[ty], ri3, -kB[ty]);
          R ftot = fr;  // f/r of all the pair's terms (FP64 path)
          if constexpr (!FX) {
#pragma unroll
            for (int t = 0; t < 3; ++t) F[a][t] -= fr * d[t];
          }
          R du[3], dot = 0;
          if (ADJ) {
#pragma unroll
            for (int t = 0; t < 3; ++t) du[t] = uo[a][t] - uj[b][t];
            dot = P::fma_(d[2], du[2], P::fma_(d[1], du[1], d[0] * du[0]));
            // (k3A r'^-3) r'^-2: the FP32 path's r'^-5 alone would be subnormal
            const R gd = ((k3A[ty] * ri3) * ri2) * dot;
#pragma unroll
            for (int t = 0; t < 3; ++t) hh[a][t] = P::fma_(gd, d[t], P::fma_(fr, du[t], hh[a][t]));
          }
          if (S == 0 && a == 0 && b == 0) {
            const R s2 = sig2 * ri2, s6 = s2 * s2 * s2, s12 = s6 * s6;
            const R frl0 = (eps24 * ri2) * P::fma_((R)2, s12, -s6);
            const R frl = frl0;
            if constexpr (FX) {
              ftot = add_nc(ftot, frl);
            } else {
#pragma unroll
              for (int t = 0; t < 3; ++t) F[0][t] -= frl * d[t];
            }
            if (ADJ) {
              const R gdl = -eps24 * ri2 * ri2 * ((R)28 * s12 - (R)8 * s6) * dot;
              const R fs = epss * ri2 * ((R)4 * s12 - s6);
#pragma unroll
              for (int t = 0; t < 3; ++t) {
                hh[0][t] = P::fma_(gdl, d[t], P::fma_(frl, du[t], hh[0][t]));
                Flj[t] = P::fma_(-frl, d[t], Flj[t]);
                Fs[t] = P::fma_(-fs, d[t], Fs[t]);
              }
            }
            if (mlp) {
              // f(r) = -u'(r) by cubic Hermite between table nodes k, k+1
              const R rr = (r2 * ri) * lu;
              R tk = (rr - mrlo) * minvh;
              if (tk < (R)0) atomicOr(&A.flags[FLAG_MLP_RANGE], 1);
              tk = tk < (R)0 ? (R)0 : tk;
              int kk = (int)tk;
              kk = kk > MLP_NT - 2 ? MLP_NT - 2 : kk;
              R tt = tk - (R)kk;
              tt = tt > (R)1 ? (R)1 : tt;
              R fk0, dk0, fk1, dk1;
              if (sizeof(R) == 4) {
                const float2 e0 = stab[kk], e1 = stab[kk + 1];
                fk0 = (R)e0.x; dk0 = (R)e0.y; fk1 = (R)e1.x; dk1 = (R)e1.y;
              } else {
                const double2* tb = (const double2*)A.mtab;
                const double2 e0 = tb[kk], e1 = tb[kk + 1];
                fk0 = (R)e0.x; dk0 = (R)e0.y; fk1 = (R)e1.x; dk1 = (R)e1.y;
              }
              R hb[4], hdb[4];
              hermite<R>(tt, hb, hdb);
              const R f = hb[0] * fk0 + hb[1] * mh * dk0 + hb[2] * fk1 + hb[3] * mh * dk1;
              const R fp = (hdb[0] * fk0 + hdb[1] * mh * dk0 + hdb[2] * fk1 + hdb[3] * mh * dk1) * minvh;
              const R frm = f * ri;
              if constexpr (FX) {
                ftot = add_nc(ftot, frm);
              } else {
#pragma unroll
                for (int t = 0; t < 3; ++t) F[0][t] -= frm * d[t];
              }
              if (ADJ) {
                const R g2m = (fp * lu - f * ri) * ri2;
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                  hh[0][t] += frm * du[t] + (g2m * dot) * d[t];
                  Fml[t] -= frm * d[t];
                }
                // weight VJP: w_O . dF_O/d(table) = -(w_O . r̂) ∂f/∂(table)
                const double q = -(double)((wO[0] * d[0] + wO[1] * d[1] + wO[2] * d[2]) * ri) * gscale;
                atomicAdd((unsigned long long*)&lg[2 * kk], fix_rn(q * (double)hb[0]));
                atomicAdd((unsigned long long*)&lg[2 * kk + 1], fix_rn(q * (double)(mh * hb[1])));
                atomicAdd((unsigned long long*)&lg[2 * kk + 2], fix_rn(q * (double)hb[2]));
                atomicAdd((unsigned long long*)&lg[2 * kk + 3], fix_rn(q * (double)(mh * hb[3])));
              }
            }
          }
          if constexpr (FX) {
            const R ftot2 = ftot * (R)FX_SCALE;
#pragma unroll
            for (int t = 0; t < 3; ++t) Fq[a][t] += fx_term(ftot2, d[t]);
            // |c| <= |f/r| r (FX_CMAX2); NaN (coincident atoms) fails the test too. The adjoint
            // carries the outcome through its loops and flags once (3 % faster there, fewer
            // VGPRs); the forward flags where it happens (carried, it spills at 4 waves per SIMD).
            if constexpr (ADJ) fxbad |= !(fabs(ftot2) <= (R)FX_CMAX2 * ri);
            else if (!(fabs(ftot2) <= (R)FX_CMAX2 * ri)) atomicOr(&A.flags[FLAG_FIXED_RANGE], 1);
          }
        }
      }
    }
  };
  shell(std::integral_constant<int, 0>());
  shell(std::integral_constant<int, 1>());
  shell(std::integral_constant<int, 2>());
  if (fxbad) atomicOr(&A.flags[FLAG_FIXED_RANGE], 1);

  // combine the CH lanes of the molecule in a fixed order by DPP row steps (chunk_sum_dpp): FP32
  // or FP64 partials, and the FP64 path's integer force sums (each lane's bias removed first: one
  // per term it added, i.e. per evaluated atom pair)
  auto lane_sum = [](R v) -> double { return (double)chunk_sum_dpp<CH>(v); };
  double Fd[9], hd[9], Fljd[3], Fsd[3], Fmld[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const unsigned long long nterm = nt[a];  // fixed-point terms this own atom added
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      if constexpr (FX) {
        const long long q = (long long)chunk_sum_dpp<CH>(Fq[a][t] - nterm * FX_BIAS);
        Fd[3 * a + t] = -((double)q * FX_INV);
      } else {
        Fd[3 * a + t] = lane_sum(F[a][t]);
      }
      if (ADJ) hd[3 * a + t] = lane_sum(hh[a][t]);
    }
    if (ADJ) {
      Fljd[a] = lane_sum(Flj[a]);
      Fsd[a] = lane_sum(Fs[a]);
      Fmld[a] = mlp ? lane_sum(Fml[a]) : 0.0;
    }
  }

  __shared__ double vjs[RS_BLOCK / CH][6];
  if constexpr (STEP) {
    // Forward step n of the block's molecules (the step kernel's arithmetic), then the bonded
    // forces at the new positions, which seed F for the next pair kernel. The pair forces go
    // through LDS to as few waves as hold three lanes per molecule (one wave for CH = 16): the
    // step's FP64 work then runs on 48 of 64 lanes rather than 3 of every 16.
    static_assert(!ADJ, "the step variant is forward-only");
    __shared__ double fpair[NM][9];
    if (c == 0) {
#pragma unroll
      for (int t = 0; t < 9; ++t) fpair[threadIdx.x / CH][t] = Fd[t];
    }
    __syncthreads();
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const StepLane SL = step_lane(tid);
    const int sw = SL.sw, sl = SL.sl, sml = SL.sml, sa = SL.sa, sim = SL.sim;
    const bool slv = SL.slv, sal = SL.sal;
    if (sw * MPW < NM) {  // wave-uniform
      const int m = sal ? smol[sml] : 0;
      const int atom = 3 * m + sa;
      const double mass = sa == 0 ? p.mO : p.mH;
      double xa[3] = {0.0, 0.0, 0.0};
      if (sal) {
        const StepIn& q = stin[3 * sml + sa];
        double Ft[3];
        bool fin = true;
#pragma unroll
        for (int t = 0; t < 3; ++t) {
          Ft[t] = q.f[t] + fpair[sml][3 * sa + t];
          fin = fin && isfinite(Ft[t]);
        }
        fin = step_fwd_atom(p, St.b, atom, mass, q.x, q.v, Ft, q.xi, St.dt, St.e, St.s, xa) && fin;
        if (!fin) atomicOr(&A.flags[FLAG_NONFINITE], 1);
        if (St.assign && sa == 0) file_molecule(p, St.b.cell, St.b.cell_count, m, xa[0], xa[1], xa[2]);
      }
      double xm[9];
      gather_mol3(slv ? 3 * (sl / 3) : 0, xa, xm);
      if (sal) {
        write_sorted(p, sizeof(R) == 4 ? RS_PRECISION_FP32 : RS_PRECISION_FP64, 3 * sim + sa, xa,
                     St.posq_next, St.pos64_next);
        double Fb[9], hb[9], vj[3];
        bonded_mol<false>(p, xm, nullptr, nullptr, Fb, hb, vj, A.flags, St.b.O64[sim]);
#pragma unroll
        for (int t = 0; t < 3; ++t) A.F[3 * (size_t)atom + t] = sa == 0 ? Fb[t] : (sa == 1 ? Fb[3 + t] : Fb[6 + t]);
      }
    }
  } else if (c == 0) {
    double vj[6] = {0, 0, 0, 0, 0, 0};
    if (act) {
      const int m = A.order[i];
      double* Fm = A.F + 9 * (size_t)m;
      if (ADJ) {
        double wm[9];
#pragma unroll
        for (int t = 0; t < 9; ++t) wm[t] = A.wsrc[9 * (size_t)m + t] * A.wscale[t < 3 ? 0 : 1];
        // pair-part VJPs on this molecule's atoms: F_RF = F_pair - F_LJ - F_MLP (LJ and the MLP
        // act on O only)
        double sq = 0.0, se = 0.0, ss = 0.0;
#pragma unroll
        for (int t = 0; t < 9; ++t) sq += wm[t] * (Fd[t] - (t < 3 ? Fljd[t] + Fmld[t] : 0.0));
#pragma unroll
        for (int t = 0; t < 3; ++t) {
          se += wm[t] * Fljd[t];
          ss += wm[t] * Fsd[t];
        }
        vj[0] = ss;
        vj[1] = p.eps != 0.0 ? se / p.eps : 0.0;
        vj[2] = 2.0 * sq / p.qH;
        vj[3] = A.vjb[3 * (size_t)m];
        vj[4] = A.vjb[3 * (size_t)m + 1];
        vj[5] = A.vjb[3 * (size_t)m + 2];
        double* hm = A.h + 9 * (size_t)m;
#pragma unroll
        for (int t = 0; t < 9; ++t) hm[t] += hd[t] * A.hscale;
      }
      // F was seeded with the molecule's bonded forces; add the pair forces
      bool fin = true;
#pragma unroll
      for (int t = 0; t < 9; ++t) {
        const double f = Fm[t] + Fd[t];
        Fm[t] = f;
        fin = fin && isfinite(f);
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
      A.slots[(size_t)blk * 8 + threadIdx.x] += s;  // this block's slot, summed over steps
    }
    if (mlp) {
      // this block's table gradient of the step, accumulated over the trajectory in its own slot
      const double inv = 1.0 / gscale;
      double* sl = A.mslot + (size_t)blk * MLP_NG;
      for (int q = threadIdx.x; q < MLP_NG; q += RS_BLOCK)
        if (lg[q]) sl[q] += (double)lg[q] * inv;
    }
  }
}

// dL/dθ += Σ_blocks of the adjoint pair kernel's θ-VJP partials: one block per parameter, each
// summing the partials in block order (deterministic; the kernel boundary orders it after them).
__global__ __launch_bounds__(RS_BLOCK) void k_vjp_reduce(const double* __restrict__ slots, int nslots,
                                                         double* dth) {
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
  if (threadIdx.x == 0) dth[k] += red[0];
}

template <class P, int CH, bool ADJ, bool L16>
void dispatch_ch(const Launch& L, const PairArgs& a, const StepArgs* st) {
  const int g = pair_blocks(L.p);
  StepArgs none{};
  if constexpr (!ADJ) {
    if (st) {
      if (a.mlp) k_pair<P, CH, false, true, true, L16><<<g, RS_BLOCK, 0, L.s>>>(a, *st);
      else k_pair<P, CH, false, false, true, L16><<<g, RS_BLOCK, 0, L.s>>>(a, *st);
      return;
    }
  }
  if (a.mlp) k_pair<P, CH, ADJ, true, false, L16><<<g, RS_BLOCK, 0, L.s>>>(a, none);
  else k_pair<P, CH, ADJ, false, false, L16><<<g, RS_BLOCK, 0, L.s>>>(a, none);
}

// lanes per molecule (8 or 16, rs_create) and list entry width (p.list16)
template <class P, bool ADJ>
void dispatch(const Launch& L, const PairArgs& a, const StepArgs* st) {
  if (L.p.CH == 8) {
    if (L.p.list16) dispatch_ch<P, 8, ADJ, true>(L, a, st);
    else dispatch_ch<P, 8, ADJ, false>(L, a, st);
  } else {
    if (L.p.list16) dispatch_ch<P, 16, ADJ, true>(L, a, st);
    else dispatch_ch<P, 16, ADJ, false>(L, a, st);
  }
}

}  // namespace

namespace {

void pair_launch(const Launch& L, bool adjoint, const double* wsrc, double wscale_O,
                 double wscale_H, const StepArgs* st) {
  PairArgs a;
  a.mlp = L.use_mlp;
  a.mtab = L.b.mlp_tab;
  a.mrlo = MLP_RLO;
  a.mh = (L.mlp_rc - MLP_RLO) / (MLP_NT - 1);
  a.minvh = 1.0 / a.mh;
  a.mslot = L.b.mlp_slot;
  a.wmax = L.b.wmax + 64 * (L.parity & 1);
  a.p = L.p;
  a.nbr = L.b.nbr;
  a.ncnt = L.b.ncnt;
  a.order = L.b.order;
  a.wsrc = wsrc;
  a.wscale[0] = wscale_O;
  a.wscale[1] = wscale_H;
  a.vjb = L.b.vjb;
  a.F = L.b.F;
  a.h = L.b.h;
  a.slots = L.b.slots;
  a.flags = L.b.flags;
  const SysParams& p = L.p;
  const bool f32 = L.precision == RS_PRECISION_FP32;
  const double lu = f32 ? p.L[0] / 4294967296.0 : 1.0;
  for (int t = 0; t < 3; ++t) {
    a.kA[t] = p.kqq[t] / (lu * lu);
    a.kB[t] = p.kqq[t] * 2.0 * p.krf * lu;
    a.k3A[t] = -3.0 * p.kqq[t] / (lu * lu);
  }
  a.sig2A = p.sig * p.sig / (lu * lu);
  a.eps24A = 24.0 * p.eps / lu;
  a.epssA = 144.0 * p.eps / (p.sig * lu);
  a.rc2 = f32 ? (double)p.rc2q : p.rc2;
  a.lu = lu;
  a.hscale = 1.0 / lu;
  if (f32) {
    a.pos = L.b.posq;
    a.u = nullptr;
    a.pu = L.b.pu;
    if (adjoint) dispatch<P32, true>(L, a, nullptr);
    else dispatch<P32, false>(L, a, st);
  } else {
    a.pos = L.b.pos64;
    a.u = L.b.ud;
    if (adjoint) dispatch<P64, true>(L, a, nullptr);
    else dispatch<P64, false>(L, a, st);
  }
}

}  // namespace

void launch_pair(const Launch& L, bool adjoint, const double* wsrc, double wscale_O, double wscale_H) {
  pair_launch(L, adjoint, wsrc, wscale_O, wscale_H, nullptr);
}

void launch_pair_step(const Launch& L, int step, uint64_t seed, double dt, double e, double s,
                      double kT, bool assign) {
  StepArgs st;
  st.b = L.b;
  st.posq_next = L.b.posq_alt;
  st.pos64_next = L.b.pos64_alt;
  st.step = step;
  st.assign = assign;
  st.seed = seed;
  st.dt = dt;
  st.e = e;
  st.s = s;
  st.kT = kT;
  pair_launch(L, false, nullptr, 0.0, 0.0, &st);
}

void launch_vjp_reduce(const Launch& L) {
  k_vjp_reduce<<<6, RS_BLOCK, 0, L.s>>>(L.b.slots, pair_blocks(L.p), L.b.dth);
}
