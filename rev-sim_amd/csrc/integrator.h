// integrator.h — per-atom pieces of the forward Langevin step (forward_sim loop,
// water/water.jl:1030-1037), shared by the integrator kernel and the pair kernel variant that
// ends with the step:
//   v += F dt/m; x += v dt/2; v = e v + s ξ_n; x += v dt/2; wrap       e = exp(-γdt), s = √(1-e²)
// FP64, evaluated without contraction in the reference's broadcast order.
#pragma once

#include "internal.h"

// Sorted pair-input slot of one atom: quantized position (FP32 mode) or the FP64 position.
__device__ __forceinline__ void write_sorted(const SysParams& p, int precision, int slot,
                                             const double x[3], int* posq, double4* pos64) {
  if (precision == RS_PRECISION_FP32)
    for (int k = 0; k < 3; ++k) posq[3 * slot + k] = (int)quant1(x[k], p.qscale[k]);
  else
    pos64[slot] = make_double4(x[0], x[1], x[2], 0.0);
}

// ξ_n of one atom: the caller's noise array if given, else σ N(0,1) from the Philox counter
// (step n, atom) — the same bits in the forward and the reverse pass.
__device__ __forceinline__ void atom_noise(const DevBufs& b, const SysParams& p, int step,
                                           uint64_t seed, int atom, double mass, double kT,
                                           double xi[3]) {
  if (b.ext_noise) {
    const double* src = b.ext_noise + (size_t)(step - 1) * 3 * p.N + 3 * (size_t)atom;
    xi[0] = src[0];
    xi[1] = src[1];
    xi[2] = src[2];
  } else {
    double nn[3];
    normals3(seed, step, atom, nn);
    const double sd = sqrt(kT / mass);
    xi[0] = sd * nn[0];
    xi[1] = sd * nn[1];
    xi[2] = sd * nn[2];
  }
}

// The forward step of one atom from its position x0, velocity v0 and total force F; stores the
// new x and v and returns the new position in xa. False when x or v is not finite.
__device__ __forceinline__ bool step_fwd_atom(const SysParams& p, const DevBufs& b, int atom,
                                              double mass, const double x0[3], const double v0[3],
                                              const double F[3], const double xi[3], double dt,
                                              double e, double s, double xa[3]) {
#pragma clang fp contract(off)
  bool fin = true;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const size_t o = 3 * (size_t)atom + k;
    double v = v0[k] + (F[k] * dt) / mass;
    double x = x0[k] + (v * dt) / 2.0;
    v = v * e + xi[k] * s;
    x = x + (v * dt) / 2.0;
    x = wrap1(x, p.L[k]);
    b.v[o] = v;
    b.x[o] = x;
    xa[k] = x;
    fin = fin && isfinite(x) && isfinite(v);
  }
  return fin;
}

// The molecule's nine values (O, H1, H2) of a per-atom 3-vector held by lanes base .. base+2 of
// the wave (atom a in lane base + a). Every lane of the wave takes part.
__device__ __forceinline__ void gather_mol3(int base, const double v[3], double out[9]) {
#pragma unroll
  for (int b = 0; b < 3; ++b)
#pragma unroll
    for (int k = 0; k < 3; ++k) out[3 * b + k] = __shfl(v[k], base + b, RS_WAVE);
}

// Per-step arguments of the forward step (the pair kernel variant that ends with it).
struct StepArgs {
  DevBufs b;
  int* posq_next;       // sorted pair inputs of the new positions (the other buffer)
  double4* pos64_next;
  int step, assign;     // assign: file each molecule into its cell for the list build that follows
  uint64_t seed;
  double dt, e, s, kT;
};
