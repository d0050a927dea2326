// step.hip — streaming kernel of the reverse-time adjoint step, and the per-molecule bonded,
// seed, noise and clipping kernels.
//
// The forward step n (forward_sim loop, water/water.jl:1030-1037; integrator.h) runs at the end
// of the forward pair kernel. It is inverted here, FP64 state, one lane per atom:
// Reverse step (reverse_sim, water/water.jl:1496-1521), fused as [part 2 of step n+1][part 1 of n]:
//   part 2: v -= F dt/m; B = e B + dt² h; A += B   (dL/dθ is accumulated by the pair kernel)
//   part 1: x -= v dt/2; v = (v - s ξ_n)/e; x -= v dt/2; wrap; D = e D + dt² h (1+e)/2; C += D
// ξ_n is regenerated from the Philox counter (step n, atom) — the same bits forward and reverse.
// Every lane also writes its atom's slot of the sorted adjoint pair inputs for the next pair
// kernel.
// Expressions follow the reference's broadcast order and are evaluated without contraction.
#include "internal.h"
#include "bonded.h"
#include "integrator.h"

namespace {

// Per-molecule kernels run one wave per block so that the few waves there are spread over all CUs.
constexpr int MOL_BLOCK = 64;

// max over the block's (one wave's) molecules of |w| on O atoms, folded into one of 64 slots (a
// non-negative double kept as its uint64 bit pattern, which orders like the value): the per-step
// fixed-point scale of the MLP table gradient. Spreading the atomics over 64 addresses keeps
// them from serialising on one.
__device__ __forceinline__ void wmax_fold(double v, double* slots) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0 && v > 0.0)
    atomicMax((unsigned long long*)&slots[blockIdx.x & 63], (unsigned long long)__double_as_longlong(v));
}

__global__ void k_wrap(SysParams p, double* x) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 3 * p.N) return;
  x[t] = wrap1(x[t], p.L[t % 3]);
}

// The step kernels run one lane per atom: a 64-lane block holds ATOM_MOLS molecules in lanes
// 3*mol + atom (lane 63 idles), so the FP64 state streams per atom, and the three lanes of a
// molecule exchange coordinates by shuffles and evaluate its bonded terms (redundantly, in step).
constexpr int ATOM_MOLS = 21;

struct AtomLane {
  int mol, a, atom, base;  // molecule, atom within it (0 = O), atom index, first lane of molecule
  bool act;
};

__device__ __forceinline__ AtomLane atom_lane(const SysParams& p) {
  AtomLane L;
  const int l = threadIdx.x;
  const int ml = l < 3 * ATOM_MOLS ? l / 3 : ATOM_MOLS - 1;
  L.a = l - 3 * ml;
  L.mol = blockIdx.x * ATOM_MOLS + ml;
  L.act = l < 3 * ATOM_MOLS && L.mol < p.M;
  L.atom = 3 * L.mol + (L.a < 3 ? L.a : 0);
  L.base = 3 * ml;
  return L;
}

// The molecule's nine values (O, H1, H2) of a per-atom 3-vector held by its three lanes.
__device__ __forceinline__ void gather_mol(const AtomLane& L, const double v[3], double out[9]) {
  gather_mol3(L.base, v, out);
}

// Reverse step: [part 2 of step n+1][part 1 of step n]; after part 1 the bonded forces, HVP and
// θ-VJP at the reconstructed positions seed F, h and vjb for the adjoint pair kernel of step n.
__global__ __launch_bounds__(RS_WAVE) void k_step_rev(SysParams p, int precision, DevBufs b,
                                                      int do_r2, int do_r1, int step,
                                                      uint64_t seed, double dt, double e,
                                                      double s, double kT, int nslots, int clip,
                                                      int mlp, int assign) {
#pragma clang fp contract(off)
  const AtomLane L = atom_lane(p);
  const double dt2 = dt * dt;
  // MLP scale slots alternate by step parity: this step's pair kernel reads slot (step & 1)
  if (mlp && do_r1 && blockIdx.x == 0) b.wmax[64 * ((step + 1) & 1) + threadIdx.x] = 0.0;
  double xa[3] = {0.0, 0.0, 0.0}, ua[3] = {0.0, 0.0, 0.0}, wa[3] = {0.0, 0.0, 0.0};
  const double mass = L.a == 0 ? p.mO : p.mH;
  // every load is issued before any store (one memory round trip per wave; the state arrays may
  // alias as far as the compiler knows)
  double X[3] = {0, 0, 0}, V[3] = {0, 0, 0}, H[3] = {0, 0, 0}, Aa[3] = {0, 0, 0};
  double Fo[3] = {0, 0, 0}, Bo[3] = {0, 0, 0}, Do[3] = {0, 0, 0}, Co[3] = {0, 0, 0};
  int rk = 0;
  double4 ob = make_double4(0.0, 0.0, 0.0, 0.0);
  double cs = 1.0;
  if (L.act) {
    if (do_r1) rk = b.rank[L.mol];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const size_t o = 3 * (size_t)L.atom + k;
      X[k] = b.x[o];
      V[k] = b.v[o];
      H[k] = b.h[o];
      Aa[k] = b.A[o];
      if (do_r2) {
        Fo[k] = b.F[o];
        Bo[k] = b.B[o];
      }
      if (do_r1) {
        Do[k] = b.D[o];
        Co[k] = b.C[o];
      }
    }
    if (clip) cs = b.scal[4];
    if (do_r1) ob = b.O64[rk];
  }
  if (L.act) {
    double xi[3] = {0.0, 0.0, 0.0};
    if (do_r1) atom_noise(b, p, step, seed, L.atom, mass, kT, xi);
    bool fin = true;
    double vn[3], Bn[3], An[3], Dn[3], Cn[3], hn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double x = X[k], v = V[k], h = H[k], A = Aa[k];
      if (clip) h = h * cs;
      hn[k] = h;
      Bn[k] = 0.0;
      if (do_r2) {
        v = v - (Fo[k] * dt) / mass;
        double B = e * Bo[k] + dt2 * h;
        Bn[k] = B;
        A = A + B;
      }
      An[k] = A;
      Dn[k] = 0.0;
      Cn[k] = 0.0;
      if (do_r1) {
        x = x - (v * dt) / 2.0;
        v = (v - xi[k] * s) / e;
        x = x - (v * dt) / 2.0;
        x = wrap1(x, p.L[k]);
        double D = e * Do[k] + ((dt2 * h) * (1.0 + e)) / 2.0;
        double C = Co[k] + D;
        Dn[k] = D;
        Cn[k] = C;
        ua[k] = C / mass;
        wa[k] = A * (dt2 / (2.0 * mass));
      }
      vn[k] = v;
      xa[k] = x;
      fin = fin && isfinite(x) && isfinite(v);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const size_t o = 3 * (size_t)L.atom + k;
      if (do_r2) {
        b.B[o] = Bn[k];
        b.A[o] = An[k];
      }
      if (do_r1) {
        b.D[o] = Dn[k];
        b.C[o] = Cn[k];
      } else if (clip) {
        b.h[o] = hn[k];
      }
      b.x[o] = xa[k];
      b.v[o] = vn[k];
    }
    if (!fin) atomicOr(&b.flags[FLAG_NONFINITE], 1);
    if (assign && do_r1 && L.a == 0) file_molecule(p, b.cell, b.cell_count, L.mol, xa[0], xa[1], xa[2]);
  }
  double wv = 0.0;  // max |w| on this molecule's O (MLP scale)
  if (do_r1) {
    double xm[9], um[9], wm[9];
    gather_mol(L, xa, xm);
    gather_mol(L, ua, um);
    gather_mol(L, wa, wm);
    if (L.act) {
      // the adjoint pair inputs: FP32, the record (positions and u) alone; posq is not read
      // before the next list build repacks it
      if (precision == RS_PRECISION_FP32) {
        int q[3];
        for (int k = 0; k < 3; ++k) q[k] = (int)quant1(xa[k], p.qscale[k]);
        write_pair_rec(b.pu, rk, L.a, q, ua);
      } else {
        b.pos64[3 * rk + L.a] = make_double4(xa[0], xa[1], xa[2], 0.0);
        b.ud[3 * rk + L.a] = make_double4(ua[0], ua[1], ua[2], 0.0);
      }
      wv = fmax(fabs(wm[0]), fmax(fabs(wm[1]), fabs(wm[2])));
      double F[9], h[9], vj[3];
      bonded_mol<true>(p, xm, um, wm, F, h, vj, b.flags, ob);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        b.F[3 * (size_t)L.atom + k] = F[3 * L.a + k];
        b.h[3 * (size_t)L.atom + k] = h[3 * L.a + k];
      }
      if (L.a == 0) {
        b.vjb[3 * (size_t)L.mol] = vj[0];
        b.vjb[3 * (size_t)L.mol + 1] = vj[1];
        b.vjb[3 * (size_t)L.mol + 2] = vj[2];
      }
    }
  }
  if (mlp && do_r1) wmax_fold(wv, b.wmax + 64 * (step & 1));
}

// Bonded terms of the current state alone (start of a trajectory and single evaluations):
// F (and h, vjb) are seeded as in the step kernels. u = csrc*cscale, w = wsrc*wscale per atom type.
template <bool ADJ>
__global__ __launch_bounds__(MOL_BLOCK) void k_bonded(SysParams p, DevBufs b, const double* csrc,
                                                      double cO, double cH, const double* wsrc,
                                                      double wO, double wH, int mlp) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  double wv = 0.0;
  if (m < p.M) {
  double xm[9], um[9], wm[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    xm[t] = b.x[9 * (size_t)m + t];
    if (ADJ) {
      um[t] = csrc[9 * (size_t)m + t] * (t < 3 ? cO : cH);
      wm[t] = wsrc[9 * (size_t)m + t] * (t < 3 ? wO : wH);
    }
  }
  if (ADJ) wv = fmax(fabs(wm[0]), fmax(fabs(wm[1]), fabs(wm[2])));
  double F[9], h[9], vj[3];
  bonded_mol<ADJ>(p, xm, um, wm, F, h, vj, b.flags, b.O64[b.rank[m]]);
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    b.F[9 * (size_t)m + t] = F[t];
    if (ADJ) b.h[9 * (size_t)m + t] = h[t];
  }
  if (ADJ) {
    b.vjb[3 * (size_t)m] = vj[0];
    b.vjb[3 * (size_t)m + 1] = vj[1];
    b.vjb[3 * (size_t)m + 2] = vj[2];
  }
  }
  if (ADJ && mlp) wmax_fold(wv, b.wmax);
}

// Gradient seed g (b.grad) into the accumulators: reset (A = B = g(1+e), C = 0,
// D = g(1+2e+e²)/(2e), h = 0; water/water.jl:1480-1483,1540-1545) or additive (same terms added).
__global__ void k_seed(SysParams p, DevBufs b, int reset, double e) {
#pragma clang fp contract(off)
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= 3 * p.N) return;
  const double g = b.grad[o];
  const double ab = g * (1.0 + e);
  const double d = g * (1.0 + 2.0 * e + e * e) / (2.0 * e);
  if (reset) {
    b.A[o] = ab;
    b.B[o] = ab;
    b.C[o] = 0.0;
    b.D[o] = d;
    b.h[o] = 0.0;
  } else {
    b.A[o] += ab;
    b.B[o] += ab;
    b.D[o] += d;
  }
}

__global__ void k_noise(SysParams p, int step, uint64_t seed, double kT, const double* ext,
                        double* xi) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= p.N) return;
  if (ext) {
    for (int k = 0; k < 3; ++k) xi[3 * t + k] = ext[(size_t)(step - 1) * 3 * p.N + 3 * t + k];
    return;
  }
  const double mass = (t % 3 == 0) ? p.mO : p.mH;
  double nn[3];
  normals3(seed, step, t, nn);
  const double sd = sqrt(kT / mass);
  for (int k = 0; k < 3; ++k) xi[3 * t + k] = sd * nn[k];
}

// ||h||: per-block partial sums, then one block sums them in order and sets the clip scale.
__global__ void k_hnorm_part(SysParams p, const double* h, double* part) {
  __shared__ double red[RS_BLOCK];
  double acc = 0.0;
  for (int o = blockIdx.x * RS_BLOCK + threadIdx.x; o < 3 * p.N; o += gridDim.x * RS_BLOCK)
    acc += h[o] * h[o];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = RS_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_hnorm_final(int nparts, const double* part, double clip, double* scal) {
  if (threadIdx.x != 0) return;
  double s = 0.0;
  for (int q = 0; q < nparts; ++q) s += part[q];
  double nn = sqrt(s);
  scal[4] = nn > clip ? clip / nn : 1.0;
}

}  // namespace

void launch_wrap_state(const Launch& L) {
  k_wrap<<<(3 * L.p.N + RS_BLOCK - 1) / RS_BLOCK, RS_BLOCK, 0, L.s>>>(L.p, L.b.x);
}

void launch_bonded(const Launch& L, bool adjoint, const double* csrc, double cO, double cH,
                   const double* wsrc, double wO, double wH) {
  const int g = (L.p.M + MOL_BLOCK - 1) / MOL_BLOCK;
  if (adjoint)
    k_bonded<true><<<g, MOL_BLOCK, 0, L.s>>>(L.p, L.b, csrc, cO, cH, wsrc, wO, wH, L.use_mlp);
  else
    k_bonded<false><<<g, MOL_BLOCK, 0, L.s>>>(L.p, L.b, csrc, cO, cH, wsrc, wO, wH, 0);
}

void launch_step_rev(const Launch& L, bool do_r2, bool do_r1, int step, uint64_t seed, double dt,
                     double e, double s, double kT, int nslots, bool clip, bool assign) {
  k_step_rev<<<(L.p.M + ATOM_MOLS - 1) / ATOM_MOLS, RS_WAVE, 0, L.s>>>(
      L.p, L.precision, L.b, do_r2, do_r1, step, seed, dt, e, s, kT, nslots, clip, L.use_mlp,
      assign && do_r1);
}

void launch_seed(const Launch& L, bool reset, double e) {
  k_seed<<<(3 * L.p.N + RS_BLOCK - 1) / RS_BLOCK, RS_BLOCK, 0, L.s>>>(L.p, L.b, reset, e);
}

void launch_noise(const Launch& L, int step, uint64_t seed, double kT, double* xi) {
  k_noise<<<(L.p.N + RS_BLOCK - 1) / RS_BLOCK, RS_BLOCK, 0, L.s>>>(L.p, step, seed, kT,
                                                                    L.b.ext_noise, xi);
}

void launch_hnorm(const Launch& L, double clip) {
  const int parts = 256;
  k_hnorm_part<<<parts, RS_BLOCK, 0, L.s>>>(L.p, L.b.h, L.b.hpart);
  k_hnorm_final<<<1, 64, 0, L.s>>>(parts, L.b.hpart, clip, L.b.scal);
}
