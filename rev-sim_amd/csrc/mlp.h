// mlp.h — the radial-basis MLP pair term as the pair kernels see it.
//
// The network u(r) (16 Gaussian RBF features with a cosine cutoff -> 64 -> 64 -> 1, SiLU, BF16
// weights) depends on the O-O distance only, so for the weights of a trajectory the pair force
// f(r) = -u'(r) is tabulated once on MLP_NT nodes r_k = r_lo + k h of [r_lo, r_c] together with
// its derivative f'(r) = -u''(r) (mlp.hip: forward-mode jets, on MFMA in BF16 with FP32
// accumulation in the FP32 path, FP64 in the parity path). Pairs use the cubic Hermite
// interpolant of (f, f') between nodes, which has a continuous derivative, so forces, the
// Hessian-vector product and the weight VJP are all exact for the interpolated model; its
// distance from the network is below 1e-8 relative in f and 1e-6 in f' for h = 1.66 pm.
#pragma once

#include "common.h"

#define MLP_NF 16
#define MLP_NH 64
#define MLP_NW (MLP_NH * MLP_NF + MLP_NH + MLP_NH * MLP_NH + MLP_NH + MLP_NH + 1)
#define MLP_W1 0
#define MLP_B1 (MLP_NH * MLP_NF)
#define MLP_W2 (MLP_B1 + MLP_NH)
#define MLP_B2 (MLP_W2 + MLP_NH * MLP_NH)
#define MLP_W3 (MLP_B2 + MLP_NH)
#define MLP_B3 (MLP_W3 + MLP_NH)
#define MLP_NT 512          // table nodes
#define MLP_RLO 0.15        // nm; O-O pairs below it are flagged (LJ keeps them far above)
#define MLP_NG (2 * MLP_NT)  // table-gradient coefficients: dL/df_k, dL/df'_k

struct MlpTable {
  const void* tab;  // float2 (FP32 path) or double2 (FP64 path) per node: (f, f')
  double rlo, h, inv_h;
  float rlof, hf, inv_hf;
};

// Hermite basis at t in [0,1] and its t-derivative.
template <typename R>
__device__ __forceinline__ void hermite(R t, R b[4], R db[4]) {
  const R t2 = t * t, t3 = t2 * t;
  b[0] = (R)2 * t3 - (R)3 * t2 + (R)1;
  b[1] = t3 - (R)2 * t2 + t;
  b[2] = (R)-2 * t3 + (R)3 * t2;
  b[3] = t3 - t2;
  db[0] = (R)6 * t2 - (R)6 * t;
  db[1] = (R)3 * t2 - (R)4 * t + (R)1;
  db[2] = (R)-6 * t2 + (R)6 * t;
  db[3] = (R)3 * t2 - (R)2 * t;
}
