// mlp.hip — the radial-basis MLP correction on O-O pairs: tabulation of f = -u' and f' = -u'' by
// forward-mode jets through the network (BF16 MFMA with FP32 accumulation in the FP32 path, the
// activations split into three BF16 terms so that only FP32 rounding remains; FP64
// in the parity path), and dL/dW from the table gradient the adjoint pair kernels accumulate.
//
// Network (features of one O-O distance r < r_c):
//   φ_k(r) = exp(-(r-μ_k)^2/(2w^2)) (cos(π r/r_c)+1)/2, μ_k = 0.2 + k 0.8/15 nm, w = 0.05 nm
//   a1 = SiLU(W1 φ + b1) (64), a2 = SiLU(W2 a1 + b2) (64), u = w3.a2 + b3   [kJ/mol per pair]
// Jets: z' = W φ', z'' = W φ''; SiLU(z)' = σ1 z', SiLU(z)'' = σ2 z'^2 + σ1 z''.
#include "internal.h"
#include "mlp.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest, ties to even
  return (short)(u >> 16);
}

// φ, φ', φ'' of feature k at r.
template <typename R>
__device__ __forceinline__ void rbf1(R r, R rc, int k, R& p0, R& p1, R& p2) {
  const R w = (R)0.05, iw2 = (R)1 / (w * w), a = (R)M_PI / rc;
  const R fc = (R)0.5 * (cos(a * r) + (R)1), fc1 = (R)-0.5 * a * sin(a * r),
          fc2 = (R)-0.5 * a * a * cos(a * r);
  const R mu = (R)0.2 + (R)k * ((R)0.8 / (R)15);
  const R d = r - mu;
  const R g = exp((R)-0.5 * d * d * iw2), g1 = -d * iw2 * g, g2 = (d * d * iw2 - (R)1) * iw2 * g;
  p0 = g * fc;
  p1 = g1 * fc + g * fc1;
  p2 = g2 * fc + (R)2 * g1 * fc1 + g * fc2;
}

template <typename R>
__device__ __forceinline__ void silu4(R z, R& f0, R& f1, R& f2, R& f3) {
  const R s = (R)1 / ((R)1 + exp(-z));
  const R s1 = s * ((R)1 - s), s2 = s1 * ((R)1 - (R)2 * s), s3 = s2 * ((R)1 - (R)2 * s) - (R)2 * s1 * s1;
  f0 = z * s;
  f1 = s + z * s1;
  f2 = (R)2 * s1 + z * s2;
  f3 = (R)3 * s2 + z * s3;
}

// FP64 tabulation: one thread per node.
__global__ void k_mlp_tab64(const double* __restrict__ W, double rc, double rlo, double h,
                            double2* tab) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= MLP_NT) return;
  const double r = rlo + k * h;
  double p[MLP_NF], pa[MLP_NF], pb[MLP_NF];
  for (int f = 0; f < MLP_NF; ++f) rbf1<double>(r, rc, f, p[f], pa[f], pb[f]);
  double a1[MLP_NH], a1a[MLP_NH], a1b[MLP_NH];
  for (int n = 0; n < MLP_NH; ++n) {
    double z = W[MLP_B1 + n], za = 0.0, zb = 0.0;
    for (int f = 0; f < MLP_NF; ++f) {
      const double wv = W[MLP_W1 + n * MLP_NF + f];
      z += wv * p[f];
      za += wv * pa[f];
      zb += wv * pb[f];
    }
    double s0, s1, s2, s3;
    silu4<double>(z, s0, s1, s2, s3);
    a1[n] = s0;
    a1a[n] = s1 * za;
    a1b[n] = s2 * za * za + s1 * zb;
  }
  double u1 = 0.0, u2 = 0.0;
  for (int m = 0; m < MLP_NH; ++m) {
    double z = W[MLP_B2 + m], za = 0.0, zb = 0.0;
    for (int n = 0; n < MLP_NH; ++n) {
      const double wv = W[MLP_W2 + m * MLP_NH + n];
      z += wv * a1[n];
      za += wv * a1a[n];
      zb += wv * a1b[n];
    }
    double s0, s1, s2, s3;
    silu4<double>(z, s0, s1, s2, s3);
    u1 += W[MLP_W3 + m] * (s1 * za);
    u2 += W[MLP_W3 + m] * (s2 * za * za + s1 * zb);
  }
  tab[k] = make_double2(-u1, -u2);
}

// BF16 MFMA tabulation: one wave per 16 nodes. v_mfma_f32_16x16x16_bf16 operand layout: lane l
// holds A[row l%16][k 4(l/16)+v], B[k 4(l/16)+v][col l%16] and D[row 4(l/16)+v][col l%16], v=0..3.
// The three jet streams (value, d/dr, d²/dr²) are three GEMMs sharing the weight operand. The
// weights are BF16 values, so only the activation operand needs care: each FP32 activation is
// split into three BF16 terms x = hi + mid + lo (24 significant bits, the split itself exact to
// FP32 rounding) and the three partial GEMMs accumulate in FP32 on the MFMA, so the table carries
// FP32 rounding only, not BF16 rounding of the features and activations.
__device__ __forceinline__ void split3(float x, short& hi, short& mid, short& lo) {
  hi = f2bf(x);
  const float r1 = x - __uint_as_float((unsigned)(unsigned short)hi << 16);
  mid = f2bf(r1);
  lo = f2bf(r1 - __uint_as_float((unsigned)(unsigned short)mid << 16));
}

// D = A B + C over the three BF16 terms of the activation operand (FP32 accumulation).
__device__ __forceinline__ f32x4 mfma3(const float x[4], const bf16x4& b, f32x4 acc) {
  bf16x4 xh, xm, xl;
  for (int v = 0; v < 4; ++v) {
    short h, m, lo;
    split3(x[v], h, m, lo);
    xh[v] = h;
    xm[v] = m;
    xl[v] = lo;
  }
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xl, b, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xm, b, acc, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xh, b, acc, 0, 0, 0);
}

__global__ __launch_bounds__(64) void k_mlp_tab_bf16(const unsigned short* __restrict__ Wb,
                                                     const double* __restrict__ W, double rc,
                                                     double rlo, double h, float2* tab) {
  __shared__ float A1[3][16][MLP_NH + 1];
  const int l = threadIdx.x, lr = l & 15, lq = l >> 4;
  const int node0 = blockIdx.x * 16;
  // layer 1: A = features of node (node0 + lr), k = 4 lq + v
  const float r = (float)(rlo + (node0 + lr) * h);
  float a[3][4];
  for (int v = 0; v < 4; ++v) rbf1<float>(r, (float)rc, 4 * lq + v, a[0][v], a[1][v], a[2][v]);
  for (int t = 0; t < 4; ++t) {
    const int n = 16 * t + lr;
    bf16x4 b;
    for (int v = 0; v < 4; ++v) b[v] = (short)Wb[MLP_W1 + n * MLP_NF + 4 * lq + v];
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    const f32x4 d0 = mfma3(a[0], b, zero);
    const f32x4 d1 = mfma3(a[1], b, zero);
    const f32x4 d2 = mfma3(a[2], b, zero);
    const float b1 = (float)W[MLP_B1 + n];
    for (int v = 0; v < 4; ++v) {
      const int row = 4 * lq + v;
      float s0, s1, s2, s3;
      silu4<float>(d0[v] + b1, s0, s1, s2, s3);
      A1[0][row][n] = s0;
      A1[1][row][n] = s1 * d1[v];
      A1[2][row][n] = s2 * d1[v] * d1[v] + s1 * d2[v];
    }
  }
  __syncthreads();
  float u1[4] = {0.f, 0.f, 0.f, 0.f}, u2[4] = {0.f, 0.f, 0.f, 0.f};
  for (int t2 = 0; t2 < 4; ++t2) {
    const int out = 16 * t2 + lr;
    f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0, acc2 = acc0;
    for (int ks = 0; ks < 4; ++ks) {
      bf16x4 b;
      float x0[4], x1[4], x2[4];
      for (int v = 0; v < 4; ++v) {
        const int kin = 16 * ks + 4 * lq + v;
        b[v] = (short)Wb[MLP_W2 + out * MLP_NH + kin];
        x0[v] = A1[0][lr][kin];
        x1[v] = A1[1][lr][kin];
        x2[v] = A1[2][lr][kin];
      }
      acc0 = mfma3(x0, b, acc0);
      acc1 = mfma3(x1, b, acc1);
      acc2 = mfma3(x2, b, acc2);
    }
    const float b2 = (float)W[MLP_B2 + out], w3 = (float)W[MLP_W3 + out];
    for (int v = 0; v < 4; ++v) {
      float s0, s1, s2, s3;
      silu4<float>(acc0[v] + b2, s0, s1, s2, s3);
      u1[v] += w3 * (s1 * acc1[v]);
      u2[v] += w3 * (s2 * acc1[v] * acc1[v] + s1 * acc2[v]);
    }
  }
  // sum over the 16 output columns held by lanes with the same lq
  for (int o = 8; o > 0; o >>= 1)
    for (int v = 0; v < 4; ++v) {
      u1[v] += __shfl_xor(u1[v], o, 64);
      u2[v] += __shfl_xor(u2[v], o, 64);
    }
  if (lr == 0)
    for (int v = 0; v < 4; ++v) tab[node0 + 4 * lq + v] = make_float2(-u1[v], -u2[v]);
}

// dL/d(table) = Σ over pair-kernel blocks, in block order.
__global__ void k_mlp_gsum(const double* __restrict__ slotT, int nblocks, double* G) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= MLP_NG) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; ++b) s += slotT[(size_t)b * MLP_NG + c];
  G[c] = s;
}

// dL/dW = Σ_k (-G_f[k]) ∂u'(r_k)/∂W + (-G_d[k]) ∂u''(r_k)/∂W: reverse mode through the jets,
// FP64. Block b handles nodes [64 b, 64 b + 64); thread j owns neuron j's weight rows.
__global__ __launch_bounds__(64) void k_mlp_backprop(const double* __restrict__ W, const double* G,
                                                     double rc, double rlo, double h, double* part) {
  __shared__ double p[3][MLP_NF];
  __shared__ double a1s[3][MLP_NH];
  __shared__ double zb2s[3][MLP_NH];
  const int j = threadIdx.x;
  double dW1[MLP_NF], dW2[MLP_NH], db1 = 0.0, db2 = 0.0, dw3 = 0.0;
  for (int f = 0; f < MLP_NF; ++f) dW1[f] = 0.0;
  for (int n = 0; n < MLP_NH; ++n) dW2[n] = 0.0;
  for (int kk = 0; kk < 64; ++kk) {
    const int k = blockIdx.x * 64 + kk;
    const double r = rlo + k * h;
    const double g1 = -G[2 * k], g2 = -G[2 * k + 1];
    if (j < MLP_NF) rbf1<double>(r, rc, j, p[0][j], p[1][j], p[2][j]);
    __syncthreads();
    // layer 1, neuron j
    double z1 = W[MLP_B1 + j], z1a = 0.0, z1b = 0.0;
    for (int f = 0; f < MLP_NF; ++f) {
      const double wv = W[MLP_W1 + j * MLP_NF + f];
      z1 += wv * p[0][f];
      z1a += wv * p[1][f];
      z1b += wv * p[2][f];
    }
    double t0, t1, t2, t3;
    silu4<double>(z1, t0, t1, t2, t3);
    a1s[0][j] = t0;
    a1s[1][j] = t1 * z1a;
    a1s[2][j] = t2 * z1a * z1a + t1 * z1b;
    __syncthreads();
    // layer 2, neuron m = j
    double z2 = W[MLP_B2 + j], z2a = 0.0, z2b = 0.0;
    for (int n = 0; n < MLP_NH; ++n) {
      const double wv = W[MLP_W2 + j * MLP_NH + n];
      z2 += wv * a1s[0][n];
      z2a += wv * a1s[1][n];
      z2b += wv * a1s[2][n];
    }
    double q0, q1, q2, q3;
    silu4<double>(z2, q0, q1, q2, q3);
    const double w3 = W[MLP_W3 + j];
    dw3 += g1 * (q1 * z2a) + g2 * (q2 * z2a * z2a + q1 * z2b);
    const double ab = g1 * w3, abb = g2 * w3;
    const double zb2b = abb * q1;
    const double zb2a = abb * q2 * 2.0 * z2a + ab * q1;
    const double zb2 = abb * (q3 * z2a * z2a + q2 * z2b) + ab * q2 * z2a;
    db2 += zb2;
    for (int n = 0; n < MLP_NH; ++n) dW2[n] += zb2 * a1s[0][n] + zb2a * a1s[1][n] + zb2b * a1s[2][n];
    zb2s[0][j] = zb2;
    zb2s[1][j] = zb2a;
    zb2s[2][j] = zb2b;
    __syncthreads();
    // back into layer 1, neuron n = j
    double ab1 = 0.0, ab1a = 0.0, ab1b = 0.0;
    for (int m = 0; m < MLP_NH; ++m) {
      const double wv = W[MLP_W2 + m * MLP_NH + j];
      ab1 += wv * zb2s[0][m];
      ab1a += wv * zb2s[1][m];
      ab1b += wv * zb2s[2][m];
    }
    const double zb1b = ab1b * t1;
    const double zb1a = ab1b * t2 * 2.0 * z1a + ab1a * t1;
    const double zb1 = ab1b * (t3 * z1a * z1a + t2 * z1b) + ab1a * t2 * z1a + ab1 * t1;
    db1 += zb1;
    for (int f = 0; f < MLP_NF; ++f) dW1[f] += zb1 * p[0][f] + zb1a * p[1][f] + zb1b * p[2][f];
    __syncthreads();
  }
  double* out = part + (size_t)blockIdx.x * MLP_NW;
  for (int f = 0; f < MLP_NF; ++f) out[MLP_W1 + j * MLP_NF + f] = dW1[f];
  out[MLP_B1 + j] = db1;
  for (int n = 0; n < MLP_NH; ++n) out[MLP_W2 + j * MLP_NH + n] = dW2[n];
  out[MLP_B2 + j] = db2;
  out[MLP_W3 + j] = dw3;
  if (j == 0) out[MLP_B3] = 0.0;
}

__global__ void k_mlp_dwsum(const double* __restrict__ part, int nparts, double* dW) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= MLP_NW) return;
  double s = 0.0;
  for (int b = 0; b < nparts; ++b) s += part[(size_t)b * MLP_NW + q];
  dW[q] = s;
}

}  // namespace

void launch_mlp_tabulate(const Launch& L, double rc) {
  const double h = (rc - MLP_RLO) / (MLP_NT - 1);
  if (L.precision == RS_PRECISION_FP32)
    k_mlp_tab_bf16<<<MLP_NT / 16, 64, 0, L.s>>>(L.b.mlp_wb, L.b.mlp_w, rc, MLP_RLO, h, (float2*)L.b.mlp_tab);
  else
    k_mlp_tab64<<<(MLP_NT + 63) / 64, 64, 0, L.s>>>(L.b.mlp_w, rc, MLP_RLO, h, (double2*)L.b.mlp_tab);
}

void launch_mlp_grad(const Launch& L, double rc) {
  const double h = (rc - MLP_RLO) / (MLP_NT - 1);
  k_mlp_gsum<<<(MLP_NG + 255) / 256, 256, 0, L.s>>>(L.b.mlp_slot, pair_blocks(L.p), L.b.mlp_G);
  k_mlp_backprop<<<MLP_NT / 64, 64, 0, L.s>>>(L.b.mlp_w, L.b.mlp_G, rc, MLP_RLO, h, L.b.mlp_part);
  k_mlp_dwsum<<<(MLP_NW + 255) / 256, 256, 0, L.s>>>(L.b.mlp_part, MLP_NT / 64, L.b.mlp_dW);
}
