// loss.hip — O-O RDF loss and O mean-squared-displacement loss with their position gradients.
//
// RDF (rdf_differentiable / loss_rdf_part / rrule, water/water.jl:631-736), re-binned per the
// config: bins x_b = (b+1/2)Δ, Δ = rmax/bins; each O-O distance r spreads weights
// e_b = exp(-(x_b - r)^2/width) over the bins within ±window, normalised by their sum (Wang 2023
// eq. 4/6 as restated by the reference); g_b = 2 pre/x_b^2 Σ_{i<j} φ_b(r_ij),
// pre = V/(4π N_O^2 Δ); L = Σ_b (g_b - target_b)^2.
// Bin sums are 64-bit fixed point (2^-32 per unit weight), accumulated with LDS and global integer
// atomics: order-independent and identical to the oracle bit for bit (same r, same exp sequence).
// Distances are FP64 in both precisions (the FP64 state): the narrow smoothing makes φ_b about
// 160/nm sensitive to r, and FP32 distances put ~1e-5 relative noise into the histogram.
//
// MSD (calculate_D / loss_fn_final, diffusion/diffusion.jl:861-911) on the O atoms of 2K frames:
// D = (1/6K) Σ_s MSD(s, s+K)/(K Δt_frame) (Einstein relation at lag K frames, K origins) with
// unwrapped displacements from summed minimum-image
// increments; L = w 1e18 (1e-6 D - target)^2; dL/dx_{frame} from the telescoping increments.
#include "internal.h"

namespace {

struct RdfArgs {
  int nb;
  double rmax, width, window, wgt, pre;
};

// Window of bins within ±window of one distance.
__device__ __forceinline__ void window64(const RdfArgs& r, double rd, int& lo, int& hi) {
#pragma clang fp contract(off)
  const double invd = r.nb / r.rmax, c = r.window;
  lo = (int)ceil((rd - c) * invd - 0.5);
  hi = (int)floor((rd + c) * invd - 0.5);
  if (lo < 0) lo = 0;
  if (hi > r.nb - 1) hi = r.nb - 1;
}

__device__ __forceinline__ void contrib64(const RdfArgs& r, double rd, unsigned long long* lh) {
#pragma clang fp contract(off)
  int lo, hi;
  window64(r, rd, lo, hi);
  if (lo > hi) return;
  const double dx = r.rmax / r.nb, iw = 1.0 / r.width;
  double den = 0.0;
  for (int b = lo; b <= hi; ++b) {
    double t = ((double)b + 0.5) * dx - rd;
    den += rs_exp(-(t * t) * iw);
  }
  for (int b = lo; b <= hi; ++b) {
    double t = ((double)b + 0.5) * dx - rd;
    double e = rs_exp(-(t * t) * iw);
    long long v = __double2ll_rn((e / den) * 4294967296.0);
    atomicAdd(&lh[b], (unsigned long long)v);
  }
}

#define RDF_MAX_BINS 1024

template <int CH, bool L16>
__global__ __launch_bounds__(RS_BLOCK) void k_rdf_hist(SysParams p, RdfArgs r,
                                                        const typename ListEntry<L16>::type* __restrict__ nbr,
                                                        const int4* __restrict__ ncnt,
                                                        const int* __restrict__ order,
                                                        const double* __restrict__ x,
                                                        long long* hist) {
  __shared__ unsigned long long lh[RDF_MAX_BINS];
  for (int b = threadIdx.x; b < r.nb; b += RS_BLOCK) lh[b] = 0ull;
  __syncthreads();
  const int gtid = blockIdx.x * RS_BLOCK + threadIdx.x;
  const int i = gtid / CH;
  if (i < p.M) {
    const int4 nc4 = ncnt[gtid];
    for (int s = 0; s < LIST_SHELLS; ++s) {
      const int n = s == 0 ? nc4.x : (s == 1 ? nc4.y : nc4.z);
      const typename ListEntry<L16>::type* L = nbr + list_base(p, gtid, s);
      const double* xi = x + 9 * (size_t)order[i];
      for (int k = 0; k < n; ++k) {
        const int j = L[(size_t)k * RS_WAVE];
        if (j <= i) continue;
        const double* xj = x + 9 * (size_t)order[j];
        const double dx = mic1(xj[0] - xi[0], p.L[0], p.invL[0]);
        const double dy = mic1(xj[1] - xi[1], p.L[1], p.invL[1]);
        const double dz = mic1(xj[2] - xi[2], p.L[2], p.invL[2]);
        contrib64(r, sqrt(r2_of(dx, dy, dz)), lh);
      }
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < r.nb; b += RS_BLOCK)
    if (lh[b]) atomicAdd((unsigned long long*)&hist[b], lh[b]);
}

// g_b, L and dL/dg_b (one block). The loss sum runs over bins in order, as on the host.
__global__ void k_rdf_final(RdfArgs r, const long long* hist, const double* target, double* g,
                            double* G, double* scal, double* loss_dst) {
#pragma clang fp contract(off)
  const double dx = r.rmax / r.nb;
  for (int b = threadIdx.x; b < r.nb; b += blockDim.x) {
    double xb = ((double)b + 0.5) * dx;
    double gb = 2.0 * r.pre / (xb * xb) * ((double)hist[b] / 4294967296.0);
    g[b] = gb;
    G[b] = 2.0 * (gb - target[b]) * 2.0 * r.pre / (xb * xb);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double L = 0.0;
    for (int b = 0; b < r.nb; ++b) {
      double d = g[b] - target[b];
      L += d * d;
    }
    scal[1] = L;
    if (loss_dst) {
      // loss_dst is the scalar block: [0] the trajectory loss, [8] its RDF term
      loss_dst[0] += r.wgt * L;
      loss_dst[8] += r.wgt * L;
    }
  }
}

// dL/dx_O (times the block weight) by owner-computes over the full list, FP64.
template <int CH, bool L16>
__global__ __launch_bounds__(RS_BLOCK) void k_rdf_grad(SysParams p, RdfArgs r,
                                                        const typename ListEntry<L16>::type* __restrict__ nbr,
                                                        const int4* __restrict__ ncnt,
                                                        const int* __restrict__ order,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ G,
                                                        double* grad) {
  const int gtid = blockIdx.x * RS_BLOCK + threadIdx.x;
  const int i = gtid / CH, c = gtid % CH;
  const bool act = i < p.M;
  double acc[3] = {0.0, 0.0, 0.0};
  const int mi = act ? order[i] : 0;
  if (act) {
    const double* xi = x + 9 * (size_t)mi;
    const int4 nc4 = ncnt[gtid];
    for (int s = 0; s < LIST_SHELLS; ++s) {
      const int n = s == 0 ? nc4.x : (s == 1 ? nc4.y : nc4.z);
      const typename ListEntry<L16>::type* L = nbr + list_base(p, gtid, s);
      const double dxb = r.rmax / r.nb, iw = 1.0 / r.width, rmaxw = r.rmax + r.window;
      for (int k = 0; k < n; ++k) {
        const int j = L[(size_t)k * RS_WAVE];
        const double* xj = x + 9 * (size_t)order[j];
        double d[3];
        for (int t = 0; t < 3; ++t) d[t] = mic1(xj[t] - xi[t], p.L[t], p.invL[t]);
        const double rr = sqrt(r2_of(d[0], d[1], d[2]));
        if (rr > rmaxw + 0.01) continue;
        int lo, hi;
        window64(r, rr, lo, hi);
        if (lo > hi) continue;
        double den = 0.0, dden = 0.0, s1 = 0.0, s2 = 0.0;
        for (int bb = lo; bb <= hi; ++bb) {
          double t = ((double)bb + 0.5) * dxb - rr;
          double e = rs_exp(-(t * t) * iw);
          double de = e * 2.0 * t * iw;
          den += e;
          dden += de;
          s1 += G[bb] * de;
          s2 += G[bb] * e;
        }
        const double dLdr = s1 / den - dden * s2 / (den * den);
        for (int t = 0; t < 3; ++t) acc[t] -= dLdr * d[t] / rr;
      }
    }
  }
  for (int t = 0; t < 3; ++t) acc[t] = chunk_sum<CH>(acc[t]);
  if (act && c == 0) {
    for (int t = 0; t < 3; ++t) {
      grad[9 * (size_t)mi + t] = r.wgt * acc[t];
      grad[9 * (size_t)mi + 3 + t] = 0.0;
      grad[9 * (size_t)mi + 6 + t] = 0.0;
    }
  }
}

__global__ void k_zero_ll(long long* a, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0;
}
__global__ void k_zero_d(double* a, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0.0;
}

__global__ void k_frame_copy(SysParams p, const double* x, double* frames, int f) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= p.M) return;
  for (int k = 0; k < 3; ++k) frames[(size_t)f * 3 * p.M + 3 * m + k] = x[9 * (size_t)m + k];
}

// Per origin s and molecule m, the unwrapped displacement Δ over K frame intervals and |Δ|^2 (no
// contraction). The |Δ|^2 are then summed in a fixed order that the oracle follows too, so D agrees
// bit for bit: in index order within chunks of MSD_CHUNK molecules (k_msd_chunks, one thread per
// chunk and origin), then the chunk sums in order (k_msd_final).
__global__ void k_msd_del(SysParams p, int K, const double* frames, double* del, double* sq_out) {
#pragma clang fp contract(off)
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= p.M) return;
  const size_t F = 3 * (size_t)p.M;
  for (int s = 0; s < K; ++s) {
    double a[3] = {0.0, 0.0, 0.0};
    for (int f = s; f < s + K; ++f)
      for (int k = 0; k < 3; ++k)
        a[k] += mic1(frames[(f + 1) * F + 3 * m + k] - frames[f * F + 3 * m + k], p.L[k], p.invL[k]);
    double sq = 0.0;
    for (int k = 0; k < 3; ++k) {
      del[((size_t)s * p.M + m) * 3 + k] = a[k];
      sq += a[k] * a[k];
    }
    sq_out[(size_t)s * p.M + m] = sq;
  }
}

constexpr int MSD_CHUNK = 256;

__global__ void k_msd_chunks(SysParams p, int K, const double* __restrict__ sq, double* part) {
  const int nch = (p.M + MSD_CHUNK - 1) / MSD_CHUNK;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // (origin, chunk)
  if (t >= K * nch) return;
  const int s = t / nch, c = t % nch;
  const double* q = sq + (size_t)s * p.M;
  const int m1 = min(p.M, (c + 1) * MSD_CHUNK);
  double a = 0.0;
#pragma unroll 8
  for (int m = c * MSD_CHUNK; m < m1; ++m) a += q[m];
  part[t] = a;
}

// D = Σ_s (Σ_chunks (Σ_m |Δ_sm|^2) / M) / lag / (6 K): one thread per origin sums its chunks in
// order (once per trajectory), thread 0 the origins in order.
__global__ void k_msd_final(SysParams p, int K, double lag, double w, double target,
                            const double* part, double* scal, double* loss_dst) {
#pragma clang fp contract(off)
  __shared__ double msd_s[64];
  const int nch = (p.M + MSD_CHUNK - 1) / MSD_CHUNK;
  if ((int)threadIdx.x < K) {
    double msd = 0.0;
    const double* q = part + (size_t)threadIdx.x * nch;
    for (int c = 0; c < nch; ++c) msd += q[c];
    msd_s[threadIdx.x] = msd / p.M;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double Dsum = 0.0;
  for (int s = 0; s < K; ++s) Dsum += msd_s[s] / lag;
  double D = Dsum / (6.0 * K);
  double diff = 1e-6 * D - target;
  double L = w * 1e18 * diff * diff;
  scal[2] = L;
  scal[3] = D;
  scal[5] = (w * 1e18 * 2.0 * diff * 1e-6) * 2.0 / (6.0 * K * lag * p.M);
  if (loss_dst) *loss_dst += L;
}

__global__ void k_msd_grad(SysParams p, int K, const double* del, const double* scal, double* out) {
  const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;  // (f, m, k)
  const size_t F = 3 * (size_t)p.M;
  if (t >= 2 * (size_t)K * F) return;
  const int f = (int)(t / F);
  const size_t r = t - f * F;
  const double c = scal[5];
  out[t] = f >= K ? c * del[(size_t)(f - K) * F + r] : -(c * del[(size_t)f * F + r]);
}

__global__ void k_msd_inject(SysParams p, const double* mg, int f, double* grad) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= p.M) return;
  for (int k = 0; k < 3; ++k) grad[9 * (size_t)m + k] += mg[(size_t)f * 3 * p.M + 3 * m + k];
}

}  // namespace

void launch_rdf(const Launch& L, int nb, double rmax, double width, double window, double wgt,
                bool grad, double* loss_dst) {
  const SysParams& p = L.p;
  RdfArgs r;
  r.nb = nb;
  r.rmax = rmax;
  r.width = width;
  r.window = window;
  r.wgt = wgt;
  const double V = p.L[0] * p.L[1] * p.L[2];
  const double dx = rmax / nb;
  r.pre = V / (4.0 * M_PI * (double)p.M * (double)p.M * dx);
  k_zero_ll<<<(nb + RS_BLOCK - 1) / RS_BLOCK, RS_BLOCK, 0, L.s>>>(L.b.hist, nb);
  int g = pair_blocks(p);
  // p.CH is 8 or 16, the entry width follows p.list16
  auto run = [&](auto ch, auto l16) {
    constexpr int CH = decltype(ch)::value;
    constexpr bool L16 = decltype(l16)::value;
    typedef typename ListEntry<L16>::type LT;
    const LT* nbr = (const LT*)L.b.nbr;
    k_rdf_hist<CH, L16><<<g, RS_BLOCK, 0, L.s>>>(p, r, nbr, L.b.ncnt, L.b.order, L.b.x, L.b.hist);
    k_rdf_final<<<1, RS_BLOCK, 0, L.s>>>(r, L.b.hist, L.b.target, L.b.rdf_g, L.b.rdfG, L.b.scal, loss_dst);
    if (grad) k_rdf_grad<CH, L16><<<g, RS_BLOCK, 0, L.s>>>(p, r, nbr, L.b.ncnt, L.b.order, L.b.x, L.b.rdfG, L.b.grad);
  };
  auto by_width = [&](auto ch) {
    if (p.list16) run(ch, std::true_type());
    else run(ch, std::false_type());
  };
  if (p.CH == 8) by_width(std::integral_constant<int, 8>());
  else by_width(std::integral_constant<int, 16>());
}

void launch_zero(const Launch& L, double* a, size_t n) {
  k_zero_d<<<(unsigned)((n + RS_BLOCK - 1) / RS_BLOCK), RS_BLOCK, 0, L.s>>>(a, (int)n);
}

void launch_frame_copy(const Launch& L, int f) {
  k_frame_copy<<<(L.p.M + RS_BLOCK - 1) / RS_BLOCK, RS_BLOCK, 0, L.s>>>(L.p, L.b.x, L.b.frames, f);
}

void launch_msd_bufs(const Launch& L, int nf, double lag, double w_msd, double target, const double* frames,
                     double* del, double* sq, double* part, double* scal, double* grad, double* loss_dst) {
  const SysParams& p = L.p;
  const int K = nf / 2;
  k_msd_del<<<(p.M + RS_BLOCK - 1) / RS_BLOCK, RS_BLOCK, 0, L.s>>>(p, K, frames, del, sq);
  const int nch = (p.M + MSD_CHUNK - 1) / MSD_CHUNK;
  k_msd_chunks<<<(K * nch + RS_BLOCK - 1) / RS_BLOCK, RS_BLOCK, 0, L.s>>>(p, K, sq, part);
  k_msd_final<<<1, 64, 0, L.s>>>(p, K, lag, w_msd, target, part, scal, loss_dst);
  size_t tot = (size_t)nf * 3 * p.M;
  k_msd_grad<<<(unsigned)((tot + RS_BLOCK - 1) / RS_BLOCK), RS_BLOCK, 0, L.s>>>(p, K, del, scal, grad);
}

void launch_msd(const Launch& L, int nf, double lag, double w_msd, double target, double* loss_dst) {
  launch_msd_bufs(L, nf, lag, w_msd, target, L.b.frames, L.b.msd_del, L.b.msd_sq, L.b.msd_part, L.b.scal,
                  L.b.msd_grad, loss_dst);
}

void launch_msd_inject(const Launch& L, int f) {
  k_msd_inject<<<(L.p.M + RS_BLOCK - 1) / RS_BLOCK, RS_BLOCK, 0, L.s>>>(L.p, L.b.msd_grad, f, L.b.grad);
}
