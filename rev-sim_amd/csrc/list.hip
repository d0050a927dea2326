// list.hip — cell sort of molecules and the chunked molecule neighbour list.
//
// Role of Molly's find_neighbors / CellListMap (called at water/water.jl:1025,1038,1488,1504).
// Membership is by molecule: j is listed for i when the O-O minimum-image distance at build time
// is <= r_list = r_c + 2*h_ext + skin (FP64 rule, identical to the oracle's), which contains every
// atom pair within r_c as long as each O moves less than skin/2 and each O-H stays below h_ext
// until the next build (both checked every step by the integrator kernels, which evaluate the
// bonded terms).
//
// The list is full (both i->j and j->i) so each molecule's forces are summed by its own lanes:
// no atomics, deterministic. CH lanes cooperate on one molecule; lane c owns sub-list c, built
// here by the same lane mapping the pair kernel uses (gtid = i*CH + c), in ELL-by-wave layout so
// that each of the pair kernel's list reads is one coalesced 128-byte row (16-bit entries, at most
// 65,536 molecules) or 256-byte row (32-bit).
//
// Each molecule's list is split into three shells by the O-O distance at the build: which atom
// pairs of a listed molecule can come within r_c before the next build follows from it (O-O only
// up to r_c + skin, O-H only up to r_c + h_ext + skin), so the pair kernels evaluate nine, eight or
// four atom pairs per entry, shell by shell, each shell in its own loop.
#include "internal.h"

namespace {

// Cell of each molecule's O and the per-cell counts. Counts go through a block-private LDS
// histogram (molecules of one block share few cells) and one global atomic per nonzero bin.
// Builds that follow an integrator step skip this kernel: the step files the molecules itself.
constexpr int LDS_CELLS = 8192;

__global__ __launch_bounds__(RS_BLOCK) void k_cells_assign(SysParams p, const double* __restrict__ x,
                                                           int* cell, int* count) {
  __shared__ int hist[LDS_CELLS];
  const bool lds = p.ncell <= LDS_CELLS;
  if (lds)
    for (int c = threadIdx.x; c < p.ncell; c += RS_BLOCK) hist[c] = 0;
  __syncthreads();
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < p.M) {
    const int c = p.brute ? 0 : cell_index(p, x[9 * m], x[9 * m + 1], x[9 * m + 2]);
    cell[m] = c;
    if (lds) atomicAdd(&hist[c], 1);
    else atomicAdd(&count[c], 1);
  }
  __syncthreads();
  if (lds)
    for (int c = threadIdx.x; c < p.ncell; c += RS_BLOCK)
      if (hist[c]) atomicAdd(&count[c], hist[c]);
}

// Exclusive scan of cell counts (one 1,024-thread block); clears the fill counters, and the
// counts for the next build (they start at zero when allocated). Thread t takes SCAN_PER
// consecutive cells per round, loaded together; waves scan by shuffles, the block through LDS.
constexpr int SCAN_BLOCK = 1024, SCAN_PER = 4;

__global__ __launch_bounds__(SCAN_BLOCK) void k_cells_scan(int ncell, int* count, int* start, int* fill) {
  __shared__ int wsum[SCAN_BLOCK / RS_WAVE];
  __shared__ int carry;
  const int t = threadIdx.x, lane = t & (RS_WAVE - 1), w = t / RS_WAVE;
  if (t == 0) carry = 0;
  for (int base = 0; base < ncell; base += SCAN_BLOCK * SCAN_PER) {
    const int c0 = base + t * SCAN_PER;
    int v[SCAN_PER], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k) {
      v[k] = c0 + k < ncell ? count[c0 + k] : 0;
      s += v[k];
    }
    int incl = s;
#pragma unroll
    for (int o = 1; o < RS_WAVE; o <<= 1) {
      const int u = __shfl_up(incl, o, RS_WAVE);
      if (lane >= o) incl += u;
    }
    __syncthreads();  // carry of the previous round is final; wsum free
    if (lane == RS_WAVE - 1) wsum[w] = incl;
    __syncthreads();
    int before = carry;
    for (int q = 0; q < w; ++q) before += wsum[q];
    int acc = before + incl - s;
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k)
      if (c0 + k < ncell) {
        start[c0 + k] = acc;
        fill[c0 + k] = 0;
        count[c0 + k] = 0;
        acc += v[k];
      }
    __syncthreads();  // every thread has read carry
    if (t == SCAN_BLOCK - 1) carry = acc;
  }
  __syncthreads();
  if (t == 0) start[ncell] = carry;
}

// Places each molecule into its cell's range (any order inside the cell; k_cells_order then sorts
// by molecule index). Block-private LDS offsets, one global atomic per nonzero bin.
__global__ __launch_bounds__(RS_BLOCK) void k_cells_scatter(SysParams p, const int* cell,
                                                            const int* start, int* fill, int* tmp) {
  __shared__ int hist[LDS_CELLS];
  __shared__ int base[LDS_CELLS];
  const bool lds = p.ncell <= LDS_CELLS;
  if (lds)
    for (int c = threadIdx.x; c < p.ncell; c += RS_BLOCK) hist[c] = 0;
  __syncthreads();
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = m < p.M ? cell[m] : 0;
  int local = 0;
  if (m < p.M) {
    if (lds) local = atomicAdd(&hist[c], 1);
    else tmp[start[c] + atomicAdd(&fill[c], 1)] = m;
  }
  __syncthreads();
  if (lds) {
    for (int q = threadIdx.x; q < p.ncell; q += RS_BLOCK)
      if (hist[q]) base[q] = atomicAdd(&fill[q], hist[q]);
    __syncthreads();
    if (m < p.M) tmp[start[c] + base[c] + local] = m;
  }
}

// Sorted copies of atom a of molecule m, now at sorted index i: the pair inputs for the current
// permutation.
__device__ __forceinline__ void pack_atom(const SysParams& p, int precision, int i, int a, int m,
                                          const double* __restrict__ x, const double* __restrict__ C,
                                          int* posq, double4* pos64, float4* pu, double4* ud,
                                          double4* O64, int4* Oq) {
  const int t = 3 * i + a;  // sorted atom slot
  const double* xa = x + 9 * (size_t)m + 3 * a;
  const double x0 = xa[0], x1 = xa[1], x2 = xa[2];
  const int q[3] = {(int)quant1(x0, p.qscale[0]), (int)quant1(x1, p.qscale[1]), (int)quant1(x2, p.qscale[2])};
  if (precision == RS_PRECISION_FP32) {
    posq[3 * t] = q[0];
    posq[3 * t + 1] = q[1];
    posq[3 * t + 2] = q[2];
  } else {
    pos64[t] = make_double4(x0, x1, x2, 0.0);
  }
  if (a == 0) {
    O64[i] = make_double4(x0, x1, x2, 0.0);
    Oq[i] = make_int4(q[0], q[1], q[2], 0);
  }
  if (C) {
    const double im = 1.0 / (a == 0 ? p.mO : p.mH);
    const double* ca = C + 9 * (size_t)m + 3 * a;
    const double u[3] = {ca[0] * im, ca[1] * im, ca[2] * im};
    if (precision == RS_PRECISION_FP32)
      write_pair_rec(pu, i, a, q, u);
    else
      ud[t] = make_double4(u[0], u[1], u[2], 0.0);
  }
}

// Orders each cell's members by molecule index (one wave per cell): deterministic sorted order;
// the lane that places a molecule also writes its packed pair inputs (C: accum_C for the adjoint
// records, or null).
__global__ void k_cells_order(SysParams p, int precision, const int* start, const int* tmp, int* order,
                              int* rank, const double* __restrict__ x, const double* __restrict__ C,
                              int* posq, double4* pos64, float4* pu, double4* ud, double4* O64, int4* Oq) {
  int wave = (blockIdx.x * blockDim.x + threadIdx.x) / RS_WAVE;
  int lane = threadIdx.x & (RS_WAVE - 1);
  if (wave >= p.ncell) return;
  int s = start[wave], n = start[wave + 1] - s;
  for (int t = lane; t < n; t += RS_WAVE) {
    int m = tmp[s + t];
    int r = 0;
    for (int u = 0; u < n; ++u) r += (tmp[s + u] < m);
    order[s + r] = m;
    rank[m] = s + r;
#pragma unroll
    for (int a = 0; a < 3; ++a) pack_atom(p, precision, s + r, a, m, x, C, posq, pos64, pu, ud, O64, Oq);
  }
}

template <int CH, bool L16>
__global__ __launch_bounds__(RS_BLOCK) void k_list_build(SysParams p, const double4* __restrict__ O64,
                                                         const int4* __restrict__ Oq, int* tmp,
                                                         typename ListEntry<L16>::type* nbr, int4* ncnt,
                                                         int* flags) {
  // All-pairs build for boxes too small for cells. Pass 1: lane c tests every CH-th molecule and
  // keeps the accepted ones in discovery order in its slot column of tmp. Everything goes to
  // shell 0 (all nine atom pairs evaluated).
  const int gtid = blockIdx.x * RS_BLOCK + threadIdx.x;
  const int i = gtid / CH, c = gtid % CH;
  const bool act = i < p.M;
  const int lane = gtid & (RS_WAVE - 1);
  const size_t wbase = (size_t)(gtid / RS_WAVE) * p.capc * RS_WAVE;
  int* tout = tmp + wbase + lane;
  int n = 0;
  if (act) {
    const double4 oi = O64[i];
    const int4 qi = Oq[i];
    auto test = [&](int j) {
      if (j == i) return;
      // FP32 screen on quantized positions; the FP64 rule decides inside the uncertainty band
      const int4 qj = Oq[j];
      const float r2f = r2f_of(qdiff((uint32_t)qi.x, (uint32_t)qj.x, p.unitf[0]),
                               qdiff((uint32_t)qi.y, (uint32_t)qj.y, p.unitf[1]),
                               qdiff((uint32_t)qi.z, (uint32_t)qj.z, p.unitf[2]));
      bool in = r2f < p.rlo2f;
      if (!in && r2f <= p.rhi2f) {
        const double4 oj = O64[j];
        const double dx = mic1(oj.x - oi.x, p.L[0], p.invL[0]);
        const double dy = mic1(oj.y - oi.y, p.L[1], p.invL[1]);
        const double dz = mic1(oj.z - oi.z, p.L[2], p.invL[2]);
        in = r2_of(dx, dy, dz) <= p.rlist2;
      }
      if (in) {
        if (n < p.capc) tout[(size_t)n * RS_WAVE] = j;
        ++n;
      }
    };
    for (int j = c; j < p.M; j += CH) test(j);
  }
  // pass 2 reads entries other lanes of the wave wrote in pass 1
  __threadfence_block();
  // Pass 2: the molecule's accepted candidates in (lane, discovery) order get consecutive global
  // ranks g = (exclusive scan of the lane counts) + k and are dealt round-robin: rank g goes to
  // sub-list g % CH, slot g / CH. Sub-lists then differ in length by at most one, so the lanes of a
  // wave run the same number of pair iterations.
  int incl = n;
#pragma unroll
  for (int o = 1; o < CH; o <<= 1) {
    const int t = __shfl_up(incl, o, CH);
    if (c >= o) incl += t;
  }
  const int T = __shfl(incl, CH - 1, CH);
  const bool ovf = n > p.capc || T > CH * p.capc;
  // Written slot row by slot row so that every store is a full coalesced row: the lane filling
  // rank g = slot*CH + c finds the discovering lane d (the last with inclusive count <= g) among
  // the molecule's CH prefix sums and reads its entry g - first_d from the scratch.
  int incl_of[CH];
#pragma unroll
  for (int d = 0; d < CH; ++d) incl_of[d] = __shfl(incl, d, CH);
  int mine = (T - c + CH - 1) / CH;
  mine = mine < p.capc ? mine : p.capc;
  if (act && !ovf) {
    const int* src0 = tmp + wbase + (lane - c);
    for (int slot = 0; slot < mine; ++slot) {
      const int g = slot * CH + c;
      int d = 0, before = 0;
#pragma unroll
      for (int q = 0; q < CH - 1; ++q)
        if (incl_of[q] <= g) {
          d = q + 1;
          before = incl_of[q];
        }
      nbr[list_base(p, gtid, 0) + (size_t)slot * RS_WAVE] =
          (typename ListEntry<L16>::type)src0[(size_t)(g - before) * RS_WAVE + d];
    }
  }
  if (act) {
    if (ovf) atomicOr(&flags[FLAG_LIST_OVERFLOW], 1);
    ncnt[gtid] = make_int4(mine, 0, 0, 0);
  }
}

// Cell-centric build (the cell-stencil path): one block per home cell. The stencil's quantized O
// positions are staged in LDS in chunks, in stencil order (rows (dy, dz), then cells along x, then
// sorted index); each wave takes the cell's molecules in turn and tests 64 staged candidates per
// iteration. Per shell, a wave ballot ranks the accepted ones in stencil order, so candidate g of
// molecule i's shell s goes straight to its dealt place, sub-list g % CH, slot g / CH of shell s in
// the pair kernels' ELL-by-wave layout: no scratch, deterministic, sub-lists balanced to within one
// entry in every shell. The shell test errs only towards lower shells (more pairs evaluated).
constexpr int LC_CHUNK = 1280;  // staged candidates per pass (a 98k-atom stencil holds ~1,200)
constexpr int LC_MAXH = 256;    // molecules per home cell
constexpr int LC_MAXR = 2 * RS_WAVE;  // stencil ranges: (2 cs + 1)^2 <= 25 rows, two slots each

template <int CH, bool L16>
__global__ __launch_bounds__(RS_BLOCK) void k_list_cells(SysParams p, const int* __restrict__ start,
                                                         const double4* __restrict__ O64,
                                                         const int4* __restrict__ Oq,
                                                         typename ListEntry<L16>::type* nbr,
                                                         int4* ncnt, int* flags) {
  typedef typename ListEntry<L16>::type LT;
  // quantized O positions, one array per axis, and the candidates' sorted indices (22 KB of LDS
  // per block in all with 16-bit indices: seven blocks per CU; 24.6 KB and six with 32-bit ones)
  __shared__ int sqx[LC_CHUNK], sqy[LC_CHUNK], sqz[LC_CHUNK];
  __shared__ LT sj[LC_CHUNK];  // sorted indices, at the list's entry width
  __shared__ int rbeg[LC_MAXR], roff[LC_MAXR + 1];
  __shared__ int Ts[LIST_SHELLS][LC_MAXH];
  __shared__ int nr;
  // home cell, XCD-aware: the grid is rounded up to a multiple of 8 and the padding exits
  const int h = xcd_block(blockIdx.x, gridDim.x);
  if (h >= p.ncell) return;
  const int h0 = start[h], nh = start[h + 1] - h0;
  if (nh == 0) return;  // uniform for the block
  if (nh > LC_MAXH) {
    if (threadIdx.x == 0) atomicOr(&flags[FLAG_CELL_OVERFLOW], 1);
    return;
  }
  if (threadIdx.x < RS_WAVE) {
    // lane r < (2S+1)^2 takes stencil row r = (dy, dz): one contiguous molecule range, or two
    // where the row wraps around the box in x; offsets by a prefix sum over the wave
    const int S = p.cs, W = 2 * S + 1, lane = threadIdx.x;
    int b0 = 0, n0 = 0, b1 = 0, n1 = 0;
    if (lane < W * W) {
      const int cx = h % p.nc[0], cy = (h / p.nc[0]) % p.nc[1], cz = h / (p.nc[0] * p.nc[1]);
      const int dy = lane % W - S, dz = lane / W - S;
      int ey = cy + dy, ez = cz + dz;
      ey += (ey < 0) ? p.nc[1] : (ey >= p.nc[1] ? -p.nc[1] : 0);
      ez += (ez < 0) ? p.nc[2] : (ez >= p.nc[2] ? -p.nc[2] : 0);
      const int row = (ez * p.nc[1] + ey) * p.nc[0];
      const int x0 = cx - S, x1 = cx + S;
      int a0, a1, c0 = 0, c1 = -1;  // cells a0..a1 and c0..c1 of the row
      if (x0 < 0) {
        a0 = row + x0 + p.nc[0];
        a1 = row + p.nc[0] - 1;
        c0 = row;
        c1 = row + x1;
      } else if (x1 >= p.nc[0]) {
        a0 = row + x0;
        a1 = row + p.nc[0] - 1;
        c0 = row;
        c1 = row + x1 - p.nc[0];
      } else {
        a0 = row + x0;
        a1 = row + x1;
      }
      b0 = start[a0];
      n0 = start[a1 + 1] - b0;
      if (c1 >= c0) {
        b1 = start[c0];
        n1 = start[c1 + 1] - b1;
      }
    }
    int incl = n0 + n1;
#pragma unroll
    for (int o = 1; o < RS_WAVE; o <<= 1) {
      const int t = __shfl_up(incl, o, RS_WAVE);
      if (lane >= o) incl += t;
    }
    const int excl = incl - n0 - n1;
    // ranges in row order, the wrapped row's two pieces in x order: stencil order
    rbeg[2 * lane] = b0;
    roff[2 * lane] = excl;
    rbeg[2 * lane + 1] = b1;
    roff[2 * lane + 1] = excl + n0;
    if (lane == RS_WAVE - 1) {
      roff[2 * RS_WAVE] = incl;
      nr = 2 * W * W;
    }
  }
  for (int m = threadIdx.x; m < nh; m += RS_BLOCK) Ts[0][m] = Ts[1][m] = Ts[2][m] = 0;
  __syncthreads();
  const int total = roff[nr];
  const int wave = threadIdx.x / RS_WAVE, lane = threadIdx.x & (RS_WAVE - 1);
  for (int base = 0; base < total; base += LC_CHUNK) {
    const int nchunk = min(LC_CHUNK, total - base);
    // one range per wave at a time, lanes along it (no search for an element's range)
    for (int r = wave; r < nr; r += RS_BLOCK / RS_WAVE) {
      const int b = rbeg[r], o = roff[r] - base, len = roff[r + 1] - roff[r];
      for (int t = lane; t < len; t += RS_WAVE) {
        const int gi = o + t;
        if (gi >= 0 && gi < nchunk) {
          sj[gi] = (LT)(b + t);
          const int4 q = Oq[b + t];
          sqx[gi] = q.x;
          sqy[gi] = q.y;
          sqz[gi] = q.z;
        }
      }
    }
    __syncthreads();
    for (int m = wave; m < nh; m += RS_BLOCK / RS_WAVE) {
      const int i = h0 + m;
      const int4 qi = Oq[i];
      const double4 oi = O64[i];
      const size_t g0 = (size_t)i * CH;
      LT* out0 = nbr + list_base(p, g0, 0);
      const unsigned shell_words = (unsigned)p.capc * RS_WAVE;  // one shell of a wave's list
      unsigned T0 = Ts[0][m], T1 = Ts[1][m], T2 = Ts[2][m];
      const bool cubic = p.rlo2q >= 0.0f;  // uniform: quantized shells and screen
#pragma unroll 2
      for (int c0 = 0; c0 < nchunk; c0 += RS_WAVE) {
        const int t = c0 + lane;
        // read at a clamped index and mask, so the LDS reads do not wait on a branch
        const int tc = t < nchunk ? t : nchunk - 1;
        const int j = (int)sj[tc];
        // FP32 screen in quantized units (cubic box: one unit for all axes): accept below rlo2q,
        // reject above rhi2q, the FP64 rule decides in between (a non-cubic box has rlo2q < 0
        // and rhi2q = inf: every candidate goes to the FP64 rule, shells included)
        const float r2q = r2f_of(qdiffq((uint32_t)qi.x, (uint32_t)sqx[tc]),
                                 qdiffq((uint32_t)qi.y, (uint32_t)sqy[tc]),
                                 qdiffq((uint32_t)qi.z, (uint32_t)sqz[tc]));
        const bool lt = t < nchunk, ne = j != i;
        const bool ql = r2q < p.rlo2q, q0 = r2q <= p.sh2q[0], q01 = r2q <= p.sh2q[1];
        // Wave masks, each the ballot of one compare (the compare writes it to a scalar pair
        // directly); the per-lane flags are the compares themselves.
        const unsigned long long vm = __ballot(lt) & __ballot(ne);
        unsigned long long ba = vm & __ballot(ql);
        unsigned long long m0 = __ballot(q0), m01 = __ballot(q01);
        const unsigned long long bandm = vm & ~ba & __ballot(r2q <= p.rhi2q);
        bool in = lt && ne && ql, s0 = q0, s01 = q01;
        if (bandm) {  // uniform, rare in a cubic box: FP64 decisions patch masks and flags
          bool in64 = false, s064 = false, s0164 = false;
          if ((bandm >> lane) & 1ull) {
            const double4 oj = O64[j];
            const double dx = mic1(oj.x - oi.x, p.L[0], p.invL[0]);
            const double dy = mic1(oj.y - oi.y, p.L[1], p.invL[1]);
            const double dz = mic1(oj.z - oi.z, p.L[2], p.invL[2]);
            const double r2 = r2_of(dx, dy, dz);
            in64 = r2 <= p.rlist2;
            s064 = r2 <= p.sh2[0];
            s0164 = r2 <= p.sh2[1];
          }
          ba |= __ballot(in64);
          in = in || in64;
          if (!cubic) {
            m0 = __ballot(s064);
            m01 = __ballot(s0164);
            s0 = s064;
            s01 = s0164;
          }
        }
        const unsigned long long b0 = ba & m0, b01 = ba & m01;
        const unsigned long long m1 = b01 & ~b0, m2 = ba & ~b01;
        if (in) {
          // rank in the lane's shell: the shell mask's bits below the lane (mbcnt) plus the
          // shell's entries so far; candidate g goes to sub-list g % CH, slot g / CH
          const unsigned g0 = __builtin_amdgcn_mbcnt_hi((unsigned)(b0 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)b0, T0));
          const unsigned g1 = __builtin_amdgcn_mbcnt_hi((unsigned)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m1, T1));
          const unsigned g2 = __builtin_amdgcn_mbcnt_hi((unsigned)(m2 >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m2, T2));
          const unsigned g = s0 ? g0 : (s01 ? g1 : g2);
          const unsigned slot = g / CH;
          const unsigned shoff = s0 ? 0u : (s01 ? shell_words : 2u * shell_words);
          if (slot < (unsigned)p.capc) out0[shoff + slot * RS_WAVE + (g % CH)] = (LT)j;
        }
        T0 += __popcll(b0);
        T1 += __popcll(m1);
        T2 += __popcll(m2);
      }
      if (lane == 0) {
        Ts[0][m] = T0;
        Ts[1][m] = T1;
        Ts[2][m] = T2;
      }
    }
    __syncthreads();
  }
  for (int m = wave; m < nh; m += RS_BLOCK / RS_WAVE) {
    const int i = h0 + m;
    int mine[LIST_SHELLS];
    bool ovf = false;
    for (int s = 0; s < LIST_SHELLS; ++s) {
      const int T = Ts[s][m];
      mine[s] = min((T - lane + CH - 1) / CH, p.capc);
      ovf = ovf || T > CH * p.capc;
    }
    if (lane < CH) ncnt[(size_t)i * CH + lane] = make_int4(mine[0], mine[1], mine[2], 0);
    if (lane == 0 && ovf) atomicOr(&flags[FLAG_LIST_OVERFLOW], 1);
  }
}

}  // namespace

int pair_blocks(const SysParams& p) { return (p.M * p.CH + RS_BLOCK - 1) / RS_BLOCK; }

void launch_list_build(const Launch& L, bool with_u, bool filed) {
  const SysParams& p = L.p;
  const DevBufs& b = L.b;
  int gm = (p.M + RS_BLOCK - 1) / RS_BLOCK;
  if (!filed) k_cells_assign<<<gm, RS_BLOCK, 0, L.s>>>(p, b.x, b.cell, b.cell_count);
  k_cells_scan<<<1, SCAN_BLOCK, 0, L.s>>>(p.ncell, b.cell_count, b.cell_start, b.cell_fill);
  k_cells_scatter<<<gm, RS_BLOCK, 0, L.s>>>(p, b.cell, b.cell_start, b.cell_fill, b.tmp);
  int gw = (p.ncell * RS_WAVE + RS_BLOCK - 1) / RS_BLOCK;
  k_cells_order<<<gw, RS_BLOCK, 0, L.s>>>(p, L.precision, b.cell_start, b.tmp, b.order, b.rank, b.x,
                                          with_u ? b.C : nullptr, b.posq, b.pos64, b.pu, b.ud, b.O64, b.Oq);
  // p.CH is 8 or 16 (rs_create), the entry width follows p.list16
  auto build = [&](auto ch, auto l16) {
    constexpr int CH = decltype(ch)::value;
    constexpr bool L16 = decltype(l16)::value;
    typedef typename ListEntry<L16>::type LT;
    if (!p.brute)
      k_list_cells<CH, L16><<<(p.ncell + RS_XCDS - 1) / RS_XCDS * RS_XCDS, RS_BLOCK, 0, L.s>>>(
          p, b.cell_start, b.O64, b.Oq, (LT*)b.nbr, b.ncnt, b.flags);
    else
      k_list_build<CH, L16><<<pair_blocks(p), RS_BLOCK, 0, L.s>>>(p, b.O64, b.Oq, b.nbr_tmp, (LT*)b.nbr,
                                                                   b.ncnt, b.flags);
  };
  auto by_width = [&](auto ch) {
    if (p.list16) build(ch, std::true_type());
    else build(ch, std::false_type());
  };
  if (p.CH == 8) by_width(std::integral_constant<int, 8>());
  else by_width(std::integral_constant<int, 16>());
}
