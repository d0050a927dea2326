// internal.h — host-side launch interface between the librevsim translation units.
#pragma once

#include "common.h"

// Device buffers of one context. Arrays named "orig" are in the caller's atom order; arrays named
// "sorted" are indexed by the cell-sorted molecule index of the last list build.
struct DevBufs {
  // state, original order, 3N doubles each
  double *x, *v, *F, *h, *A, *B, *C, *D, *grad;
  // cell sort
  int *cell, *cell_count, *cell_start, *cell_fill, *tmp, *order, *rank;
  // sorted pair inputs
  int* posq;       // FP32 mode: quantized positions, 9 ints per molecule (O, H1, H2)
  double4* pos64;  // FP64 mode: positions, 3 per molecule
  // the other buffer of each, written by the pair kernel that ends with the integrator step; the
  // host swaps the pair after such a launch
  int* posq_alt;
  double4* pos64_alt;
  // FP32 mode, adjoint pair inputs: one 80-byte record per molecule, H1, H2, O, each as its
  // quantized position (int bits) then accum_C / m, and two unused words (pair_rec in common.h)
  float4* pu;
  double4* ud;     // FP64 mode: accum_C / m
  double4* O64;    // O position at the last list build, sorted
  int4* Oq;        // the same, quantized (2^32 per box edge), for the list build's FP32 screen
  // molecule list, ELL by wave and distance shell (list_base in common.h), entries of
  // ListEntry<p.list16>; per-lane lengths
  void* nbr;
  int4* ncnt;
  int* nbr_tmp;    // all-pairs list-build scratch: [((gtid/64)*capc + k)*64 + gtid%64]
  int* flags;
  double* slots;   // per-block θ-VJP partials [nblocks * 8]
  double* vjb;     // bonded θ-VJP per molecule [3M]: (k_bond, r0, k_angle)
  double* dth;     // dL/dθ accumulator [8]
  double* scal;    // device scalars: [0] loss, [1] rdf loss (last), [2] msd loss, [3] D, [4] clip scale,
                   // [5] msd gradient factor, [6] enthalpy loss (last), [7] its c', [8] trajectory RDF
                   // term, [9] trajectory enthalpy term
  double* wmax;    // MLP: max|w_O| per step parity, spread over 64 slots (uint64 bits of doubles)
  long long* hist; // RDF bins (fixed point 2^32)
  double *rdfG, *target, *rdf_g;
  double* frames;  // MSD frames [nf * 3M]
  double* msd_del; // [K * M * 3]
  double* msd_sq;     // |Δ|^2 per origin and molecule [K * M]
  double* msd_part;   // their sums over chunks of 256 molecules [K * ceil(M/256)]
  double* msd_grad;   // dL/dframes [nf * 3M]
  double* ext_noise;  // [steps * 3N] or null
  double* hpart;      // h-norm partials
  // radial-basis MLP (mlp.h)
  double* mlp_w;            // weights (BF16 values) as FP64
  unsigned short* mlp_wb;   // weights as BF16 bits
  void* mlp_tab;            // (f, f') per node: float2 (FP32 path) or double2
  double* mlp_slot;         // per pair-block table gradient [blocks * MLP_NG]
  double *mlp_G, *mlp_part, *mlp_dW;
  // potential energy (energy.hip)
  double* Fe;      // -dU/dx of the last energy evaluation, caller order [3N]
  double* eslots;  // per pair-block partial sums of (U, dU/dθ) [blocks * 8]
  double* eout;    // (U, dU/dθ[6]) [8]
  double* rwacc;   // reweighting accumulators [19]: ΣL, ΣdL/dθ, ΣdU/dθ, ΣL dU/dθ
  double* xbak;    // positions before a barostat move [3N]
};

struct Launch {
  SysParams p;
  DevBufs b;
  int precision;
  hipStream_t s;
  int use_mlp;     // MLP pair term on (weights set and tabulated)
  double mlp_rc;   // cutoff of the MLP features (= pair cutoff)
  int parity;      // step parity selecting the max|w_O| slot of the adjoint MLP accumulation
};

// list (filed: the integrator kernel that wrote the positions has already filed each molecule
// into its cell and counted it, so the build starts at the scan)
void launch_list_build(const Launch& L, bool with_u, bool filed = false);
// pair
void launch_pair(const Launch& L, bool adjoint, const double* wsrc, double wscale_O, double wscale_H);
// forward pair kernel followed, in the same kernel, by the forward step n (Philox seed, dt,
// e = exp(-γ dt), s = sqrt(1 - e^2), kT; assign: file each molecule into its cell for the list
// build that follows); writes the new sorted positions to posq_alt / pos64_alt, which the caller
// swaps in
void launch_pair_step(const Launch& L, int step, uint64_t seed, double dt, double e, double s,
                      double kT, bool assign);
void launch_vjp_reduce(const Launch& L);
void launch_bonded(const Launch& L, bool adjoint, const double* csrc, double cO, double cH,
                   const double* wsrc, double wO, double wH);
// integrator (assign: file each molecule into its cell for the list build that follows)
void launch_wrap_state(const Launch& L);
void launch_step_rev(const Launch& L, bool do_r2, bool do_r1, int step, uint64_t seed, double dt,
                     double e, double s, double kT, int nslots, bool clip, bool assign = false);
void launch_seed(const Launch& L, bool reset, double e);
void launch_noise(const Launch& L, int step, uint64_t seed, double kT, double* xi);
void launch_hnorm(const Launch& L, double clip);
// losses
void launch_rdf(const Launch& L, int nb, double rmax, double width, double window, double wgt,
                bool grad, double* loss_dst);
void launch_frame_copy(const Launch& L, int f);
void launch_msd(const Launch& L, int nf, double lag, double w_msd, double target, double* loss_dst);
// The same kernels on caller-given buffers: frames [nf 3M], del [K 3M], sq [K M],
// part [K ceil(M/256)], scal [>= 6] (writes [2] loss, [3] D, [5] gradient factor), grad [nf 3M].
void launch_msd_bufs(const Launch& L, int nf, double lag, double w_msd, double target, const double* frames,
                     double* del, double* sq, double* part, double* scal, double* grad, double* loss_dst);
void launch_msd_inject(const Launch& L, int f);
void launch_zero(const Launch& L, double* a, size_t n);
int pair_blocks(const SysParams& p);
void launch_mlp_tabulate(const Launch& L, double rc);
// energy
void launch_energy(const Launch& L);
void launch_dhvap(const Launch& L, double w, double target, double ugas, double RT, double wgt,
                  bool direct, bool seed);
void launch_reweight_acc(const Launch& L, double w_rdf, bool dh);
void launch_scale_molecules(const Launch& L, double l, const double* Lnew);
void launch_mlp_grad(const Launch& L, double rc);
