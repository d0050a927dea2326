/*
 * revsim.h — C ABI of the MI355X reversible-simulation gradient library (librevsim.so).
 *
 * The library runs the reversible-simulation gradient loop for flexible 3-site water on one
 * gfx950 GPU: forward Langevin MD (cell-sorted molecule neighbour list, one fused kernel for
 * Lennard-Jones + reaction-field Coulomb + harmonic bond/angle forces, velocity-Verlet/Langevin
 * step with counter-based Philox noise) followed by the reverse-time adjoint (Hessian-vector
 * product dF/dx^T.lambda and parameter VJP dF/dtheta^T.lambda in the same fused pass), with the
 * O-O RDF and O mean-squared-displacement losses and their gradients computed on the device.
 *
 * Reference interfaces this ABI replaces (greener-group/rev-sim):
 *   rs_forward          forward_sim(coords, velocities, boundary, noise_seeds, γ, n_steps, n_steps_loss,
 *                                   params, dt, n_threads, inter_type, loss_type, run_grads=false)
 *                       water/water.jl:977-979,1029-1045,1104
 *   rs_reverse          reverse_sim(coords, velocities, boundary, noise_seeds, γ, n_steps, n_steps_loss,
 *                                   n_steps_trunc, params, dt, n_threads, fwd_loggers,
 *                                   clip_norm_threshold, inter_type, loss_type)
 *                       water/water.jl:1435-1558
 *   rs_value_and_grad   the forward_sim + reverse_sim pair as called by train(), water/water.jl:1662-1669
 *   rs_forces           calc_F!            water/water.jl:877-975
 *   rs_adjoint_terms    calc_F_and_grads!  water/water.jl:1212-1433 (plus the weighted sum of 2.c-2.e,
 *                                          water/water.jl:1517-1519)
 *   rs_noise            generate_noise!    water/water.jl:867-873 (Philox4x64-10 instead of Xoshiro)
 *   rs_rdf              rdf_differentiable + loss_rdf_part + its rrule, water/water.jl:631-736
 *   rs_msd              calculate_D + loss_fn_final + rrule, diffusion/diffusion.jl:861-911
 *   rs_molecule_pairs   Molly find_neighbors (CellListMap), called at water/water.jl:1025,1038
 *   rs_energy           potential_energy and its θ-gradient gradient(potential_energy, sys, ...),
 *                       water/water.jl:743,1183-1191
 *   rs_reweight         reweighting_sim (ensemble-reweighting gradient estimator),
 *                       water/water.jl:1107-1210
 *
 * Conventions
 *   units      nm, ps, u (g/mol), kJ/mol, e.
 *   layout     atoms are ordered O,H1,H2 per molecule (water/water.jl:550-556); coordinate and
 *              velocity arrays are 3*n_atoms doubles, xyz interleaved per atom (the memory layout of
 *              Julia's Vector{SVector{3,Float64}}), caller-owned; the library copies in and out.
 *   theta      6 doubles (sigma_O, eps_O, q_H, k_bond, r0_bond, k_angle); q_O = -2 q_H
 *              (water/water.jl:548-549); the angle rest value theta0 is a fixed config field.
 *   errors     every call returns 0 on success or a negative RS_E_* code; rs_last_error() gives the
 *              message. The library never aborts the host process.
 *   threading  a context is not thread-safe; calls are synchronous on the context's HIP stream.
 *              Different contexts may be used from different host threads.
 */
#ifndef REVSIM_H
#define REVSIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RS_ABI_VERSION 2

enum {
  RS_OK = 0,
  RS_E_CONFIG = -1,     /* invalid configuration or argument */
  RS_E_HIP = -2,        /* HIP runtime error */
  RS_E_OVERFLOW = -3,   /* neighbour-list capacity or fixed-point range exceeded */
  RS_E_NONFINITE = -4,  /* NaN/Inf in the state */
  RS_E_STALE = -5,      /* a molecule moved further than skin/2 between list builds */
  RS_E_STATE = -6,      /* call order violated (e.g. rs_reverse without a matching rs_forward) */
  RS_E_NODEVICE = -7    /* no usable gfx950 device */
};

enum { RS_PRECISION_FP64 = 0, RS_PRECISION_FP32 = 1 };

typedef struct rs_config {
  int32_t abi_version;     /* must be RS_ABI_VERSION */
  int32_t n_molecules;     /* atoms = 3 * n_molecules, order O,H1,H2 */
  double  box[3];          /* orthorhombic periodic box edges (nm) */
  double  cutoff;          /* pair cutoff r_c (nm); requires r_c <= box/2 */
  double  skin;            /* molecule-list skin (nm); list radius = r_c + 2*h_extent + skin */
  int32_t nl_every;        /* neighbour-list rebuild cadence (steps) */
  int32_t precision;       /* RS_PRECISION_FP64 or RS_PRECISION_FP32 (pair arithmetic) */
  double  dt;              /* time step (ps) */
  double  gamma;           /* Langevin friction (1/ps) */
  double  temperature;     /* K */
  double  eps_rf;          /* reaction-field dielectric */
  double  coulomb_k;       /* kJ mol^-1 nm e^-2 */
  double  kB;              /* kJ mol^-1 K^-1 */
  double  mass_O, mass_H;  /* u */
  double  theta0;          /* H-O-H rest angle (rad), fixed */
  /* O-O RDF loss: bins of width rdf_rmax/rdf_bins, Gaussian smoothing exp(-(x-r)^2/rdf_width)
     over bins within +-rdf_window of r, per-distance normalised (water/water.jl:631-691). */
  int32_t rdf_bins;
  double  rdf_rmax, rdf_width, rdf_window;
  double  w_rdf;           /* weight of sum_b (g_b - target_b)^2 */
  int32_t n_steps_loss;    /* RDF loss every n_steps_loss steps, averaged (0 = once at the end) */
  int32_t n_steps_trunc;   /* reverse truncation inside each loss block (0 = none) */
  int32_t n_steps_log;     /* forward snapshot cadence used to reset reverse drift (0 = none) */
  /* O mean-squared-displacement loss over the last msd_window_steps steps, frames every msd_every
     steps (diffusion/diffusion.jl:861-911): L = w_msd * 1e18 * (D[m^2/s] - msd_target)^2. */
  int32_t msd_window_steps;
  int32_t msd_every;
  double  msd_target;      /* m^2/s */
  double  w_msd;
  double  clip_norm;       /* clip ||dF/dx^T lambda|| at this value (<= 0 or inf: no clip) */
  int32_t use_mlp;         /* 0 = classical force field only, 1 = + radial-basis MLP on O-O pairs */
  int32_t pair_chunks;     /* lanes cooperating on one molecule's list: 8 or 16 (0 = default: 16) */
  /* Enthalpy-of-vaporization block loss (enthalpy_vaporization / loss_enth_vap,
     water/water.jl:738-754), evaluated with the RDF loss every n_steps_loss steps:
     dH = dhvap_u_gas - U/n_molecules + kB*temperature, L = w_dhvap * (dH - dhvap_target)^2.
     U includes the reaction-field shift (zero at r_c). Not available with use_mlp. */
  double  w_dhvap;         /* 0 = off */
  double  dhvap_target;    /* kJ/mol (reference: 44.12 + 2.8) */
  double  dhvap_u_gas;     /* mean gas-phase U per molecule, kJ/mol (reference: 3.69) */
  /* Monte Carlo barostat in the forward pass (MonteCarloBarostatNoUnits / apply_coupling!,
     water/water.jl:789-865; the reference equilibrates with it): every barostat_every steps one
     isotropic volume move dV = s (2u - 1), molecule centres of mass scaled, accepted with
     probability min(1, exp(-(dU + P dV - M kT ln(V'/V))/kT)); s starts at 0.01 V and is
     rescaled by 1.1 to keep acceptance in [0.25, 0.75] (capped at 0.3 V). Forward-only: a
     trajectory that moved the box has no reverse pass. Uniforms from Philox keyed by the seed. */
  int32_t barostat_every;  /* 0 = constant volume */
  double  barostat_pressure; /* bar */
  int32_t reserved[8];
} rs_config;

typedef struct rs_ctx rs_ctx;

/* Fill a config with the water defaults (the 216-water, 1.862 nm, 0.8 nm cutoff case). */
void rs_default_config(rs_config* cfg);

int rs_create(const rs_config* cfg, int32_t device, rs_ctx** out);
int rs_destroy(rs_ctx* ctx);
const char* rs_last_error(const rs_ctx* ctx);   /* ctx may be NULL: last global error */
int rs_version(void);
int32_t rs_config_size(void);                   /* sizeof(rs_config), for FFI layout checks */

int rs_set_theta(rs_ctx* ctx, const double* theta6);
int rs_set_target_rdf(rs_ctx* ctx, const double* g, int32_t n_bins);
/* Replace the Philox noise by caller noise xi[(step-1)*3N + dof] (already scaled by sqrt(kT/m)),
   for steps 1..n_steps; NULL restores the Philox stream. Used to replay another noise source. */
int rs_set_external_noise(rs_ctx* ctx, const double* xi, int32_t n_steps);

/* One forward trajectory from (x0, v0). Stores the snapshots and MSD frames that rs_reverse
   needs. x_out/v_out (nullable) receive the final state, loss_out (nullable) the loss. */
int rs_forward(rs_ctx* ctx, const double* x0, const double* v0, uint64_t seed, int32_t n_steps,
               double* x_out, double* v_out, double* loss_out);
/* The reverse-only arguments of reverse_sim (water/water.jl:1435-1436): n_steps_trunc (0: none)
   and clip_norm_threshold (<= 0 or inf: none), replacing the config's n_steps_trunc and clip_norm.
   The records of the last rs_forward stay valid, so this may come between rs_forward and
   rs_reverse. */
int rs_set_reverse_options(rs_ctx* ctx, int32_t n_steps_trunc, double clip_norm);
/* Reverse-time adjoint from the final state of the last rs_forward (same seed and n_steps).
   Returns the loss, dL/dtheta (6), and (nullable) the reconstructed initial state. */
int rs_reverse(rs_ctx* ctx, const double* xN, const double* vN, uint64_t seed, int32_t n_steps,
               double* loss_out, double* dL_dtheta6, double* x0_out, double* v0_out);
/* rs_forward followed by rs_reverse. dL_dmlp (nullable) receives the MLP weight gradient. */
int rs_value_and_grad(rs_ctx* ctx, const double* x0, const double* v0, uint64_t seed,
                      int32_t n_steps, double* loss_out, double* dL_dtheta6, float* dL_dmlp);
/* The loss terms of the last rs_forward / rs_reverse / rs_value_and_grad: terms4 =
   (l_rdf, l_ev, l_msd, D), the extra return values of the reference's reverse_sim
   (sys, l, dl_dp, l_rdf, l_ev), water/water.jl:1549-1557, and of the diffusion variant
   (sys, loss, dl_dp, D, l_diff, l_rdf, l_ev), diffusion/diffusion.jl:1605-1616: the RDF and
   enthalpy-of-vaporization terms averaged over loss blocks (weights applied), the MSD term, and the
   diffusion coefficient D in m^2/s. l_rdf + l_ev + l_msd is the loss. After rs_reverse the first
   two are the reverse pass's own block losses. */
int rs_loss_terms(const rs_ctx* ctx, double* terms4);
/* Counters (no counterpart in the reference): stats[RS_STAT_RERUNS_CALL] list-driven reruns of the
   last call that ran device work (a neighbour-list capacity overflow, a stale list, an O-H bond
   beyond the list's extent or a crowded cell made it widen the list and run the call again),
   stats[RS_STAT_BUILDS_CALL] its list builds (all attempts), stats[RS_STAT_RERUNS_TOTAL] reruns
   since the context was created. n: entries wanted (<= RS_STAT_COUNT). */
enum { RS_STAT_RERUNS_CALL = 0, RS_STAT_BUILDS_CALL, RS_STAT_RERUNS_TOTAL, RS_STAT_COUNT };
int rs_stats(const rs_ctx* ctx, int64_t* stats, int32_t n);

/* --- radial-basis MLP correction on O-O pairs (rs_config.use_mlp = 1) ---------------------
   u(r) = w3.SiLU(W2 SiLU(W1 φ(r) + b1) + b2) + b3 per O-O pair within the cutoff, 16 Gaussian
   features (centres 0.2-1.0 nm, width 0.05 nm) with a cosine cutoff. No reference counterpart:
   the reference's learned-potential variants evaluate a network through PythonCall
   (water/water_maceoff.jl:147-163 HVP/VJP by double backward); this is the MI355X version for a
   pair network. Weights are packed W1[64][16], b1[64], W2[64][64], b2[64], w3[64], b3 (5313
   floats) and rounded to bfloat16 on entry. */
#define RS_MLP_NWEIGHTS 5313
int rs_set_mlp(rs_ctx* ctx, const float* weights, int64_t n);
/* dL/dW of the last rs_reverse / rs_value_and_grad / rs_adjoint_terms call (n = 5313). */
int rs_mlp_grad(rs_ctx* ctx, double* dL_dW, int64_t n);
/* The tabulated pair force f(r_k) = -u'(r_k) and f'(r_k) = -u''(r_k) on the library's n_nodes
   nodes r_k = r_lo + k (r_c - r_lo)/(n_nodes - 1); returns n_nodes and r_lo. */
int rs_mlp_table(rs_ctx* ctx, double* f, double* df, int32_t capacity, int32_t* n_nodes, double* r_lo);

/* --- single-evaluation entry points (parity and unit tests) ------------------------------ */
/* Forces F(x) (3N doubles). */
int rs_forces(rs_ctx* ctx, const double* x, double* F);
/* One adjoint pass at x: F(x), h = (dF/dx)^T (c/m) and vjp6[k] = sum_dof w_dof dF_dof/dtheta_k,
   with c = accum_C and w = accum_A*dt^2/(2m) as in water/water.jl:1505-1519. */
int rs_adjoint_terms(rs_ctx* ctx, const double* x, const double* c, const double* w, double* F,
                     double* h, double* vjp6);
/* Langevin noise of one step (3N doubles, scaled by sqrt(kB T / m)). */
int rs_noise(rs_ctx* ctx, uint64_t seed, int32_t step, double* xi);
/* O-O RDF at x: g (rdf_bins), raw fixed-point bin sums (nullable), loss vs the target and dL/dx. */
int rs_rdf(rs_ctx* ctx, const double* x, double* g, int64_t* hist_fixed, double* loss,
           double* dL_dx);
/* MSD loss over n_frames frames of O positions (frames[f*3M + 3i + d]): D (nm^2/ps), loss and
   dL/dframes (same layout). */
int rs_msd(rs_ctx* ctx, const double* frames, int32_t n_frames, double* D, double* loss,
           double* dL_dframes);
/* rs_msd for callers that cannot allocate device memory (code under HIP graph capture, tight
   serving loops, harnesses that forbid hipMalloc); no counterpart in the reference. frames and
   dL_dframes are device arrays (layout as rs_msd); D and loss are host scalars. The call allocates
   nothing: its kernels' scratch is a static device array of 33.6 MB that each GPU holds from the
   library's load until the process ends, which limits a call to n_frames/2 * n_molecules <=
   1,048,576 (RS_E_CONFIG beyond). All calls on one GPU share that array: calls that overlap in
   time on one GPU (other host threads, other contexts) race. Synchronous on the context's stream;
   results equal rs_msd's bit for bit. */
int rs_msd_noalloc(rs_ctx* ctx, const double* frames, int32_t n_frames, double* D, double* loss,
                   double* dL_dframes);
/* Potential energy U(x) (kJ/mol: reaction field with its r_c shift, Lennard-Jones, bonds, angles),
   its gradient dU/dtheta (6, nullable) and the forces -dU/dx (3N, nullable), in FP64 in both
   precisions. Classical force field only (RS_E_CONFIG with use_mlp). */
int rs_energy(rs_ctx* ctx, const double* x, double* U, double* dU_dtheta6, double* F);
/* Ensemble-reweighting estimator over one forward trajectory (reweighting_sim,
   water/water.jl:1107-1210): at each of the B = n_steps/n_steps_loss block ends the block loss
   L_b (RDF and enthalpy-of-vaporization terms), its direct gradient dL_b/dtheta and dU/dtheta
   are taken, and
     dL/dtheta = <dL_b/dtheta> - (<L_b dU_b/dtheta> - <L_b><dU_b/dtheta>) / (kB T),
   <.> the mean over blocks; loss_out = <L_b>. x_out/v_out (nullable) receive the final state.
   The MSD loss is not part of it (RS_E_CONFIG when configured); classical force field only. */
int rs_reweight(rs_ctx* ctx, const double* x0, const double* v0, uint64_t seed, int32_t n_steps,
                double* loss_out, double* dL_dtheta6, double* x_out, double* v_out);
/* Current box edges (nm): the configured box, or where the barostat left it. */
int rs_get_box(const rs_ctx* ctx, double* box3);
/* Replace the box for the following calls (e.g. the equilibrated box of a barostat run); the
   state of the last forward trajectory is discarded. Same constraints as rs_create. */
int rs_set_box(rs_ctx* ctx, const double* box3);
/* Molecule pairs (i<j) whose O-O distance is within the list radius; pairs[2*k], pairs[2*k+1]. */
int rs_molecule_pairs(rs_ctx* ctx, const double* x, int32_t* pairs, int64_t capacity,
                      int64_t* n_pairs);
/* List radius used by rs_molecule_pairs and the force kernels (nm). */
double rs_list_radius(const rs_ctx* ctx);
/* Per-phase device time (ms) accumulated since the last call: see RS_PHASE_*. The forward pass
   runs the pair forces and the integrator step as one kernel per step, timed as FWD_PAIR
   (FWD_STEP stays zero). */
enum { RS_PHASE_FWD_PAIR = 0, RS_PHASE_FWD_STEP, RS_PHASE_REV_PAIR, RS_PHASE_REV_STEP,
       RS_PHASE_LIST, RS_PHASE_LOSS, RS_PHASE_MLP, RS_PHASE_COUNT };
int rs_timings(rs_ctx* ctx, double* ms_by_phase, int32_t n);
/* Number of timed launches per phase accumulated since the last rs_timings call (read it first). */
int rs_timing_counts(rs_ctx* ctx, int64_t* counts, int32_t n);
/* Enable/disable per-phase event timing (off by default; adds event records per launch). */
int rs_enable_timing(rs_ctx* ctx, int32_t on);
/* Measurement aid, no counterpart in the reference: `reps` launches of a plain 16-byte vector copy
   kernel (k_copy16) over `bytes` bytes between two scratch buffers allocated for the call, so
   hardware traffic counters can be scaled against a kernel whose HBM traffic is known exactly
   (bytes read + bytes written per launch). */
int rs_copy_calibration(rs_ctx* ctx, int64_t bytes, int32_t reps);

#ifdef __cplusplus
}
#endif
#endif /* REVSIM_H */
