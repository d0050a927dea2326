#!/usr/bin/env python3
"""Benchmark: forward+reverse gradient steps/s of the reversible-simulation loop on one MI355X.

Workload (default): 32,768 flexible waters (98,304 atoms) in a 9.95 nm box, r_c = 1.0 nm, list
rebuilt every 10 steps, 1,000 forward + 1,000 reverse steps per trajectory, loss = O-O RDF (200
bins, 0-0.8 nm) + O MSD slope over the last 500 fs vs 2.3e-9 m^2/s. FP32 pair arithmetic, FP64
state and dL/dθ. One bench "step" is one full trajectory (rs_value_and_grad); `value` counts
gradient steps (one forward + one reverse MD step) per second.

Also reports: the dominant kernel's roofline (HIP events on the library's stream), the FP32-path
dL/dθ relative error against the FP64 path, and the FP64 CPU oracle timed on this host on a
bounded sample of the same system (CPU baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rev-sim_amd"))

METRIC = "fwd+rev gradient steps/s at 98k atoms (plain vs +MLP); dL/dθ rel err vs FP64"
PEAK_FP32 = 157.3  # TFLOP/s, MI355X vector FP32 (packed); non-packed issue rate is half
PEAK_FP32_NONPACKED = 78.6
PEAK_HBM = 8000.0  # GB/s
# algorithmic FLOP per pair within r_c (add/sub/mul 1, FMA 2, sqrt/rcp/exp 1): base, extra for O-O LJ
FLOP_FWD = (36, 15)
FLOP_REV = (72, 33)
# HBM bytes per launch of the adjoint pair kernel at this workload: FETCH_SIZE x 2.000 + WRITE_SIZE
# from rocprofv3 counter runs, scaled by the library's copy kernel (profiles/kernels_water32768.md)
TRAFFIC_REV_PAIR = 107.8e6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def count_pairs(x, box, rc):
    """Atom pairs within r_c in different molecules, and O-O pairs, by a periodic k-d tree."""
    from scipy.spatial import cKDTree

    pos = np.mod(x.reshape(-1, 3), box)
    tree = cKDTree(pos, boxsize=box)
    n_all = int(tree.count_neighbors(tree, rc)) - pos.shape[0]  # ordered pairs incl. self
    n_all //= 2
    n_mol = pos.shape[0] // 3
    n_pairs = n_all - 3 * n_mol  # the 3 intramolecular pairs of each water are within r_c
    O = pos[0::3]
    to = cKDTree(O, boxsize=box)
    n_oo = (int(to.count_neighbors(to, rc)) - O.shape[0]) // 2
    return n_pairs, n_oo


def cpu_baseline(wl, x, v, theta, target, steps):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import revsim

    kw = dict(precision=revsim.RS_PRECISION_FP64)
    if wl.get("msd_window_steps"):
        kw.update(msd_window_steps=steps, msd_every=steps // 2, w_msd=wl.get("w_msd", 1.0))
    cfg = revsim.make_config(wl["n_molecules"], wl["box"], wl["cutoff"], **kw)
    orc = O.Oracle(cfg)
    orc.set_theta(theta)
    orc.set_target_rdf(target)
    t0 = time.perf_counter()
    orc.value_and_grad(x, v, 17, steps)
    dt = time.perf_counter() - t0
    return {
        "value": steps / dt,
        "unit": "gradient steps/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": f"FP64 C restatement of the reference CPU loop (OpenMP), same {wl['n_molecules']}-water "
                  f"system, {steps} forward + {steps} reverse steps (RDF + MSD loss over those steps), "
                  f"{dt:.1f} s wall",
    }


def fit_msd_window(extra, n_steps):
    """Shorter trajectories than the MSD window: the MSD over their second half, with at least two
    frame intervals (window a multiple of 2 * msd_every, as the library requires)."""
    ex = dict(extra)
    if ex.get("msd_window_steps", 0) > n_steps:
        every = ex["msd_every"]
        if n_steps // 2 < 2 * every:
            every = max(1, n_steps // 4)
        ex["msd_every"] = every
        ex["msd_window_steps"] = max(2 * every, (n_steps // 2) // (2 * every) * (2 * every))
    return ex


def timed_trajectories(sim, x, v, seed, n_steps, warmup, steps, tag):
    """W untimed then K timed value_and_grad trajectories (no instrumentation inside the timed
    region); then one more trajectory with per-phase HIP events recorded on the library's stream,
    whose phase times are returned (milliseconds for that one trajectory)."""
    for w in range(warmup):
        t = time.perf_counter()
        sim.value_and_grad(x, v, seed, n_steps)
        log(f"[bench] {tag} warmup {w}: {time.perf_counter() - t:.3f} s")
    t0 = time.perf_counter()
    for k in range(steps):
        loss, grad = sim.value_and_grad(x, v, seed, n_steps)[:2]
    elapsed = time.perf_counter() - t0
    sim.enable_timing(True)
    sim.timings()
    sim.value_and_grad(x, v, seed, n_steps)
    tim = sim.timings()
    sim.enable_timing(False)
    log(f"[bench] {tag}: {steps} trajectories in {elapsed:.3f} s -> {steps * n_steps / elapsed:.1f} "
        f"gradient steps/s; phases of one instrumented trajectory {json.dumps(tim)}")
    return loss, grad, elapsed, tim


def grad_error_by_horizon(wl, extra, x, v, theta, target, seed, horizons, mlp=None):
    """dL/dθ of the FP32 path against the FP64 path from the same start (the FP64 path is
    parity-tested against the FP64 CPU oracle)."""
    import revsim

    errs = {}
    for horizon in horizons:
        ex = fit_msd_window(extra, horizon)
        g, secs = {}, {}
        for p_ in (revsim.RS_PRECISION_FP32, revsim.RS_PRECISION_FP64):
            c_ = revsim.make_config(wl["n_molecules"], wl["box"], wl["cutoff"], precision=p_,
                                    use_mlp=int(mlp is not None), **ex)
            with revsim.Simulator(c_, device=0) as s_:
                s_.set_theta(theta)
                s_.set_target_rdf(target)
                if mlp is not None:
                    s_.set_mlp(mlp)
                t0 = time.perf_counter()  # first call of a context: includes its list set-up
                g[p_] = s_.value_and_grad(x, v, seed, horizon)[1]
                secs[p_] = time.perf_counter() - t0
        g32, g64 = g[revsim.RS_PRECISION_FP32], g[revsim.RS_PRECISION_FP64]
        errs[str(horizon)] = {
            "normwise": float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
            "per_param": (np.abs(g32 - g64) / np.maximum(np.abs(g64), 1e-300)).tolist(),
            "grad_norm_fp64": float(np.linalg.norm(g64)),
            # gradient steps/s of each path on this (single, untimed-warm-up) trajectory
            "steps_per_s": {"fp32": round(horizon / secs[revsim.RS_PRECISION_FP32], 1),
                            "fp64": round(horizon / secs[revsim.RS_PRECISION_FP64], 1)},
        }
        log(f"[bench] dL/dθ FP32 vs FP64 at {horizon} steps{' (+MLP)' if mlp is not None else ''}: "
            f"{errs[str(horizon)]['normwise']:.3e}")
    return errs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed trajectories")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="water32768")
    ap.add_argument("--n-steps", type=int, default=None, help="MD steps per trajectory")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--equil", type=int, default=2000, help="untimed equilibration steps")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-check", action="store_true")
    ap.add_argument("--pair-chunks", type=int, default=0)
    ap.add_argument("--no-mlp", action="store_true", help="skip the +MLP leg")
    ap.add_argument("--no-long", action="store_true",
                    help="skip the 5,000-step leg in loss blocks (the learned-correction-at-scale config)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run 5 device copies of this many MiB (a kernel of known HBM traffic, "
                         "to scale FETCH_SIZE / WRITE_SIZE counters by)")
    a = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print(json.dumps({"error": "bench.py measures one GPU; runs on more than one GPU are not supported",
                          "n_gpus": a.gpus}))
        sys.exit(2)

    import revsim

    wl = dict(revsim.WORKLOADS[a.workload])
    n_steps = a.n_steps or wl.pop("n_steps")
    wl.pop("n_steps", None)
    prec = revsim.RS_PRECISION_FP32 if a.precision == "fp32" else revsim.RS_PRECISION_FP64
    extra = {k: wl[k] for k in ("msd_window_steps", "msd_every", "w_msd", "use_mlp") if k in wl}
    extra = fit_msd_window(extra, n_steps)
    cfg = revsim.make_config(wl["n_molecules"], wl["box"], wl["cutoff"], precision=prec,
                             pair_chunks=a.pair_chunks, **extra)
    wb = revsim.water_box(wl["n_molecules"], wl["box"])
    theta = revsim.DEFAULT_THETA
    log(f"[bench] workload {a.workload}: {wb.n_atoms} atoms, {n_steps}+{n_steps} steps, {a.precision}")
    sim = revsim.Simulator(cfg, device=0)
    sim.set_theta(theta)

    # untimed equilibration from the lattice, then the synthetic target RDF
    x, v = wb.coords, wb.velocities
    if a.equil > 0:
        cfg_eq = revsim.make_config(wl["n_molecules"], wl["box"], wl["cutoff"], precision=prec, w_rdf=0.0)
        with revsim.Simulator(cfg_eq, device=0) as eq:
            eq.set_theta(theta)
            x, v, _ = eq.forward(x, v, revsim.system.SEED_NOISE + 100, a.equil)
    g0, _, _, _ = sim.rdf(x)
    target = revsim.perturbed_target(g0)
    sim.set_target_rdf(target)
    seed = revsim.system.SEED_NOISE

    loss, grad, elapsed, tim = timed_trajectories(sim, x, v, seed, n_steps, a.warmup, a.steps, "plain")
    value = a.steps * n_steps / elapsed

    # roofline of the dominant kernel (the adjoint pair kernel)
    xN, _, _ = sim.forward(x, v, seed, n_steps)
    p0 = count_pairs(x, wl["box"], wl["cutoff"])
    p1 = count_pairs(xN, wl["box"], wl["cutoff"])
    n_pairs, n_oo = (p0[0] + p1[0]) / 2, (p0[1] + p1[1]) / 2
    dom = max(("fwd_pair", "rev_pair", "fwd_step", "rev_step", "list", "loss"), key=lambda k: tim[k]["ms"])
    rp = tim["rev_pair"]
    avg_ms = rp["ms"] / max(rp["launches"], 1)
    flop = FLOP_REV[0] * n_pairs + FLOP_REV[1] * n_oo
    achieved = flop / (avg_ms * 1e-3) / 1e12
    fp = tim["fwd_pair"]
    fwd_avg = fp["ms"] / max(fp["launches"], 1)
    traffic = TRAFFIC_REV_PAIR if a.workload == "water32768" and prec == revsim.RS_PRECISION_FP32 else None
    roof = {
        "kernel": "k_pair<P32,CH,adjoint> (reverse pair + HVP + θ-VJP)",
        "bound": "valu",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32, 4),
        "frac_of_nonpacked_peak": round(achieved / PEAK_FP32_NONPACKED, 4),
        "traffic": traffic,
        "traffic_unit": "bytes per launch",
        # the same kernel's HBM rate (measured bytes per launch / event-timed launch) against 8 TB/s
        "hbm_gb_s": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
        "hbm_frac": round(traffic / (avg_ms * 1e-3) / 1e9 / PEAK_HBM, 4) if traffic else None,
        "avg_launch_ms": round(avg_ms, 5),
        "flop_per_launch": flop,
        "pairs_within_rc": n_pairs,
        "oo_pairs_within_rc": n_oo,
        "dominant_phase": dom,
        "list_radius_nm": round(sim.list_radius(), 4),
        # the forward kernel: pair forces, then the Langevin step and the bonded terms (one launch
        # per step); its rate counts the pair-term flops only
        "fwd_pair_step_avg_launch_ms": round(fwd_avg, 5),
        "fwd_pair_tflops": round((FLOP_FWD[0] * n_pairs + FLOP_FWD[1] * n_oo) / (fwd_avg * 1e-3) / 1e12, 3),
    }

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "gradient steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * elapsed / a.steps, 3),
        "ms_per_gradient_step": round(1e3 * elapsed / (a.steps * n_steps), 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if prec == revsim.RS_PRECISION_FP32 else "fp64",
        "state_dtype": "fp64",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: {wl['n_molecules']} waters / {wb.n_atoms} atoms, "
                               f"{wl['box']} nm box, r_c {wl['cutoff']} nm, {n_steps}+{n_steps} steps, "
                               f"RDF{'+MSD' if extra.get('w_msd') else ''} loss",
                   "atoms": wb.n_atoms, "n_steps": n_steps, "parallelism": "single GPU"},
        "loss": loss,
        "dL_dtheta": grad.tolist(),
        "phases_ms_one_trajectory": {k: round(vv["ms"], 3) for k, vv in tim.items()},
        # atom pairs within r_c evaluated per second (each gradient step evaluates them forward and
        # in the adjoint)
        "pair_evals_per_s": round(2 * n_pairs * value, -6),
        "roofline": roof,
    }

    if not a.no_mlp:
        # the same workload with the radial-basis MLP correction on O-O pairs (Xavier weights)
        from revsim.system import mlp_init

        cfg_m = revsim.make_config(wl["n_molecules"], wl["box"], wl["cutoff"], precision=prec,
                                   pair_chunks=a.pair_chunks, **dict(extra, use_mlp=1))
        with revsim.Simulator(cfg_m, device=0) as sm:
            sm.set_theta(theta)
            sm.set_target_rdf(target)
            sm.set_mlp(mlp_init())
            loss_m, grad_m, el_m, tim_m = timed_trajectories(sm, x, v, seed, n_steps, a.warmup, a.steps, "+MLP")
        vm = a.steps * n_steps / el_m
        out["mlp"] = {
            "value": round(vm, 2),
            "unit": "gradient steps/s",
            "ms_per_step": round(1e3 * el_m / a.steps, 3),
            "cost_ratio_vs_plain": round(el_m / elapsed, 4),
            "loss": loss_m,
            "dL_dtheta": grad_m.tolist(),
            "dL_dtheta_rel_diff_vs_plain": float(np.linalg.norm(grad_m - grad) / np.linalg.norm(grad)),
            "phases_ms_one_trajectory": {k: round(vv["ms"], 3) for k, vv in tim_m.items()},
            "network": "16 RBF -> 64 -> 64 -> 1 SiLU, bf16 weights (Xavier, seed 15); table of f, f' on "
                       "512 nodes built by bf16 MFMA per trajectory, cubic Hermite in the pair kernels",
        }

    if not a.no_mlp and not a.no_long and a.workload == "water32768":
        # The learned correction at scale: 5,000 forward steps, in 1,000-step loss blocks whose
        # reverse passes are truncated at 200 steps (the reference's training regime). Full
        # backprop over 5,000 steps is not defined in FP32: the reverse trajectory leaves the
        # forward one after ~500 steps and its friction-inverting integrator then runs away.
        from revsim.system import mlp_init

        lw = revsim.WORKLOADS["water32768_mlp"]
        T, blk, trunc = lw["n_steps"], 1000, 200
        legs = {}
        for name, use in (("plain", 0), ("mlp", 1)):
            ex = dict(extra, use_mlp=use, n_steps_loss=blk, n_steps_trunc=trunc)
            with revsim.Simulator(revsim.make_config(wl["n_molecules"], wl["box"], wl["cutoff"], precision=prec,
                                                     pair_chunks=a.pair_chunks, **ex), device=0) as sl:
                sl.set_theta(theta)
                sl.set_target_rdf(target)
                if use:
                    sl.set_mlp(mlp_init())
                sl.value_and_grad(x, v, seed, blk)  # warm-up: table, list capacities
                t0 = time.perf_counter()  # calls are synchronous on the library's stream
                L_l, g_l = sl.value_and_grad(x, v, seed, T)
                legs[name] = (time.perf_counter() - t0, L_l, g_l)
        out["long"] = {
            "config": f"{wl['n_molecules']} waters, {T} forward steps in {T // blk} loss blocks of {blk}, "
                      f"reverse truncated at {trunc} steps per block, RDF+MSD loss",
            "forward_steps_per_s": {k: round(T / vv[0], 1) for k, vv in legs.items()},
            "reverse_steps": (T // blk) * trunc,
            "mlp_cost_ratio_vs_plain": round(legs["mlp"][0] / legs["plain"][0], 4),
            "loss": {k: vv[1] for k, vv in legs.items()},
            "dL_dtheta": {k: vv[2].tolist() for k, vv in legs.items()},
            "dL_dtheta_rel_diff_mlp_vs_plain": float(np.linalg.norm(legs["mlp"][2] - legs["plain"][2]) /
                                                     np.linalg.norm(legs["plain"][2])),
        }

    if not a.no_fp64_check and prec == revsim.RS_PRECISION_FP32:
        horizons = sorted({min(100, n_steps), n_steps})
        out["grad_rel_err_vs_fp64"] = {
            "by_horizon_steps": grad_error_by_horizon(wl, extra, x, v, theta, target, seed, horizons),
            "reference": "same start, FP64 pair arithmetic (parity-tested against the FP64 CPU oracle)",
            "note": "beyond ~200 steps the adjoint grows like exp(lambda t) (chaotic MD) and FP32/FP64 "
                    "gradients decorrelate; the reference trains with n_steps_trunc = 200",
        }
        if not a.no_mlp:
            from revsim.system import mlp_init

            out["mlp"]["grad_rel_err_vs_fp64"] = grad_error_by_horizon(
                wl, extra, x, v, theta, target, seed, [min(100, n_steps)], mlp=mlp_init())

    if a.calibrate_copy_mb:
        nb = a.calibrate_copy_mb << 20
        sim.copy_calibration(nb, 5)
        out["copy_calibration"] = {"bytes_read": nb, "bytes_written": nb, "launches": 5,
                                   "kernel": "k_copy16 (16-byte vector copy, librevsim)"}

    if not a.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(wl, x, v, theta, target, a.cpu_steps)
        out["speedup_vs_cpu_baseline"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
