#!/usr/bin/env python3
"""Benchmark: forward+reverse gradient steps/s of the reversible-simulation loop on one MI355X.

Workload (default): 32,768 flexible waters (98,304 atoms) in a 9.95 nm box, r_c = 1.0 nm, list
rebuilt every 10 steps, 1,000 forward + 1,000 reverse steps per trajectory, loss = O-O RDF (200
bins, 0-0.8 nm) + O MSD slope over the last 500 fs vs 2.3e-9 m^2/s. FP32 pair arithmetic, FP64
state and dL/dθ. One bench "step" is one full trajectory (rs_value_and_grad); `value` counts
gradient steps (one forward + one reverse MD step) per second, inputs resident on the device.

Also reported, on the same start state:
  fp64        the FP64-pair path timed by the same protocol (it reproduces the FP64 oracle's
              trajectory bit for bit);
  mlp         the same workload with the radial-basis MLP correction;
  long        5,000 forward steps in 1,000-step loss blocks, reverse truncated at 200 per block;
  accuracy    dL/dθ of both paths against the FP64 CPU oracle on the first 100 steps, and the FP32
              path against the FP64 path over the full 1,000 steps;
  roofline    of the adjoint pair kernel (HIP events on the library's stream; HBM traffic from the
              counter run in profiles/, if it was taken on these kernel sources);
  cpu_baseline  the FP64 oracle (the C restatement of the reference's threaded CPU loop) timed on
              this host: the 98k-atom workload over 100 + 100 steps, and the 9,000-atom
              configuration at its full 1,000 + 1,000 steps.
"""
from __future__ import annotations

import argparse
import glob
import hashlib
import json
import os
import platform
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
PEAK_FP64 = 78.6  # TFLOP/s, MI355X vector FP64
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_water32768.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sources_digest() -> str:
    """SHA-256 over the HIP sources the library is built from (the counter run records it)."""
    h = hashlib.sha256()
    for p in sorted(glob.glob(os.path.join(ROOT, "rev-sim_amd", "csrc", "*"))):
        with open(p, "rb") as f:
            h.update(os.path.basename(p).encode() + b"\0" + f.read())
    return h.hexdigest()


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
    region; calls are synchronous on the library's stream); then one more trajectory with per-phase
    HIP events recorded on that stream, whose phase times are returned (milliseconds for that one
    trajectory). Also the list-driven reruns over the timed calls (rs_stats)."""
    for w in range(warmup):
        t = time.perf_counter()
        sim.value_and_grad(x, v, seed, n_steps)
        log(f"[bench] {tag} warmup {w}: {time.perf_counter() - t:.3f} s")
    reruns = 0
    t0 = time.perf_counter()
    for k in range(steps):
        loss, grad = sim.value_and_grad(x, v, seed, n_steps)[:2]
        reruns += sim.stats()["reruns_call"]
    elapsed = time.perf_counter() - t0
    sim.enable_timing(True)
    sim.timings()
    sim.value_and_grad(x, v, seed, n_steps)
    tim = sim.timings()
    sim.enable_timing(False)
    log(f"[bench] {tag}: {steps} trajectories in {elapsed:.3f} s -> {steps * n_steps / elapsed:.1f} "
        f"gradient steps/s, {reruns} list reruns; phases of one instrumented trajectory {json.dumps(tim)}")
    return loss, grad, elapsed, tim, reruns


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def per_param(a, b):
    return (np.abs(np.asarray(a) - np.asarray(b)) / np.maximum(np.abs(b), 1e-300)).tolist()


def make_sim(wl, extra, prec, theta, target, mlp=None, **kw):
    import revsim

    cfg = revsim.make_config(wl["n_molecules"], wl["box"], wl["cutoff"], precision=prec,
                             use_mlp=int(mlp is not None), **dict(extra, **kw))
    s = revsim.Simulator(cfg, device=0)
    s.set_theta(theta)
    s.set_target_rdf(target)
    if mlp is not None:
        s.set_mlp(mlp)
    return s


def gpu_gradient(wl, extra, prec, theta, target, x, v, seed, horizon, mlp=None):
    with make_sim(wl, fit_msd_window(extra, horizon), prec, theta, target, mlp) as s:
        t0 = time.perf_counter()
        L, g = s.value_and_grad(x, v, seed, horizon)[:2]
        return L, g, time.perf_counter() - t0


def oracle_gradient(wl, extra, theta, target, x, v, seed, horizon):
    """The FP64 oracle's loss and dL/dθ over `horizon` + `horizon` steps, and its wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import revsim

    cfg = revsim.make_config(wl["n_molecules"], wl["box"], wl["cutoff"], precision=revsim.RS_PRECISION_FP64,
                             **fit_msd_window(extra, horizon))
    orc = O.Oracle(cfg)
    orc.set_theta(theta)
    orc.set_target_rdf(target)
    t0 = time.perf_counter()
    L, g = orc.value_and_grad(x, v, seed, horizon)
    return L, g, time.perf_counter() - t0, O.num_threads()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def adj_kernel(fp64: bool, pair_chunks: int, n_molecules: int) -> str:
    """Name of the adjoint pair kernel a trajectory runs: k_pair<precision, lanes per molecule,
    adjoint, MLP compiled in, step epilogue, 16-bit list entries>. Lanes default to 16; lists have
    16-bit entries up to 65,536 molecules."""
    ch = pair_chunks or 16
    l16 = "true" if n_molecules <= 65536 else "false"
    return f"k_pair<{'P64' if fp64 else 'P32'}, {ch}, true, false, false, {l16}>"


def traffic_from_profile(digest, kernel):
    """Bytes per launch of an adjoint pair kernel from the committed counter run
    (rev-sim_amd/measure_traffic.py), used only if it was taken on these kernel sources."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None, {"file": None}
    k = rec.get("kernels", {}).get(kernel)
    match = rec.get("sources_sha256") == digest
    src = {"file": os.path.relpath(TRAFFIC_FILE, ROOT), "sources_match": match,
           "command": rec.get("command"), "scale": rec.get("scale")}
    if not k or not match:
        return None, src
    return k["fetch_bytes"] + k["write_bytes"], src


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed trajectories")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="water32768")
    ap.add_argument("--n-steps", type=int, default=None, help="MD steps per trajectory")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--equil", type=int, default=2000, help="untimed equilibration steps")
    ap.add_argument("--oracle-steps", type=int, default=100,
                    help="horizon of the FP64 CPU oracle comparison (also the CPU baseline sample)")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the oracle (comparison and timing)")
    ap.add_argument("--no-cpu-c1", action="store_true", help="skip the 9,000-atom full-length CPU timing")
    ap.add_argument("--no-fp64", action="store_true", help="skip the timed FP64-path leg")
    ap.add_argument("--no-fp64-check", action="store_true", help="skip the GPU FP32-vs-FP64 horizons")
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
    from revsim.system import mlp_init

    FP32, FP64 = revsim.RS_PRECISION_FP32, revsim.RS_PRECISION_FP64
    wl = dict(revsim.WORKLOADS[a.workload])
    n_steps = a.n_steps or wl.pop("n_steps")
    wl.pop("n_steps", None)
    prec = FP32 if a.precision == "fp32" else FP64
    extra = {k: wl[k] for k in ("msd_window_steps", "msd_every", "w_msd") if k in wl}
    extra_run = fit_msd_window(extra, n_steps)
    if a.pair_chunks:
        extra_run["pair_chunks"] = a.pair_chunks
    wb = revsim.water_box(wl["n_molecules"], wl["box"])
    theta = revsim.DEFAULT_THETA
    log(f"[bench] workload {a.workload}: {wb.n_atoms} atoms, {n_steps}+{n_steps} steps, {a.precision}")

    # untimed equilibration from the lattice, then the synthetic target RDF
    x, v = wb.coords, wb.velocities
    if a.equil > 0:
        cfg_eq = revsim.make_config(wl["n_molecules"], wl["box"], wl["cutoff"], precision=FP32, w_rdf=0.0)
        with revsim.Simulator(cfg_eq, device=0) as eq:
            eq.set_theta(theta)
            x, v, _ = eq.forward(x, v, revsim.system.SEED_NOISE + 100, a.equil)
    with revsim.Simulator(revsim.make_config(wl["n_molecules"], wl["box"], wl["cutoff"], precision=FP64),
                          device=0) as s0:
        s0.set_theta(theta)
        target = revsim.perturbed_target(s0.rdf(x)[0])
    seed = revsim.system.SEED_NOISE

    sim = make_sim(wl, extra_run, prec, theta, target)
    loss, grad, elapsed, tim, reruns = timed_trajectories(sim, x, v, seed, n_steps, a.warmup, a.steps,
                                                          "plain " + a.precision)
    value = a.steps * n_steps / elapsed
    terms = sim.loss_terms()

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
    digest = sources_digest()
    adj_name = adj_kernel(prec != FP32, a.pair_chunks, wl["n_molecules"])
    traffic, traffic_src = (traffic_from_profile(digest, adj_name) if a.workload == "water32768"
                            else (None, {"file": None}))
    roof = {
        "kernel": adj_name + " (reverse pair forces + HVP + θ-VJP)",
        "bound": "valu",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP32,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP32, 4),
        "frac_of_nonpacked_peak": round(achieved / PEAK_FP32_NONPACKED, 4),
        "traffic": traffic,
        "traffic_unit": "bytes per launch",
        "traffic_source": traffic_src,
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
    sim.close()

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
        "dtype": "fp32" if prec == FP32 else "fp64",
        "state_dtype": "fp64",
        "data": "synthetic",
        "config": {"workload": f"{a.workload}: {wl['n_molecules']} waters / {wb.n_atoms} atoms, "
                               f"{wl['box']} nm box, r_c {wl['cutoff']} nm, {n_steps}+{n_steps} steps, "
                               f"RDF{'+MSD' if extra.get('w_msd') else ''} loss",
                   "atoms": wb.n_atoms, "n_steps": n_steps, "parallelism": "single GPU"},
        "loss": loss,
        "loss_terms": terms,
        "dL_dtheta": grad.tolist(),
        "list_reruns_timed": reruns,
        "phases_ms_one_trajectory": {k: round(vv["ms"], 3) for k, vv in tim.items()},
        # atom pairs within r_c evaluated per second (each gradient step evaluates them forward and
        # in the adjoint)
        "pair_evals_per_s": round(2 * n_pairs * value, -6),
        "roofline": roof,
    }

    if not a.no_fp64 and prec == FP32:
        # the FP64-pair path by the same protocol: its trajectory is the FP64 oracle's bit for bit
        with make_sim(wl, extra_run, FP64, theta, target) as s64:
            L64, g64, el64, tim64, rr64 = timed_trajectories(s64, x, v, seed, n_steps, a.warmup, a.steps, "plain fp64")
        out["fp64"] = {
            "value": round(a.steps * n_steps / el64, 2),
            "unit": "gradient steps/s",
            "ms_per_step": round(1e3 * el64 / a.steps, 3),
            "steps": a.steps, "warmup": a.warmup,
            "loss": L64,
            "dL_dtheta": g64.tolist(),
            "list_reruns_timed": rr64,
            "phases_ms_one_trajectory": {k: round(vv["ms"], 3) for k, vv in tim64.items()},
            "note": "FP64 pair arithmetic, pair forces summed as exact 64-bit fixed point: the same "
                    "trajectory as the FP64 CPU oracle, bit for bit",
        }
        # roofline of its adjoint pair kernel against the FP64 vector peak
        rp64 = tim64["rev_pair"]
        avg64 = rp64["ms"] / max(rp64["launches"], 1)
        ach64 = flop / (avg64 * 1e-3) / 1e12
        adj64 = adj_kernel(True, a.pair_chunks, wl["n_molecules"])
        tr64, tr64_src = (traffic_from_profile(digest, adj64) if a.workload == "water32768"
                          else (None, {"file": None}))
        out["fp64"]["roofline"] = {
            "kernel": adj64, "bound": "valu", "achieved": round(ach64, 3), "peak": PEAK_FP64,
            "unit": "TFLOP/s", "frac": round(ach64 / PEAK_FP64, 4), "traffic": tr64,
            "traffic_unit": "bytes per launch", "traffic_source": tr64_src,
            "avg_launch_ms": round(avg64, 5),
            "fwd_pair_step_avg_launch_ms": round(tim64["fwd_pair"]["ms"] / max(tim64["fwd_pair"]["launches"], 1), 5),
        }

    if not a.no_mlp:
        # the same workload with the radial-basis MLP correction on O-O pairs (Xavier weights)
        with make_sim(wl, extra_run, prec, theta, target, mlp=mlp_init()) as sm:
            loss_m, grad_m, el_m, tim_m, rr_m = timed_trajectories(sm, x, v, seed, n_steps, a.warmup, a.steps, "+MLP")
        vm = a.steps * n_steps / el_m
        out["mlp"] = {
            "value": round(vm, 2),
            "unit": "gradient steps/s",
            "ms_per_step": round(1e3 * el_m / a.steps, 3),
            "cost_ratio_vs_plain": round(el_m / elapsed, 4),
            "loss": loss_m,
            "dL_dtheta": grad_m.tolist(),
            "dL_dtheta_rel_diff_vs_plain": rel(grad_m, grad),
            "list_reruns_timed": rr_m,
            "phases_ms_one_trajectory": {k: round(vv["ms"], 3) for k, vv in tim_m.items()},
            "network": "16 RBF -> 64 -> 64 -> 1 SiLU, bf16 weights (Xavier, seed 15); table of f, f' on "
                       "512 nodes built by bf16 MFMA (activations split in three bf16 terms, FP32 "
                       "accumulation) per trajectory, cubic Hermite in the pair kernels",
        }

    if not a.no_mlp and not a.no_long and a.workload == "water32768":
        # The learned correction at scale: 5,000 forward steps, in 1,000-step loss blocks whose
        # reverse passes are truncated at 200 steps (the reference's training regime). Full
        # backprop over 5,000 steps is not defined in FP32: the reverse trajectory leaves the
        # forward one after ~500 steps and its friction-inverting integrator then runs away.
        lw = revsim.WORKLOADS["water32768_mlp"]
        T, blk, trunc = lw["n_steps"], 1000, 200
        legs = {}
        for name, mw in (("plain", None), ("mlp", mlp_init())):
            with make_sim(wl, extra_run, prec, theta, target, mlp=mw, n_steps_loss=blk, n_steps_trunc=trunc) as sl:
                sl.value_and_grad(x, v, seed, blk)  # warm-up: table, list capacities
                t0 = time.perf_counter()  # calls are synchronous on the library's stream
                L_l, g_l = sl.value_and_grad(x, v, seed, T)[:2]
                legs[name] = (time.perf_counter() - t0, L_l, g_l, sl.stats()["reruns_call"])
        out["long"] = {
            "config": f"{wl['n_molecules']} waters, {T} forward steps in {T // blk} loss blocks of {blk}, "
                      f"reverse truncated at {trunc} steps per block, RDF+MSD loss",
            "forward_steps_per_s": {k: round(T / vv[0], 1) for k, vv in legs.items()},
            "reverse_steps": (T // blk) * trunc,
            "mlp_cost_ratio_vs_plain": round(legs["mlp"][0] / legs["plain"][0], 4),
            "loss": {k: vv[1] for k, vv in legs.items()},
            "dL_dtheta": {k: vv[2].tolist() for k, vv in legs.items()},
            "dL_dtheta_rel_diff_mlp_vs_plain": rel(legs["mlp"][2], legs["plain"][2]),
            "list_reruns": {k: vv[3] for k, vv in legs.items()},
        }

    acc = {}
    H = min(a.oracle_steps, n_steps)
    if not a.no_cpu_baseline and prec == FP32:
        log(f"[bench] FP64 CPU oracle, {H} + {H} steps ...")
        Lo, go, t_orc, threads = oracle_gradient(wl, extra, theta, target, x, v, seed, H)
        vs = {}
        for name, p_ in (("fp32", FP32), ("fp64", FP64)):
            Lg, gg, tg = gpu_gradient(wl, extra, p_, theta, target, x, v, seed, H)
            vs[name] = {"normwise": rel(gg, go), "per_param": per_param(gg, go),
                        "loss_rel": abs(Lg - Lo) / abs(Lo), "loss_bit_exact": bool(Lg == Lo)}
        acc[f"vs_fp64_cpu_oracle_{H}_steps"] = dict(vs, grad_norm=float(np.linalg.norm(go)))
        log(f"[bench] dL/dθ vs FP64 oracle at {H} steps: FP32 {vs['fp32']['normwise']:.3e}, "
            f"FP64 {vs['fp64']['normwise']:.3e} (loss bit-exact: {vs['fp64']['loss_bit_exact']})")
        out["cpu_baseline"] = {
            "value": H / t_orc,
            "unit": "gradient steps/s",
            "cores": threads,
            "kind": "port",
            "cpu": cpu_model(),
            "sample": f"FP64 C restatement of the reference's threaded CPU loop (OpenMP, {threads} threads), "
                      f"same {wl['n_molecules']}-water start state, {H} forward + {H} reverse steps with the "
                      f"RDF + MSD loss over them, {t_orc:.1f} s wall",
        }
        out["speedup_vs_cpu_baseline"] = round(value / out["cpu_baseline"]["value"], 1)
        if not a.no_cpu_c1:
            c1 = dict(revsim.WORKLOADS["water3000"])
            n1 = c1.pop("n_steps")
            wb1 = revsim.water_box(c1["n_molecules"], c1["box"])
            log(f"[bench] CPU oracle, {c1['n_molecules']} waters, {n1} + {n1} steps ...")
            with revsim.Simulator(revsim.make_config(c1["n_molecules"], c1["box"], c1["cutoff"], precision=FP64),
                                  device=0) as s1:
                s1.set_theta(theta)
                t1 = revsim.perturbed_target(s1.rdf(wb1.coords)[0])
            L1, g1c, tc1, thr1 = oracle_gradient(c1, {}, theta, t1, wb1.coords, wb1.velocities, seed, n1)
            L1g, g1g, tg1 = gpu_gradient(c1, {}, FP64, theta, t1, wb1.coords, wb1.velocities, seed, n1)
            out["cpu_baseline"]["water3000_full_length"] = {
                "value": n1 / tc1, "unit": "gradient steps/s", "cores": thr1,
                "sample": f"{c1['n_molecules']} waters, {n1} + {n1} steps, RDF loss, {tc1:.1f} s wall",
                "gpu_fp64_same_run": {"dL_dtheta_rel_err": rel(g1g, g1c), "loss_bit_exact": bool(L1g == L1)},
            }

    if not a.no_fp64_check and prec == FP32:
        # the FP32 path against the FP64 path over the full length (the FP64 path is the oracle's
        # trajectory): chaos decorrelates them beyond ~200 steps
        for horizon in sorted({min(100, n_steps), n_steps}):
            _, g32, t32 = gpu_gradient(wl, extra, FP32, theta, target, x, v, seed, horizon)
            _, g64, t64 = gpu_gradient(wl, extra, FP64, theta, target, x, v, seed, horizon)
            acc.setdefault("fp32_vs_fp64_path_by_horizon", {})[str(horizon)] = {
                "normwise": rel(g32, g64), "per_param": per_param(g32, g64),
                "grad_norm_fp64": float(np.linalg.norm(g64))}
            log(f"[bench] dL/dθ FP32 vs FP64 path at {horizon} steps: {rel(g32, g64):.3e}")
        if not a.no_mlp:
            hm = min(100, n_steps)
            _, gm32, _ = gpu_gradient(wl, extra, FP32, theta, target, x, v, seed, hm, mlp=mlp_init())
            _, gm64, _ = gpu_gradient(wl, extra, FP64, theta, target, x, v, seed, hm, mlp=mlp_init())
            acc["mlp_fp32_vs_fp64_path"] = {str(hm): {"normwise": rel(gm32, gm64), "per_param": per_param(gm32, gm64)}}
            log(f"[bench] dL/dθ FP32 vs FP64 path at {hm} steps (+MLP): {rel(gm32, gm64):.3e}")
    if acc:
        acc["note"] = ("beyond ~200 steps the adjoint grows like exp(lambda t) (chaotic MD) and any two "
                       "differently rounded runs decorrelate: FP32 against FP64 here; the FP64 path "
                       "follows the oracle's own rounding, so it keeps matching it")
        out["accuracy"] = acc

    if a.calibrate_copy_mb:
        with make_sim(wl, extra_run, prec, theta, target) as sc:
            nb = a.calibrate_copy_mb << 20
            sc.copy_calibration(nb, 5)
        out["copy_calibration"] = {"bytes_read": nb, "bytes_written": nb, "launches": 5,
                                   "kernel": "k_copy16 (16-byte vector copy, librevsim)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
