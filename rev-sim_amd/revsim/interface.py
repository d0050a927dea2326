"""The reference's Julia call surface, restated over librevsim.

forward_sim / reverse_sim / reweighting_sim / train mirror water/water.jl:977-979 (forward_sim),
:1435-1436 (reverse_sim), :1107-1108 (reweighting_sim) and :1623-1698 (train): same argument order
and meaning, same return tuples, same argument errors. Differences, all at the boundary:

* `params` is the reference's LJ vector (σ, ε, q_O, k_b, r0, k_a, θ0) (water/water.jl:1572); the
  library's θ is (σ, ε, q_H, k_b, r0, k_a) with q_H = -q_O/2 and θ0 a fixed configuration value, so
  gradients come back as dL/dq_O = -dL/dq_H/2 and dL/dθ0 = 0.
* `noise_seeds`: the reference draws one Xoshiro seed per step; the library's noise is a pure
  function of (seed, step, atom), so one 64-bit seed replaces the vector (the first element is used
  when a vector is passed).
* `boundary` is a box-edge length or 3-vector (nm); `coords`/`velocities` are (N,3) or 3N arrays.
* `n_threads` is accepted and ignored (one GPU per context).
* `fwd_loggers` is the handle returned in the forward System: the library keeps the forward
  snapshots and MSD frames on the device.
* `loss_type` :rdf, :enth_vap and :both as in loss_fn (water/water.jl:756-769), with the RDF term
  weighted 1 (the reference's rdf_loss_weighting folds into w_rdf), plus :msd and :rdf_msd for the
  diffusion loss (diffusion/diffusion.jl:897-911).
* `use_barostat` runs the Monte Carlo barostat every 30 steps at 1 bar (MonteCarloBarostatNoUnits
  defaults, water/water.jl:803-807,1136); the returned System carries the new box.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import _abi
from .simulator import Simulator
from .system import DEFAULT_THETA, THETA0

INTER_TYPES = ("lj",)
LOSS_TYPES = ("rdf", "enth_vap", "both", "msd", "rdf_msd")
BAROSTAT_EVERY = 30      # MonteCarloBarostatNoUnits n_steps (water/water.jl:803)
BAROSTAT_PRESSURE = 1.0  # bar (water/water.jl:1136)


@dataclasses.dataclass
class ForwardRecords:
    """What forward_sim leaves for reverse_sim (the reference's fwd_loggers): the context holding
    the trajectory's snapshots and MSD frames on the device, and the trajectory's identity."""
    sim: Simulator
    seed: int
    n_steps: int
    theta: np.ndarray
    loss_type: str


@dataclasses.dataclass
class System:
    coords: np.ndarray
    velocities: np.ndarray
    boundary: np.ndarray
    loggers: object = None


def params_to_theta(params) -> np.ndarray:
    """(σ, ε, q_O, k_b, r0, k_a[, θ0]) -> (σ, ε, q_H, k_b, r0, k_a)."""
    p = np.asarray(params, dtype=np.float64)
    if p.size not in (6, 7):
        raise ValueError("params must be (σ, ε, q_O, k_b, r0, k_a, θ0)")
    return np.array([p[0], p[1], -p[2] / 2.0, p[3], p[4], p[5]])


def theta_grad_to_params(g) -> np.ndarray:
    """dL/d(σ, ε, q_H, k_b, r0, k_a) -> dL/d(σ, ε, q_O, k_b, r0, k_a, θ0) with θ0 held fixed."""
    g = np.asarray(g, dtype=np.float64)
    return np.array([g[0], g[1], -g[2] / 2.0, g[3], g[4], g[5], 0.0])


def _box(boundary, default=None):
    if boundary is None:
        if default is None:
            raise ValueError("boundary must be given")
        return default
    b = np.asarray(boundary, dtype=np.float64).reshape(-1)
    return np.repeat(b, 3) if b.size == 1 else b


def _seed(noise_seeds) -> int:
    if np.isscalar(noise_seeds):
        return int(noise_seeds) & 0xFFFFFFFFFFFFFFFF
    return int(np.asarray(noise_seeds).reshape(-1)[0]) & 0xFFFFFFFFFFFFFFFF


def _check_types(inter_type, loss_type):
    if str(inter_type).lstrip(":") not in INTER_TYPES:
        raise ValueError("inter_type not recognised")  # water/water.jl:148
    if str(loss_type).lstrip(":") not in LOSS_TYPES:
        raise ValueError("loss_type not recognised")  # water/water.jl:766


class Context:
    """Owns the Simulator that forward_sim and reverse_sim share (the reference rebuilds its System
    on each call; here the device context persists so the forward records stay resident)."""

    def __init__(self, n_molecules: int, boundary, cutoff: float = 1.0, precision=_abi.RS_PRECISION_FP32,
                 device: int = 0, **cfg):
        self.box = _box(boundary)
        self.base = dict(n_molecules=n_molecules, cutoff=cutoff, precision=precision, **cfg)
        self.device = device
        self.sim = None
        self.key = None

    def simulator(self, boundary=None, **overrides) -> Simulator:
        """The context's Simulator for this configuration, at `boundary` (nm) when given (the
        reference passes the box with each call; a barostat run moves it)."""
        if boundary is not None:
            self.box = _box(boundary)
        kw = dict(self.base)
        kw.update(overrides)
        key = tuple(sorted(kw.items()))
        if self.sim is None or key != self.key:
            if self.sim is not None:
                self.sim.close()
            n, rc = kw.pop("n_molecules"), kw.pop("cutoff")
            cfg = _abi.make_config(n, list(self.box), rc, **kw)
            self.sim = Simulator(cfg, device=self.device)
            self.key = key
        elif not np.array_equal(self.sim.get_box(), self.box):
            self.sim.set_box(self.box)
        return self.sim


def _loss_cfg(loss_type, n_steps, n_steps_loss, n_steps_trunc=None, clip=math.inf, msd_window=500,
              msd_every=50):
    lt = str(loss_type).lstrip(":")
    kw = dict(n_steps_loss=int(n_steps_loss) if n_steps_loss and n_steps_loss != n_steps else 0,
              w_rdf=1.0 if lt in ("rdf", "both", "rdf_msd") else 0.0,
              w_dhvap=1.0 if lt in ("enth_vap", "both") else 0.0,
              clip_norm=float(clip))
    if n_steps_trunc is not None:
        kw["n_steps_trunc"] = 0 if n_steps_trunc == (n_steps_loss or n_steps) else int(n_steps_trunc)
    if lt in ("msd", "rdf_msd"):
        w = min(msd_window, n_steps)
        w -= w % (2 * msd_every)
        if w < 2 * msd_every:
            raise ValueError("trajectory too short for the MSD loss window")
        kw.update(msd_window_steps=w, msd_every=msd_every, w_msd=1.0)
    return kw


def forward_sim(coords, velocities, boundary, noise_seeds, gamma, n_steps, n_steps_loss, params, dt,
                n_threads=None, inter_type="lj", loss_type="rdf", run_grads=False, calc_loss=True,
                use_barostat=False, *, context: Context, target_rdf=None):
    """water/water.jl:977-1105 with run_grads=false: returns (System, loss, grads=None)."""
    _check_types(inter_type, loss_type)
    if run_grads:
        raise ValueError("run_grads (the direct-differentiation gradient) is not part of this library")
    if n_steps_loss and n_steps % n_steps_loss:
        raise ValueError("n_steps_loss must divide n_steps")
    kw = _loss_cfg(loss_type, n_steps, n_steps_loss)
    if not calc_loss:
        kw.update(w_rdf=0.0, w_dhvap=0.0, w_msd=0.0, msd_window_steps=0, msd_every=0)
    if use_barostat:
        kw.update(barostat_every=BAROSTAT_EVERY, barostat_pressure=BAROSTAT_PRESSURE)
    sim = context.simulator(boundary, gamma=float(gamma), dt=float(dt), **kw)
    sim.set_theta(params_to_theta(params))
    if target_rdf is not None:
        sim.set_target_rdf(target_rdf)
    x, v, loss = sim.forward(np.asarray(coords, dtype=np.float64).reshape(-1),
                             np.asarray(velocities, dtype=np.float64).reshape(-1), _seed(noise_seeds), n_steps)
    box = sim.get_box()
    context.box = box.copy()
    theta = params_to_theta(params)
    sys = System(x.reshape(-1, 3), v.reshape(-1, 3), box,
                 loggers=ForwardRecords(sim, _seed(noise_seeds), n_steps, theta, str(loss_type).lstrip(":")))
    return sys, loss, None


def reweighting_sim(coords, velocities, boundary, noise_seeds, gamma, n_steps, n_steps_loss, params, dt,
                    n_threads=None, inter_type="lj", loss_type="rdf", use_barostat=False, *,
                    context: Context, target_rdf=None):
    """water/water.jl:1107-1210: one forward trajectory with the ensemble-reweighting gradient
    estimator over its loss blocks; returns (System, loss, dl_dp, l_rdf, l_ev) with dl_dp in the
    reference's 7-parameter layout and l_rdf, l_ev the block-averaged terms (water/water.jl:1204-1209)."""
    _check_types(inter_type, loss_type)
    if str(loss_type).lstrip(":") in ("msd", "rdf_msd"):
        raise ValueError("loss_type not recognised")  # reweighting_sim has the water losses only
    if n_steps_loss and n_steps % n_steps_loss:
        raise ValueError("n_steps_loss must divide n_steps")
    kw = _loss_cfg(loss_type, n_steps, n_steps_loss)
    if use_barostat:
        kw.update(barostat_every=BAROSTAT_EVERY, barostat_pressure=BAROSTAT_PRESSURE)
    sim = context.simulator(boundary, gamma=float(gamma), dt=float(dt), **kw)
    sim.set_theta(params_to_theta(params))
    if target_rdf is not None:
        sim.set_target_rdf(target_rdf)
    loss, g, x, v = sim.reweight(np.asarray(coords, dtype=np.float64).reshape(-1),
                                 np.asarray(velocities, dtype=np.float64).reshape(-1), _seed(noise_seeds), n_steps)
    box = sim.get_box()
    context.box = box.copy()
    t = sim.loss_terms()
    return System(x.reshape(-1, 3), v.reshape(-1, 3), box), loss, theta_grad_to_params(g), t["l_rdf"], t["l_ev"]


def reverse_sim(coords, velocities, boundary, noise_seeds, gamma, n_steps, n_steps_loss, n_steps_trunc, params,
                dt, n_threads, fwd_loggers, clip_norm_threshold, inter_type="lj", loss_type="rdf"):
    """water/water.jl:1435-1558: returns (System, loss, dl_dp, l_rdf, l_ev) with dl_dp in the
    reference's 7-parameter layout; for the MSD loss types the diffusion variant's tuple
    (System, loss, dl_dp, D, l_diff, l_rdf, l_ev), diffusion/diffusion.jl:1616, D in m^2/s.
    `fwd_loggers` is the forward System's `loggers` (ForwardRecords). The truncation and the
    clip threshold are reverse_sim's own arguments: they are passed to the context here and do not
    disturb the forward records."""
    _check_types(inter_type, loss_type)
    n_loss = n_steps_loss or n_steps
    if n_steps_trunc > n_loss:
        raise ValueError("n_steps_trunc must not be greater than n_steps_loss")  # water/water.jl:1437
    if not isinstance(fwd_loggers, ForwardRecords):
        raise ValueError("fwd_loggers from forward_sim are required")
    rec = fwd_loggers
    if rec.n_steps != n_steps or rec.seed != _seed(noise_seeds):
        raise ValueError("fwd_loggers belong to a different trajectory")
    if not np.array_equal(rec.theta, params_to_theta(params)):
        raise ValueError("params differ from those of the forward trajectory")
    lt = str(loss_type).lstrip(":")
    if lt != rec.loss_type:
        raise ValueError("loss_type differs from that of the forward trajectory")
    sim = rec.sim
    trunc = 0 if n_steps_trunc in (None, n_loss) else int(n_steps_trunc)
    sim.set_reverse_options(trunc, float(clip_norm_threshold))
    loss, g, x0, v0 = sim.reverse(np.asarray(coords, dtype=np.float64).reshape(-1),
                                  np.asarray(velocities, dtype=np.float64).reshape(-1), rec.seed, n_steps)
    t = sim.loss_terms()
    sys = System(x0.reshape(-1, 3), v0.reshape(-1, 3), _box(boundary, sim.get_box()))
    if lt in ("msd", "rdf_msd"):
        return sys, loss, theta_grad_to_params(g), t["D"], t["l_msd"], t["l_rdf"], t["l_ev"]
    return sys, loss, theta_grad_to_params(g), t["l_rdf"], t["l_ev"]


def train_log_line(epoch, loss, terms, params, grads, msd=False) -> str:
    """One line of train_<run_n>.log: "epoch loss l_rdf l_ev params... grads..." as written by
    water/water.jl:1673-1680, or with the MSD loss "epoch loss D l_diff l_rdf l_ev ..." as written
    by diffusion/diffusion.jl:1727-1733."""
    head = [epoch, loss] + ([terms["D"], terms["l_msd"]] if msd else []) + [terms["l_rdf"], terms["l_ev"]]
    return " ".join(map(str, head)) + " " + " ".join(map(str, params)) + " " + " ".join(map(str, grads))


class Adam:
    """Adam (Kingma & Ba 2015) with Optimisers.jl's defaults (β = (0.9, 0.999), ε = 1e-8)."""

    def __init__(self, lr: float, n: int, beta1=0.9, beta2=0.999, eps=1e-8):
        self.lr, self.b1, self.b2, self.eps = lr, beta1, beta2, eps
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.t = 0

    def step(self, p: np.ndarray, g: np.ndarray) -> np.ndarray:
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return p - self.lr * mh / (np.sqrt(vh) + self.eps)


def train(sim: Simulator, x0, v0, params_start, n_epochs: int, n_steps: int, n_steps_equil: int = 0,
          learning_rate: float = 2e-3, grad_clamp_val: float = 1e3, scale_params: bool = True,
          seed: int = 1, log=None, mlp_weights=None, eval_fn=None, reweighting: bool = False,
          equil_sim: Simulator | None = None, velocity_temperature: float | None = None):
    """The epoch loop of water/water.jl:1623-1698 on one context: fresh noise seed per epoch,
    optional equilibration, forward + reverse, a log line per epoch, Adam on parameters scaled by
    their starting values with the gradient clamped to ±grad_clamp_val (water/water.jl:1682-1696).

    With `mlp_weights` (packed 5313-value vector) the radial-basis MLP correction on O-O pairs is
    trained jointly: one Adam state over (scaled θ, MLP weights), the weights kept in FP64 as the
    master copy and uploaded (rounded to bfloat16 by the library) before every epoch. The context
    must have been created with use_mlp = 1.

    As in the reference's loop: with `reweighting` the gradient is the ensemble-reweighting
    estimate (reweighting_sim) instead of the reversible one; `equil_sim` (e.g. a context with the
    barostat on) equilibrates each epoch from (x0, v0) and the epoch's trajectory then runs at the
    box it reached (boundary_start, water/water.jl:1651-1669); `velocity_temperature` draws fresh
    Maxwell-Boltzmann velocities each epoch (random_velocity, water/water.jl:1648).

    Returns (params, history) — or (params, mlp_weights, history) when training the MLP — with
    history the per-epoch (loss, dL/dparams[, dL/dW]) and, if `eval_fn` is given, its value on the
    epoch's parameters appended."""
    if reweighting and mlp_weights is not None:
        raise ValueError("the reweighting estimator covers the classical parameters only")
    params = np.asarray(params_start, dtype=np.float64).copy()
    p0 = params.copy()
    joint = mlp_weights is not None
    if joint:
        if not sim.cfg.use_mlp:
            raise ValueError("MLP training needs a context created with use_mlp = 1")
        W = np.asarray(mlp_weights, dtype=np.float64).reshape(-1).copy()
        if W.size != _abi.MLP_NW:
            raise ValueError(f"mlp_weights must hold {_abi.MLP_NW} values")
    n_p = params.size
    opt = Adam(learning_rate, n_p + (W.size if joint else 0))
    scaled = np.ones_like(params)
    rng = np.random.Generator(np.random.Philox(seed))
    history = []
    x, v = np.asarray(x0, dtype=np.float64).reshape(-1), np.asarray(v0, dtype=np.float64).reshape(-1)
    x0_box = equil_sim.get_box() if equil_sim is not None else None
    for epoch in range(1, n_epochs + 1):
        s_eq, s_run = (int(k) for k in rng.integers(0, 2 ** 63, size=2))
        sim.set_theta(params_to_theta(params))
        if joint:
            sim.set_mlp(W.astype(np.float32))
        xs, vs = x, v
        if velocity_temperature is not None:
            from .io import maxwell_boltzmann

            vs = maxwell_boltzmann(x.size // 9, velocity_temperature, s_eq ^ 0x5A5A)
        if n_steps_equil:
            if equil_sim is not None:
                equil_sim.set_theta(params_to_theta(params))
                if joint:
                    equil_sim.set_mlp(W.astype(np.float32))
                equil_sim.set_box(x0_box)
                xs, vs, _ = equil_sim.forward(xs, vs, s_eq, n_steps_equil)
                sim.set_box(equil_sim.get_box())
            else:
                xs, vs, _ = sim.forward(xs, vs, s_eq, n_steps_equil)
        if reweighting:
            loss, g, _, _ = sim.reweight(xs, vs, s_run, n_steps)
        elif joint:
            loss, g, gW = sim.value_and_grad(xs, vs, s_run, n_steps, mlp_grad_size=W.size)
            gW = gW.astype(np.float64)
        else:
            loss, g = sim.value_and_grad(xs, vs, s_run, n_steps)
        grads = theta_grad_to_params(g)[:n_p]
        rec = [loss, grads.copy()] + ([gW.copy()] if joint else [])
        if eval_fn is not None:
            rec.append(eval_fn(params, W if joint else None))
        history.append(tuple(rec))
        if log is not None:
            log(train_log_line(epoch, loss, sim.loss_terms(), params, grads,
                               msd=bool(sim.cfg.w_msd and sim.cfg.msd_window_steps)))
        gp = grads * p0 if scale_params else grads
        gall = np.clip(np.concatenate([gp, gW]) if joint else gp, -grad_clamp_val, grad_clamp_val)
        pall = np.concatenate([scaled if scale_params else params, W]) if joint else (
            scaled if scale_params else params)
        pall = opt.step(pall, gall)
        if scale_params:
            scaled = pall[:n_p]
            params = scaled * p0
        else:
            params = pall[:n_p].copy()
        if joint:
            W = pall[n_p:].copy()
    return (params, W, history) if joint else (params, history)


__all__ = ["System", "ForwardRecords", "Context", "forward_sim", "reverse_sim", "reweighting_sim", "train",
           "train_log_line", "Adam", "params_to_theta",
           "theta_grad_to_params", "DEFAULT_THETA", "THETA0"]
