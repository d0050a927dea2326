"""CPU tests of the training loop's host logic (revsim.interface.train, water/water.jl:1623-1698),
run over the FP64 oracle through a small adapter with the Simulator's call surface.

Pins: Adam's first step has the closed form p1 = p0 - lr * g / (|g| + eps) elementwise (m̂ = g,
v̂ = g²), applied to the scaled parameters (scale p0, gradient g * p0) and to the MLP weights; the
clamp acts before Adam; the gradients logged per epoch are the oracle's at that epoch's
parameters. The training gain of the MLP correction is measured on the GPU (test_gpu_train.py)."""
import numpy as np
import pytest


class OracleSim:
    """The Simulator methods train() uses, over the oracle (tests only)."""

    def __init__(self, cfg):
        import oracle as O

        self.cfg = cfg
        self.o = O.Oracle(cfg)

    def set_theta(self, th):
        self.o.set_theta(th)

    def set_mlp(self, w):
        self.o.set_mlp(w)

    def set_target_rdf(self, g):
        self.o.set_target_rdf(g)

    def forward(self, x, v, seed, n):
        return self.o.forward(x, v, seed, n)

    def reweight(self, x, v, seed, n):
        return self.o.reweight(x, v, seed, n)

    def value_and_grad(self, x, v, seed, n, mlp_grad_size=0):
        L, g = self.o.value_and_grad(x, v, seed, n)
        if mlp_grad_size:
            return L, g, self.o.mlp_grad().astype(np.float32)
        return L, g

    def loss_terms(self):
        return self.o.loss_terms()


def _setup(use_mlp, n_steps_loss=0):
    import revsim

    wb = revsim.water_box(216, 1.862)
    cfg = revsim.make_config(216, 1.862, 0.8, use_mlp=use_mlp, n_steps_loss=n_steps_loss)
    sim = OracleSim(cfg)
    sim.set_theta(revsim.DEFAULT_THETA)
    g0 = sim.o.rdf(wb.coords)[0]
    sim.set_target_rdf(revsim.perturbed_target(g0))
    return sim, wb


def test_adam_closed_form_first_step():
    from revsim.interface import Adam

    rng = np.random.default_rng(0)
    p, g = rng.standard_normal(7), rng.standard_normal(7)
    lr = 1e-3
    opt = Adam(lr, 7)
    p1 = opt.step(p, g)
    assert np.allclose(p1, p - lr * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    # second step with the same gradient: m̂ = g, v̂ = g² again
    p2 = opt.step(p1, g)
    assert np.allclose(p2, p1 - lr * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)


@pytest.mark.parametrize("joint", [False, True])
def test_train_first_epoch_update(joint):
    import revsim
    from revsim import interface as I
    from revsim.system import mlp_init

    sim, wb = _setup(int(joint))
    p0 = np.array([0.315, 0.636, -0.834, 462750.0, 0.09572, 836.8, 1.82421813418])
    W0 = mlp_init().astype(np.float64)
    lr = 1e-3
    kw = dict(mlp_weights=W0) if joint else {}
    out = I.train(sim, wb.coords, wb.velocities, p0, 1, 10, learning_rate=lr, **kw)
    hist = out[-1]
    # the logged gradient is the oracle's at p0 with the epoch's seed
    rng = np.random.Generator(np.random.Philox(1))
    _, s_run = (int(k) for k in rng.integers(0, 2 ** 63, size=2))
    sim.set_theta(I.params_to_theta(p0))
    if joint:
        sim.set_mlp(W0.astype(np.float32))
        L, g, gW = sim.value_and_grad(wb.coords, wb.velocities, s_run, 10, mlp_grad_size=W0.size)
    else:
        L, g = sim.value_and_grad(wb.coords, wb.velocities, s_run, 10)
    gp = I.theta_grad_to_params(g)
    assert hist[0][0] == L and np.array_equal(hist[0][1], gp)
    # one Adam step on the scaled parameters, gradient clamped to ±1e3 first
    gs = np.clip(gp * p0, -1e3, 1e3)
    assert np.allclose(out[0], p0 * (1 - lr * gs / (np.abs(gs) + 1e-8)), rtol=1e-14, atol=0)
    if joint:
        gw = np.clip(gW.astype(np.float64), -1e3, 1e3)
        assert np.array_equal(hist[0][2], gW.astype(np.float64))
        assert np.allclose(out[1], W0 - lr * gw / (np.abs(gw) + 1e-8), rtol=0, atol=1e-12)


def test_train_rejects_mlp_without_use_mlp():
    from revsim import interface as I
    from revsim.system import mlp_init

    sim, wb = _setup(0)
    with pytest.raises(ValueError):
        I.train(sim, wb.coords, wb.velocities, np.ones(7), 1, 4, mlp_weights=mlp_init())
    sim.cfg.use_mlp = 1
    with pytest.raises(ValueError):
        I.train(sim, wb.coords, wb.velocities, np.ones(7), 1, 4, mlp_weights=np.zeros(10))


def test_theta_param_mapping_round_trip():
    from revsim import interface as I

    p = np.array([0.3, 0.6, -0.8, 4e5, 0.0957, 800.0, 1.9])
    th = I.params_to_theta(p)
    assert th[2] == 0.4 and th.size == 6
    g = np.arange(1.0, 7.0)
    gp = I.theta_grad_to_params(g)
    # L(q_O) = L(q_H = -q_O/2): dL/dq_O = -dL/dq_H / 2; θ0 held fixed
    assert gp[2] == -1.5 and gp[6] == 0.0 and np.array_equal(gp[[0, 1, 3, 4, 5]], g[[0, 1, 3, 4, 5]])


def test_train_with_reweighting_uses_the_estimator():
    """reweighting = true (water/water.jl:1625,1656-1659): the epoch's gradient is the reweighting
    estimate over its loss blocks."""
    import revsim
    from revsim import interface as I

    wb = revsim.water_box(216, 1.862)
    cfg = revsim.make_config(216, 1.862, 0.8, n_steps_loss=5, w_dhvap=1.0)
    sim = OracleSim(cfg)
    sim.set_theta(revsim.DEFAULT_THETA)
    sim.set_target_rdf(revsim.perturbed_target(sim.o.rdf(wb.coords)[0]))
    p0 = np.array([0.315, 0.636, -0.834, 462750.0, 0.09572, 836.8, 1.82421813418])
    _, hist = I.train(sim, wb.coords, wb.velocities, p0, 1, 10, reweighting=True)
    rng = np.random.Generator(np.random.Philox(1))
    _, s_run = (int(k) for k in rng.integers(0, 2 ** 63, size=2))
    sim.set_theta(I.params_to_theta(p0))
    L, g, _, _ = sim.reweight(wb.coords, wb.velocities, s_run, 10)
    assert hist[0][0] == L and np.array_equal(hist[0][1], I.theta_grad_to_params(g))


def test_loss_types_map_to_weights():
    """loss_fn's :rdf, :enth_vap, :both (water/water.jl:756-769) and the diffusion losses."""
    from revsim import interface as I

    for lt, (wr, wd, wm) in {"rdf": (1, 0, 0), ":enth_vap": (0, 1, 0), "both": (1, 1, 0),
                              "msd": (0, 0, 1), "rdf_msd": (1, 0, 1)}.items():
        kw = I._loss_cfg(lt, 1000, 100)
        assert (kw["w_rdf"], kw["w_dhvap"], kw.get("w_msd", 0.0)) == (wr, wd, wm), lt
    with pytest.raises(ValueError):
        I._check_types("lj", "energy")


@pytest.mark.parametrize("msd", [False, True])
def test_train_log_line_format(msd):
    """The per-epoch log line (water/water.jl:1673-1680; diffusion/diffusion.jl:1727-1733 with the
    MSD loss): epoch, loss, [D, l_diff,] l_rdf, l_ev, the 7 parameters, the 7 gradients; the terms
    are the oracle's for the epoch's trajectory and sum to the loss."""
    import revsim
    from revsim import interface as I

    wb = revsim.water_box(216, 1.862)
    kw = dict(msd_window_steps=10, msd_every=5, w_msd=1e-3) if msd else {}
    cfg = revsim.make_config(216, 1.862, 0.8, n_steps_loss=10, w_dhvap=0.5, **kw)
    sim = OracleSim(cfg)
    sim.set_theta(revsim.DEFAULT_THETA)
    sim.set_target_rdf(revsim.perturbed_target(sim.o.rdf(wb.coords)[0]))
    p0 = np.array([0.315, 0.636, -0.834, 462750.0, 0.09572, 836.8, 1.82421813418])
    lines = []
    _, hist = I.train(sim, wb.coords, wb.velocities, p0, 1, 20, log=lines.append)
    f = [float(v) for v in lines[0].split()]
    nh = 6 if msd else 4
    assert len(f) == nh + 14 and f[0] == 1 and f[1] == hist[0][0]
    rng = np.random.Generator(np.random.Philox(1))
    _, s_run = (int(k) for k in rng.integers(0, 2 ** 63, size=2))
    sim.set_theta(I.params_to_theta(p0))
    L, g = sim.value_and_grad(wb.coords, wb.velocities, s_run, 20)
    t = sim.loss_terms()
    want = ([t["D"], t["l_msd"]] if msd else []) + [t["l_rdf"], t["l_ev"]]
    assert f[2:nh] == pytest.approx(want, rel=1e-15)
    assert t["l_rdf"] > 0 and t["l_ev"] > 0 and (t["l_msd"] > 0) == msd
    assert L == pytest.approx(t["l_rdf"] + t["l_ev"] + t["l_msd"], rel=1e-12)
    assert f[nh:nh + 7] == pytest.approx(list(p0), rel=1e-15)
    assert f[nh + 7:] == pytest.approx(list(I.theta_grad_to_params(g)), rel=1e-15)
