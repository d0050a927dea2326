"""CPU pins of the oracle's radial-basis MLP on O-O pairs (exact per-pair FP64 evaluation):
derivatives by finite differences, the weight VJP by finite differences in weight space, forces
and HVP of an MLP-only system, and the trajectory weight gradient against the stored-trajectory
adjoint."""
import numpy as np
import pytest

from conftest import rel_err


def _mlp_only(n=216, box=1.862, rc=0.8, **kw):
    """Charges ~0 and no LJ: only bonds, angles and the MLP act (no cutoff-crossing jumps)."""
    import oracle as O
    import revsim
    from revsim.system import mlp_init

    o = O.Oracle(revsim.make_config(n, box, rc, **kw))
    th = revsim.DEFAULT_THETA.copy()
    th[1] = 0.0
    th[2] = 1e-9
    o.set_theta(th)
    W = mlp_init().astype(np.float64)
    o.set_mlp(W, exact=True)
    return o, W


def test_mlp_derivatives():
    o, _ = _mlp_only()
    r = np.linspace(0.22, 0.79, 13)
    u, u1, u2 = o.mlp_eval(r)
    h = 1e-6
    assert np.allclose(u1, (o.mlp_eval(r + h)[0] - o.mlp_eval(r - h)[0]) / (2 * h), rtol=1e-6, atol=1e-8)
    assert np.allclose(u2, (o.mlp_eval(r + h)[1] - o.mlp_eval(r - h)[1]) / (2 * h), rtol=1e-6, atol=1e-6)
    # smooth cutoff: the force vanishes at r_c
    assert abs(o.mlp_eval(np.array([0.8]))[1][0]) < 1e-12


def test_mlp_weight_vjp():
    o, W = _mlp_only()
    rng = np.random.default_rng(0)
    r = rng.uniform(0.24, 0.8, 9)
    g1, g2 = rng.standard_normal(9), rng.standard_normal(9)
    dW = o.mlp_vjp(r, g1, g2)
    d = rng.standard_normal(W.size)
    e = 1e-6
    o.set_mlp(W + e * d, exact=True)
    _, a1, a2 = o.mlp_eval(r)
    o.set_mlp(W - e * d, exact=True)
    _, b1, b2 = o.mlp_eval(r)
    fd = (g1 @ (a1 - b1) + g2 @ (a2 - b2)) / (2 * e)
    assert abs(dW @ d - fd) < 1e-6 * abs(fd)


def test_bf16_weights():
    """set_mlp stores round-to-nearest-even bfloat16 values."""
    import oracle as O
    import revsim
    from revsim.system import mlp_init

    w = mlp_init()
    o = O.Oracle(revsim.make_config(216, 1.862, 0.8))
    o.set_mlp(w)
    b = w.view(np.uint32).astype(np.uint64)
    b = (b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000
    w16 = b.astype(np.uint32).view(np.float32).astype(np.float64)
    o2 = O.Oracle(revsim.make_config(216, 1.862, 0.8))
    o2.set_mlp(w16, exact=True)
    r = np.linspace(0.3, 0.7, 5)
    assert np.array_equal(o.mlp_eval(r)[1], o2.mlp_eval(r)[1])


def test_mlp_system_forces_hvp_and_weight_vjp(jittered216, box216):
    o, W = _mlp_only()
    o.set_float_sums()  # steps of 1e-6 resolve below the fixed-point unit of the force sums
    x = jittered216
    F, _ = o.forces(x, with_energy=True)
    rng = np.random.default_rng(1)
    h = 1e-6
    d = rng.standard_normal(x.size)
    d /= np.linalg.norm(d)
    fd = (o.forces(x + h * d, True)[1] - o.forces(x - h * d, True)[1]) / (2 * h)
    assert abs(fd + F @ d) < 1e-6 * np.linalg.norm(F)
    m = np.repeat(box216.masses(), 3)
    u = rng.standard_normal(x.size)
    w = rng.standard_normal(x.size)
    _, hv, _, dW = o.adjoint_terms_mlp(x, u * m, w)
    fdh = (o.forces(x + h * u) - o.forces(x - h * u)) / (2 * h)
    assert rel_err(hv, fdh) < 1e-7
    dd = rng.standard_normal(W.size)
    e = 1e-6
    o.set_mlp(W + e * dd, exact=True)
    Fp = o.forces(x)
    o.set_mlp(W - e * dd, exact=True)
    Fm = o.forces(x)
    fdw = w @ (Fp - Fm) / (2 * e)
    assert abs(dW @ dd - fdw) < 1e-6 * abs(fdw)


def test_mlp_trajectory_weight_gradient(box216):
    """Reversible dL/dW equals the stored-trajectory adjoint's, and central differences along a
    random weight direction (RDF loss with a smooth window, 15 steps)."""
    import revsim

    o, W = _mlp_only(rdf_window=0.1)
    th = revsim.DEFAULT_THETA.copy()
    o.set_theta(th)  # full force field + MLP
    g0, _, _, _ = o.rdf(box216.coords)
    o.set_target_rdf(revsim.perturbed_target(g0))
    L, g = o.value_and_grad(box216.coords, box216.velocities, 13, 15)
    dW = o.mlp_grad()
    Ls, gs = o.value_and_grad_stored(box216.coords, box216.velocities, 13, 15)
    dWs = o.mlp_grad()
    assert rel_err(dW, dWs) < 1e-9
    assert rel_err(g, gs) < 1e-9
    rng = np.random.default_rng(3)
    dd = rng.standard_normal(W.size)
    e = 1e-5
    o.set_mlp(W + e * dd, exact=True)
    lp = o.forward(box216.coords, box216.velocities, 13, 15)[2]
    o.set_mlp(W - e * dd, exact=True)
    lm = o.forward(box216.coords, box216.velocities, 13, 15)[2]
    fd = (lp - lm) / (2 * e)
    assert abs(dW @ dd - fd) < 2e-3 * abs(fd), (dW @ dd, fd)
