"""CPU pins of the FP64 oracle (no GPU): each check is one a plausible mistake in the restatement
would fail — energy finite differences for forces, force finite differences for the HVP and the
θ-VJP, Hessian symmetry, numpy's Philox for the noise bits, closed forms for the MSD and the RDF
normalisation, reversibility, an independent textbook adjoint, and trajectory finite differences.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import rel_err

HERE = os.path.dirname(os.path.abspath(__file__))


def _orc(n=216, box=1.862, rc=0.8, **kw):
    import oracle as O
    import revsim

    o = O.Oracle(revsim.make_config(n, box, rc, **kw))
    o.set_theta(revsim.DEFAULT_THETA)
    return o


def test_philox_matches_numpy_golden(oracle_mod):
    d = json.load(open(os.path.join(HERE, "golden", "philox4x64.json")))
    for case in d["cases"]:
        ctr = [int(v, 16) for v in case["counter"]]
        key = [int(v, 16) for v in case["key"]]
        assert oracle_mod.philox4x64(ctr, key) == [int(v, 16) for v in case["out"]]


def test_noise_statistics(oracle_mod):
    import revsim

    o = _orc(n=3000, box=4.48, rc=1.0)
    xi = np.concatenate([o.noise(13, s) for s in (1, 2, 3)]).reshape(-1, 3)
    m = np.tile(revsim.WaterBox(np.zeros(27000), np.zeros(27000), np.ones(3)).masses(), 3)
    z = xi / np.sqrt(0.00831446262 * 300.0 / m)[:, None]
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.01
    # different steps and seeds give different streams; the same (seed, step) is reproducible
    assert not np.allclose(o.noise(13, 1), o.noise(13, 2))
    assert not np.allclose(o.noise(13, 1), o.noise(14, 1))
    assert np.array_equal(o.noise(13, 5), o.noise(13, 5))


def test_exp_polynomial_accuracy(oracle_mod):
    L = oracle_mod.lib()
    t = np.linspace(-4.0, 0.0, 2001)
    e64 = np.array([L.or_exp(float(v)) for v in t])
    assert np.max(np.abs(e64 / np.exp(t) - 1.0)) < 4e-16


def test_forces_are_minus_energy_gradient(jittered216):
    o = _orc()
    F, _ = o.forces(jittered216, with_energy=True)
    rng = np.random.default_rng(0)
    for _ in range(3):
        d = rng.standard_normal(jittered216.size)
        d /= np.linalg.norm(d)
        h = 1e-6
        Ep = o.forces(jittered216 + h * d, True)[1]
        Em = o.forces(jittered216 - h * d, True)[1]
        assert abs((Ep - Em) / (2 * h) + F @ d) < 1e-6 * np.linalg.norm(F)


@pytest.mark.parametrize("n,box,rc", [(216, 1.862, 0.8), (3000, 4.48, 1.0)])
def test_fp32_membership_rule_agrees_with_fp64(n, box, rc):
    """The FP32 pair rule selects the same interacting atom pairs as the FP64 rule up to rounding
    at the cutoff: forces from the two rules (both evaluated in FP64) agree except on the few atoms
    of a pair sitting within ~1e-7 nm of r_c. A rule that mishandles periodic images (in a box
    whose half-width is close to r_c + the molecule size) changes dozens."""
    import oracle as O
    import revsim

    wb = revsim.water_box(n, box)
    rng = np.random.default_rng(1)
    x = wb.coords + 0.01 * rng.standard_normal(wb.coords.size)
    F = {}
    for prec in (0, 1):
        o = O.Oracle(revsim.make_config(n, box, rc, precision=prec))
        o.set_theta(revsim.DEFAULT_THETA)
        F[prec] = o.forces(x)
    d = np.linalg.norm((F[1] - F[0]).reshape(-1, 3), axis=1)
    assert np.count_nonzero(d > 1e-9 * np.sqrt(np.mean(F[0] ** 2))) <= 4


def test_newton_third_law(jittered216):
    F = _orc().forces(jittered216).reshape(-1, 3)
    assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).sum()


def test_bonded_only_two_molecules_far_apart():
    """Two waters 0.9 nm apart in a 1.862 nm box, charges and LJ switched off: only the bond and
    angle terms act. A stretched bond along x and a bent angle give closed-form forces."""
    import oracle as O
    import revsim

    o = O.Oracle(revsim.make_config(2, 1.862, 0.8))
    th = revsim.DEFAULT_THETA.copy()
    th[1] = 0.0  # no LJ
    o.set_theta(th)
    # molecule 0: O at origin-ish, H1 along +x stretched to 0.1 nm, H2 at 90 degrees (0.09572 along y)
    x = np.zeros(18)
    c0 = np.array([0.3, 0.3, 0.3])
    x[0:3] = c0
    x[3:6] = c0 + [0.1, 0, 0]
    x[6:9] = c0 + [0, 0.09572, 0]
    x[9:12] = c0 + [0.9, 0.9, 0.9]
    x[12:15] = c0 + [0.9 + 0.09572 * math.sin(1.82421813418 / 2), 0.9, 0.9 + 0.09572 * math.cos(1.82421813418 / 2)]
    x[15:18] = c0 + [0.9 - 0.09572 * math.sin(1.82421813418 / 2), 0.9, 0.9 + 0.09572 * math.cos(1.82421813418 / 2)]
    # charges are in q_H: set a tiny q_H so RF is negligible but the config stays valid
    th[2] = 1e-12
    o.set_theta(th)
    F = o.forces(x).reshape(-1, 3)
    kb, r0, ka, t0 = th[3], th[4], th[5], 1.82421813418
    # H1: bond force -kb (0.1 - r0) along +x; angle force along +y direction of the opening
    ang = math.pi / 2
    fb = -kb * (0.1 - r0)
    fa_h1 = -ka * (ang - t0) / 0.1  # magnitude along -y for H1 (opening the angle moves H1 to -y)
    assert F[1, 0] == pytest.approx(fb, rel=1e-12)
    assert F[1, 1] == pytest.approx(-fa_h1, rel=1e-12)
    fa_h2 = -ka * (ang - t0) / 0.09572
    assert F[2, 0] == pytest.approx(-fa_h2, rel=1e-12)
    assert np.abs(F[0] + F[1] + F[2]).max() < 1e-9
    assert np.abs(F[3:]).max() < 1e-6  # the second molecule sits at rest geometry


def test_hvp_matches_force_finite_differences(jittered216, box216):
    o = _orc()
    m = np.repeat(box216.masses(), 3)
    rng = np.random.default_rng(2)
    u = rng.standard_normal(jittered216.size)
    _, h, _ = o.adjoint_terms(jittered216, u * m, np.zeros_like(u))
    e = 1e-6
    fd = (o.forces(jittered216 + e * u) - o.forces(jittered216 - e * u)) / (2 * e)
    assert rel_err(h, fd) < 1e-8


def test_hessian_is_symmetric(jittered216, box216):
    o = _orc()
    m = np.repeat(box216.masses(), 3)
    rng = np.random.default_rng(4)
    u, v = rng.standard_normal(jittered216.size), rng.standard_normal(jittered216.size)
    _, hu, _ = o.adjoint_terms(jittered216, u * m, np.zeros_like(u))
    _, hv, _ = o.adjoint_terms(jittered216, v * m, np.zeros_like(u))
    assert abs(v @ hu - u @ hv) < 1e-10 * np.linalg.norm(hu) * np.linalg.norm(v)


def test_theta_vjp_matches_finite_differences(jittered216):
    import revsim

    o = _orc()
    o.set_float_sums()  # steps of 1e-6 resolve below the fixed-point unit of the force sums
    rng = np.random.default_rng(5)
    w = rng.standard_normal(jittered216.size)
    _, _, vj = o.adjoint_terms(jittered216, np.zeros_like(w), w)
    for k in range(6):
        th = revsim.DEFAULT_THETA.copy()
        hk = 1e-6 * th[k]
        th[k] += hk
        o.set_theta(th)
        Fp = o.forces(jittered216)
        th[k] -= 2 * hk
        o.set_theta(th)
        Fm = o.forces(jittered216)
        o.set_theta(revsim.DEFAULT_THETA)
        fd = w @ (Fp - Fm) / (2 * hk)
        assert abs(vj[k] - fd) < 1e-6 * abs(fd), (k, vj[k], fd)


def test_lj_vjp_identities(jittered216):
    """LJ forces are linear in ε: w.dF/dε = w.F_LJ/ε; RF forces are quadratic in q_H:
    w.dF/dq_H = 2 w.F_RF/q_H (water/water.jl:184,1246)."""
    import revsim

    o = _orc()
    rng = np.random.default_rng(6)
    w = rng.standard_normal(jittered216.size)
    _, _, vj = o.adjoint_terms(jittered216, np.zeros_like(w), w)
    th = revsim.DEFAULT_THETA.copy()
    th_noLJ = th.copy()
    th_noLJ[1] = 0.0
    F = o.forces(jittered216)
    o.set_theta(th_noLJ)
    F_noLJ = o.forces(jittered216)
    th_noq = th.copy()
    th_noq[2] = 1e-30
    o.set_theta(th_noq)
    F_noq = o.forces(jittered216)
    o.set_theta(th)
    F_lj = F - F_noLJ
    F_rf = F - F_noq
    assert vj[1] == pytest.approx(w @ F_lj / th[1], rel=1e-9)
    assert vj[2] == pytest.approx(2 * (w @ F_rf) / th[2], rel=1e-9)


def test_molecule_pairs_brute_force(jittered216):
    from scipy.spatial import cKDTree

    o = _orc()
    r = 1.0
    P = o.molecule_pairs(jittered216, r)
    O_ = np.mod(jittered216.reshape(-1, 9)[:, :3], 1.862)
    want = np.array(sorted(cKDTree(O_, boxsize=1.862).query_pairs(r)))
    assert np.array_equal(P, want)


def test_msd_closed_form():
    """Every O drifts by δ per frame: MSD at lag K frames is |Kδ|², so D = K|δ|²/(6 Δt_frame)."""
    o = _orc(msd_window_steps=40, msd_every=5, w_msd=1.0)
    M = 216
    base = np.random.default_rng(0).uniform(0, 1.862, (M, 3))
    delta = np.array([0.01, -0.02, 0.015])
    nf, K, dtf = 8, 4, 5 * 0.001
    fr = np.concatenate([np.mod(base + f * delta, 1.862).reshape(-1) for f in range(nf)])
    D, L, _ = o.msd(fr)
    assert D == pytest.approx(K * (delta @ delta) / (6 * dtf), rel=1e-12)
    assert L == pytest.approx(1e18 * (1e-6 * D - 2.3e-9) ** 2, rel=1e-12)


def test_msd_gradient_finite_differences():
    o = _orc(msd_window_steps=40, msd_every=5, w_msd=1.0)
    rng = np.random.default_rng(1)
    fr = np.concatenate([rng.uniform(0, 1.862, 648) for _ in range(8)])
    fr = np.cumsum(np.concatenate([fr[:648], 0.02 * rng.standard_normal(7 * 648)])).copy()
    fr = np.mod(fr, 1.862)
    _, _, d = o.msd(fr)
    u = rng.standard_normal(fr.size)
    h = 1e-7
    fd = (o.msd(fr + h * u)[1] - o.msd(fr - h * u)[1]) / (2 * h)
    assert abs(fd - d @ u) < 1e-6 * abs(fd)


def test_rdf_weights_are_normalised(box216):
    """Two molecules: the histogram holds one distance; its weights sum to 1 (2^32 in fixed point)
    and are centred on the distance."""
    import oracle as O
    import revsim

    for prec in (0, 1):
        o = O.Oracle(revsim.make_config(2, 1.862, 0.8, precision=prec))
        o.set_theta(revsim.DEFAULT_THETA)
        x = np.zeros(18)
        x[0:3] = [0.2, 0.2, 0.2]
        x[3:6] = [0.25, 0.2, 0.2]
        x[6:9] = [0.2, 0.25, 0.2]
        x[9:12] = [0.2 + 0.4312, 0.2, 0.2]
        x[12:15] = [0.68, 0.2, 0.2]
        x[15:18] = [0.63, 0.25, 0.2]
        g, H, _, _ = o.rdf(x)
        # one rounding to 2^-32 per bin of the 21-bin window; FP32 weights are float quotients
        tol = 11 if prec == 0 else 2 ** 32 * 2e-6
        assert abs(int(H.sum()) - 2 ** 32) <= tol
        b = np.arange(200)
        centre = ((b + 0.5) * 0.004 * H).sum() / H.sum()
        assert abs(centre - 0.4312) < 2e-4


def test_rdf_ideal_gas_is_flat():
    """Uniformly random O positions (ideal gas): g(r) ~ 1 away from r = 0."""
    import oracle as O
    import revsim

    M, L = 3000, 4.48
    rng = np.random.default_rng(7)
    O_ = rng.uniform(0, L, (M, 3))
    x = np.repeat(O_, 3, axis=0)
    x[1::3, 0] += 0.01
    x[2::3, 1] += 0.01
    o = O.Oracle(revsim.make_config(M, L, 1.0))
    o.set_theta(revsim.DEFAULT_THETA)
    g, _, _, _ = o.rdf(np.ascontiguousarray(x.reshape(-1)))
    assert abs(g[60:190].mean() - 1.0) < 0.01
    assert np.abs(g[100:190] - 1.0).max() < 0.15


def test_rdf_gradient_finite_differences(jittered216):
    """Smooth window (0.1 nm): the analytic dL/dx matches central differences."""
    import revsim

    o = _orc(rdf_window=0.1)
    g0, _, _, _ = o.rdf(jittered216)
    o.set_target_rdf(revsim.perturbed_target(g0))
    _, _, _, d = o.rdf(jittered216)
    rng = np.random.default_rng(8)
    for _ in range(3):
        u = np.zeros_like(jittered216)
        u[0::9] = rng.standard_normal(216)
        u /= np.linalg.norm(u)
        h = 1e-6
        fd = (o.rdf(jittered216 + h * u)[2] - o.rdf(jittered216 - h * u)[2]) / (2 * h)
        assert abs(fd - d @ u) < 2e-3 * np.linalg.norm(d)


def test_reversibility_and_loss_agreement(box216):
    """water/water.jl:1613-1614: x0 recovered within 1e-9 nm and |L_rev - L_fwd| < 1e-7."""
    import revsim

    o = _orc()
    g0, _, _, _ = o.rdf(box216.coords)
    o.set_target_rdf(revsim.perturbed_target(g0))
    xN, vN, lf = o.forward(box216.coords, box216.velocities, 13, 100)
    lr, g, x0, v0 = o.reverse(xN, vN, 13, 100)
    d = x0 - np.mod(box216.coords, 1.862)
    d -= 1.862 * np.round(d / 1.862)
    assert np.abs(d).max() < 1e-9
    assert abs(lr - lf) < 1e-7
    assert rel_err(v0, box216.velocities) < 1e-9


@pytest.mark.parametrize("msd", [False, True])
def test_reversible_gradient_equals_stored_trajectory_adjoint(box216, msd):
    import revsim

    kw = dict(msd_window_steps=20, msd_every=5, w_msd=1e-3) if msd else {}
    o = _orc(**kw)
    g0, _, _, _ = o.rdf(box216.coords)
    o.set_target_rdf(revsim.perturbed_target(g0))
    L1, g1 = o.value_and_grad(box216.coords, box216.velocities, 13, 40)
    L2, g2 = o.value_and_grad_stored(box216.coords, box216.velocities, 13, 40)
    assert L1 == pytest.approx(L2, rel=1e-12)
    assert rel_err(g1, g2) < 1e-10


def test_gradient_matches_trajectory_finite_differences(box216):
    """Central differences of the 20-step trajectory loss (same noise) for σ, ε, q_H, the
    parameters the O-O RDF depends on most; smooth RDF window."""
    import revsim

    o = _orc(rdf_window=0.1)
    g0, _, _, _ = o.rdf(box216.coords)
    o.set_target_rdf(revsim.perturbed_target(g0))
    _, g = o.value_and_grad(box216.coords, box216.velocities, 13, 20)
    for k in (0, 1, 2):
        th = revsim.DEFAULT_THETA.copy()
        hk = 1e-6 * th[k]
        th[k] += hk
        o.set_theta(th)
        lp = o.forward(box216.coords, box216.velocities, 13, 20)[2]
        th[k] -= 2 * hk
        o.set_theta(th)
        lm = o.forward(box216.coords, box216.velocities, 13, 20)[2]
        o.set_theta(revsim.DEFAULT_THETA)
        fd = (lp - lm) / (2 * hk)
        assert abs(g[k] - fd) < 1e-3 * abs(fd), (k, g[k], fd)


def _single_block(o1, x0, v0, seed, t0, n):
    """Loss and gradient of an n-step single-block trajectory started from the forward state at
    step t0 of the seed's trajectory, with that trajectory's noise of steps t0+1..t0+n replayed."""
    xs, vs, _ = o1.forward(x0, v0, seed, t0) if t0 > 0 else (x0, v0, None)
    xi = np.concatenate([o1.noise(seed, t0 + k) for k in range(1, n + 1)])
    o1.set_external_noise(xi, n)
    L, g = o1.value_and_grad(xs, vs, seed, n)
    o1.set_external_noise(None, 0)
    return L, g


@pytest.mark.parametrize("n_log", [20, 0])
def test_truncated_blocks_equal_per_block_gradients(box216, n_log):
    """Reference training regime (water/water.jl:1491-1547): loss blocks of 20 steps, each block's
    reverse pass truncated after 10 steps, forward snapshots restored every n_log steps. Block b's
    gradient flows only through the last 10 steps before its end t_b, from the state at t_b - 10
    (restored from the forward pass), so the result must equal the mean over blocks of the
    single-block 10-step gradients started from the forward state at t_b - 10 with the same noise
    replayed; the loss is the mean of the block-end RDF losses. Snapshots every 20 steps and none
    (block boundaries are restored in either case) agree."""
    import revsim

    n_loss, n_trunc, T = 20, 10, 60
    o = _orc(n_steps_loss=n_loss, n_steps_trunc=n_trunc, n_steps_log=n_log)
    tgt = revsim.perturbed_target(o.rdf(box216.coords)[0])
    o.set_target_rdf(tgt)
    L, g = o.value_and_grad(box216.coords, box216.velocities, 13, T)
    o1 = _orc()
    o1.set_target_rdf(tgt)
    blocks = [_single_block(o1, box216.coords, box216.velocities, 13, tb - n_trunc, n_trunc)
              for tb in range(n_loss, T + 1, n_loss)]
    Lb = np.mean([b[0] for b in blocks])
    gb = np.mean([b[1] for b in blocks], axis=0)
    assert L == pytest.approx(Lb, rel=1e-12)
    assert rel_err(g, gb) < 1e-10, (g, gb)
    # a plausible misreading (truncation ignored: the full 20-step block adjoint) is far off
    gfull = np.mean([_single_block(o1, box216.coords, box216.velocities, 13, tb - n_loss, n_loss)[1]
                     for tb in range(n_loss, T + 1, n_loss)], axis=0)
    assert rel_err(gfull, gb) > 1e-3


def test_snapshot_restore_without_truncation_is_exact(box216):
    """Snapshots only reset reconstruction drift: with no truncation, restoring the forward state
    every 20 steps gives the same FP64 gradient as reconstructing all 60 steps (<= 1e-10), and both
    equal the mean of the three blocks' full-length gradients."""
    import revsim

    res = {}
    for n_log in (20, 0):
        o = _orc(n_steps_loss=20, n_steps_log=n_log)
        tgt = revsim.perturbed_target(o.rdf(box216.coords)[0])
        o.set_target_rdf(tgt)
        res[n_log] = o.value_and_grad(box216.coords, box216.velocities, 13, 60)
    assert res[20][0] == pytest.approx(res[0][0], rel=1e-12)
    assert rel_err(res[20][1], res[0][1]) < 1e-10
    # the block boundary resets the adjoint: each block's loss sees only its own 20 steps
    o1 = _orc()
    o1.set_target_rdf(tgt)
    gb = np.mean([_single_block(o1, box216.coords, box216.velocities, 13, tb - 20, 20)[1]
                  for tb in (20, 40, 60)], axis=0)
    assert rel_err(res[20][1], gb) < 1e-10


def _free_config(n, box, rc, **kw):
    """No interactions (Coulomb constant, ε, bond and angle constants all zero), so the integrator
    alone moves the atoms."""
    import revsim

    cfg = revsim.make_config(n, box, rc, w_rdf=0.0, **kw)
    cfg.coulomb_k = 0.0
    th = np.array(revsim.DEFAULT_THETA, dtype=np.float64)
    th[1] = th[3] = th[5] = 0.0  # ε_O, k_bond, k_angle
    return cfg, th


def free_particle_closed_form(x0, v0, n, dt, gamma, box):
    """Zero force and zero noise: v_n = e^n v_0 and, from the two half-step drifts of every step,
    x_n = x_0 + (dt/2) v_0 (1 + e)(1 - e^n)/(1 - e), e = exp(-γ dt) (a geometric series)."""
    e = np.exp(-gamma * dt)
    v = v0 * e ** n
    x = x0 + 0.5 * dt * v0 * (1.0 + e) * (1.0 - e ** n) / (1.0 - e)
    return np.mod(x.reshape(-1, 3), box).reshape(-1), v


def test_free_particle_closed_form(oracle_mod, box216):
    """The integrator's friction and half-step drifts against the closed form; molecules move
    rigidly (one velocity for their three atoms) so no bond is needed to keep them whole."""
    cfg, th = _free_config(216, 1.862, 0.8)
    o = oracle_mod.Oracle(cfg)
    o.set_theta(th)
    rng = np.random.default_rng(4)
    v0 = np.repeat(rng.normal(0.0, 0.4, (216, 1, 3)), 3, axis=1).reshape(-1)
    n = 50
    o.set_external_noise(np.zeros(n * v0.size), n)
    x, v, _ = o.forward(box216.coords, v0, 1, n)
    xc, vc = free_particle_closed_form(box216.coords, v0, n, cfg.dt, cfg.gamma, 1.862)
    d = x - xc
    d -= 1.862 * np.round(d / 1.862)
    assert np.abs(d).max() < 1e-12
    assert np.abs(v - vc).max() < 1e-13


def test_deterministic_elementary_functions_accuracy(oracle_mod):
    """rs_log, rs_sincos2pi, rs_acos and rs_rsqrt (IEEE operations and fixed polynomials or
    Newton steps, shared bit for bit with the GPU library) against extended-precision references
    over their ranges: within a few units in the last place (rs_rsqrt: 2; measured 1.7), and exact
    at the special points (log 1 = 0, sincos at quarter turns, acos(±1), acos 0, 1/√(4^k))."""
    import ctypes

    L = oracle_mod.lib()
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.random(20000) + 1e-300, 2.0 ** rng.uniform(-60, 10, 20000)])
    lg = np.array([L.or_log(float(v)) for v in x])
    ref = np.log(x.astype(np.longdouble))
    err = np.abs(lg - ref).astype(np.float64)
    assert np.all(err <= 3.0 * np.spacing(np.abs(ref.astype(np.float64))) + 1e-300)
    assert L.or_log(1.0) == 0.0
    sn, cs = ctypes.c_double(), ctypes.c_double()
    u = rng.random(20000)
    got = []
    for v in u:
        L.or_sincos2pi(float(v), ctypes.byref(sn), ctypes.byref(cs))
        got.append((sn.value, cs.value))
    got = np.array(got)
    # reference in extended precision (2πu itself rounds in double)
    ang = 2 * np.pi * u.astype(np.longdouble)
    ang += 2 * (np.longdouble("3.14159265358979323846264338327950288") - np.longdouble(np.pi)) * u
    assert np.max(np.abs(got[:, 0] - np.sin(ang).astype(np.float64))) < 3.5e-16
    assert np.max(np.abs(got[:, 1] - np.cos(ang).astype(np.float64))) < 3.5e-16
    for v, (s0, c0) in ((0.0, (0.0, 1.0)), (0.25, (1.0, 0.0)), (0.5, (0.0, -1.0)), (0.75, (-1.0, 0.0))):
        L.or_sincos2pi(v, ctypes.byref(sn), ctypes.byref(cs))
        assert (sn.value, cs.value) == (s0, c0)
    xa = np.concatenate([rng.uniform(-1, 1, 20000), [-1.0, -0.5, 0.0, 0.5, 1.0]])
    ac = np.array([L.or_acos(float(v)) for v in xa])
    acr = np.arccos(xa.astype(np.longdouble))
    assert np.all(np.abs(ac - acr).astype(np.float64) <= 3.0 * np.spacing(acr.astype(np.float64)))
    assert L.or_acos(1.0) == 0.0 and L.or_acos(-1.0) == np.pi and L.or_acos(0.0) == np.pi / 2
    # rs_rsqrt over squared pair distances (1e-4 .. 4 nm^2), a wide range of magnitudes, and the
    # exact squares whose reciprocal square roots are powers of two
    xr = np.concatenate([rng.uniform(1e-4, 4.0, 40000), 2.0 ** rng.uniform(-200, 200, 20000),
                         4.0 ** np.arange(-20, 21)])
    rr = np.array([L.or_rsqrt(float(v)) for v in xr])
    rref = 1 / np.sqrt(xr.astype(np.longdouble))
    ulp = np.spacing(rref.astype(np.float64))
    assert np.all(np.abs(rr - rref).astype(np.float64) <= 2.0 * ulp)
    assert np.all(rr[-41:] * np.sqrt(xr[-41:]) == 1.0)
    assert np.isnan(L.or_rsqrt(0.0))


@pytest.mark.parametrize("n,box,rc", [(216, 1.862, 0.8), (3000, 4.48, 1.0)])
def test_fixed_point_force_sums(n, box, rc):
    """The pair forces are summed as 64-bit integers of 2^-33 kJ/mol/nm: each term rounds by at
    most 2^-34, so the exact-integer forces differ from the same terms summed in FP64 by no more
    than (terms per atom) 2^-34 plus FP64 rounding, and they do not depend on the thread count (the
    integer sum is exact in any order), while FP64 sums do."""
    import subprocess
    import sys

    import revsim

    wb = revsim.water_box(n, box)
    rng = np.random.default_rng(3)
    x = np.ascontiguousarray(wb.coords + 0.005 * rng.standard_normal(wb.coords.size))
    o = _orc(n, box, rc)
    F = o.forces(x)
    o.set_float_sums()
    Ff = o.forces(x)
    nterms = 2 * 3 * 4.0 / 3.0 * np.pi * (rc + 0.2) ** 3 * (3 * n / box ** 3)
    assert np.abs(F - Ff).max() <= nterms * 2.0 ** -34 + 1e-12 * np.abs(Ff).max()
    assert np.abs(F - Ff).max() > 0  # the two sums do differ
    import os

    from conftest import ROOT

    paths = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "rev-sim_amd")]
    code = (
        "import sys, numpy as np; sys.path[:0] = %r\n"
        "import revsim, oracle as O\n"
        "o = O.Oracle(revsim.make_config(%d, %r, %r)); o.set_theta(revsim.DEFAULT_THETA)\n"
        "x = np.frombuffer(sys.stdin.buffer.read(), dtype=np.float64).copy()\n"
        "sys.stdout.buffer.write(o.forces(x).tobytes())\n" % (paths, n, box, rc))

    out = {}
    for th in ("1", "3"):
        r = subprocess.run([sys.executable, "-c", code], input=x.tobytes(), capture_output=True,
                           env=dict(os.environ, OMP_NUM_THREADS=th), check=True)
        out[th] = np.frombuffer(r.stdout, dtype=np.float64)
    assert np.array_equal(out["1"], out["3"]) and np.array_equal(out["1"], F)
