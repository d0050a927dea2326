"""CPU pins of the oracle's potential energy, enthalpy-of-vaporization loss and ensemble-reweighting
estimator (water/water.jl:738-754,1107-1210): central finite differences in θ and in x, an explicit
energy sum for two waters, trajectory finite differences of the enthalpy-of-vaporization loss,
agreement with the stored-trajectory adjoint, and the reweighting estimator rebuilt from the
per-block states."""
import math

import numpy as np
import pytest

from conftest import rel_err

RT300 = 0.00831446262 * 300.0


def _orc(n=216, box=1.862, rc=0.8, **kw):
    import oracle as O
    import revsim

    o = O.Oracle(revsim.make_config(n, box, rc, **kw))
    o.set_theta(revsim.DEFAULT_THETA)
    return o


def test_energy_theta_gradient_matches_finite_differences(jittered216):
    import revsim

    o = _orc()
    U, dU, _ = o.energy(jittered216)
    for k in range(6):
        th = revsim.DEFAULT_THETA.copy()
        h = 1e-6 * abs(th[k])
        th[k] += h
        o.set_theta(th)
        up = o.energy(jittered216)[0]
        th[k] -= 2 * h
        o.set_theta(th)
        um = o.energy(jittered216)[0]
        fd = (up - um) / (2 * h)
        assert dU[k] == pytest.approx(fd, rel=1e-6, abs=1e-9 * abs(U)), (k, dU[k], fd)
    o.set_theta(revsim.DEFAULT_THETA)


def test_energy_forces_are_minus_gradient(jittered216):
    """F = -dU/dx along random directions, with the reaction field shifted to zero at r_c (a pair
    crossing r_c does not make the energy jump)."""
    o = _orc()
    U, _, F = o.energy(jittered216)
    rng = np.random.default_rng(2)
    for _ in range(3):
        d = rng.standard_normal(jittered216.size)
        d /= np.linalg.norm(d)
        h = 1e-6
        up = o.energy(jittered216 + h * d)[0]
        um = o.energy(jittered216 - h * d)[0]
        assert abs((up - um) / (2 * h) + F @ d) < 1e-6 * np.linalg.norm(F)


def test_energy_of_two_waters_explicit_sum():
    """Two waters whose O-O distance is 0.5 nm: the energy is the sum, written out here, of two
    bonds and an angle per molecule, nine reaction-field pairs (shifted to zero at r_c) and one
    O-O Lennard-Jones pair; its θ-gradient likewise."""
    import revsim

    box, rc, eps_rf, kc = 3.0, 1.0, 78.3, 138.935457644
    o = _orc(n=2, box=box, rc=rc)
    th = revsim.DEFAULT_THETA
    sig, eps, qH, kb, r0, ka = th
    t0 = 1.82421813418
    rng = np.random.default_rng(4)
    x = np.zeros((6, 3))
    for m, c in enumerate(([1.0, 1.0, 1.0], [1.5, 1.0, 1.0])):
        x[3 * m] = c
        for hh, (rb, ang) in enumerate(((0.097, 0.3), (0.094, 0.3 + 1.9))):
            x[3 * m + 1 + hh] = np.array(c) + rb * np.array([0.2 * m, math.cos(ang), math.sin(ang)]) / \
                math.sqrt(1 + 0.04 * m * m)
    x += 1e-3 * rng.standard_normal(x.shape)
    U, dU, _ = o.energy(x.ravel())
    krf = (1 / rc**3) * (eps_rf - 1) / (2 * eps_rf + 1)
    crf = (1 / rc) * 3 * eps_rf / (2 * eps_rf + 1)
    q = [-2 * qH, qH, qH]
    Ue, d = 0.0, np.zeros(6)
    for m in range(2):
        O = x[3 * m]
        vs = [x[3 * m + 1] - O, x[3 * m + 2] - O]
        for v in vs:
            r = np.linalg.norm(v)
            Ue += 0.5 * kb * (r - r0) ** 2
            d[3] += 0.5 * (r - r0) ** 2
            d[4] += -kb * (r - r0)
        a = math.acos(vs[0] @ vs[1] / (np.linalg.norm(vs[0]) * np.linalg.norm(vs[1])))
        Ue += 0.5 * ka * (a - t0) ** 2
        d[5] += 0.5 * (a - t0) ** 2
    for a in range(3):
        for b in range(3):
            r = np.linalg.norm(x[3 + b] - x[a])
            e = kc * q[a] * q[b] * (1 / r + krf * r * r - crf)
            Ue += e
            d[2] += 2 * e / qH
    r = np.linalg.norm(x[3] - x[0])
    s6 = (sig / r) ** 6
    Ue += 4 * eps * (s6 * s6 - s6)
    d[0] += 24 * eps / sig * (2 * s6 * s6 - s6)
    d[1] += 4 * (s6 * s6 - s6)
    assert U == pytest.approx(Ue, rel=1e-12)
    assert rel_err(dU, d) < 1e-12


def _dhvap_setup(n_steps_loss=0, **kw):
    o = _orc(w_rdf=0.0, w_dhvap=1.0, n_steps_loss=n_steps_loss, **kw)
    return o


def test_dhvap_loss_value(box216):
    """L = w (u_gas - U/M + RT - target)^2 at the end of a one-block trajectory
    (water/water.jl:738-754)."""
    o = _dhvap_setup()
    xN, vN, L = o.forward(box216.coords, box216.velocities, 13, 5)
    U = o.energy(xN)[0]
    dH = 3.69 - U / 216 + RT300
    assert L == pytest.approx((dH - 46.92) ** 2, rel=1e-12)


def test_dhvap_gradient_matches_trajectory_finite_differences(box216):
    """Central differences of the 10-step trajectory loss (one loss block, same noise) in all six
    parameters: the enthalpy loss depends on θ through the trajectory and directly through U."""
    import revsim

    o = _dhvap_setup()
    L, g = o.value_and_grad(box216.coords, box216.velocities, 13, 10)
    for k in range(6):
        th = revsim.DEFAULT_THETA.copy()
        hk = 1e-6 * th[k]
        th[k] += hk
        o.set_theta(th)
        lp = o.forward(box216.coords, box216.velocities, 13, 10)[2]
        th[k] -= 2 * hk
        o.set_theta(th)
        lm = o.forward(box216.coords, box216.velocities, 13, 10)[2]
        o.set_theta(revsim.DEFAULT_THETA)
        fd = (lp - lm) / (2 * hk)
        assert abs(g[k] - fd) < 1e-6 * abs(fd), (k, g[k], fd)


def test_dhvap_blocks_reset_the_adjoint(box216):
    """Loss blocks reset the accumulators (reverse_sim, water/water.jl:1525-1547): the gradient of
    a two-block trajectory is the mean of each block's own single-block gradient, the second one
    started at the block boundary with the same noise replayed."""
    o2 = _dhvap_setup(n_steps_loss=5)
    _, g = o2.value_and_grad(box216.coords, box216.velocities, 13, 10)
    o1 = _dhvap_setup()
    _, g1 = o1.value_and_grad(box216.coords, box216.velocities, 13, 5)
    x5, v5, _ = o1.forward(box216.coords, box216.velocities, 13, 5)
    xi = np.concatenate([o1.noise(13, 5 + k) for k in range(1, 6)])
    o1.set_external_noise(xi, 5)
    _, g2 = o1.value_and_grad(x5, v5, 13, 5)
    assert rel_err(g, 0.5 * (g1 + g2)) < 1e-11


def test_dhvap_reversible_gradient_equals_stored_adjoint(box216):
    import revsim

    o = _orc(w_dhvap=1.0)
    g0, _, _, _ = o.rdf(box216.coords)
    o.set_target_rdf(revsim.perturbed_target(g0))  # RDF + enthalpy terms together
    L1, g1 = o.value_and_grad(box216.coords, box216.velocities, 13, 30)
    L2, g2 = o.value_and_grad_stored(box216.coords, box216.velocities, 13, 30)
    assert L1 == pytest.approx(L2, rel=1e-10)
    assert rel_err(g1, g2) < 1e-9


def test_reweighting_estimator_from_block_states(box216):
    """dL/dθ = <dL_b/dθ> - β(<L_b dU_b/dθ> - <L_b><dU_b/dθ>) (water/water.jl:1201-1207), rebuilt
    here from the block-end states (a forward run of 2k steps starts like one of k steps with the
    same seed) with the oracle's single-state RDF and energy."""
    import revsim

    o = _orc(w_dhvap=0.5, n_steps_loss=6)
    g0, _, _, _ = o.rdf(box216.coords)
    o.set_target_rdf(revsim.perturbed_target(g0))
    L, g, xN, _ = o.reweight(box216.coords, box216.velocities, 13, 18)
    Ls, dLs, dUs = [], [], []
    for n in (6, 12, 18):
        x = o.forward(box216.coords, box216.velocities, 13, n)[0]
        lr = o.rdf(x)[2]
        U, dU, _ = o.energy(x)
        dH = 3.69 - U / 216 + RT300
        le = 0.5 * (dH - 46.92) ** 2
        Ls.append(lr + le)
        dLs.append(-2 * 0.5 * (dH - 46.92) / 216 * dU)
        dUs.append(dU)
    Ls, dLs, dUs = np.array(Ls), np.array(dLs), np.array(dUs)
    beta = 1 / RT300
    ref = dLs.mean(0) - beta * ((Ls[:, None] * dUs).mean(0) - Ls.mean() * dUs.mean(0))
    assert L == pytest.approx(Ls.mean(), rel=1e-12)
    assert rel_err(g, ref) < 1e-10
    assert np.array_equal(xN, o.forward(box216.coords, box216.velocities, 13, 18)[0])


def test_loss_terms_from_block_states(box216):
    """The per-term losses the reference's reverse_sim returns beside the total (l_rdf, l_ev,
    water/water.jl:1549-1557; l_diff and D, diffusion/diffusion.jl:897-911,1605-1616): rebuilt
    from the block-end states with the single-state RDF and energy, summing to the loss, and equal
    after the forward and after the reverse pass (FP64 reconstruction)."""
    import revsim

    w_rdf, w_dh, w_msd = 0.7, 0.5, 1e-3
    o = _orc(w_rdf=w_rdf, w_dhvap=w_dh, n_steps_loss=20, msd_window_steps=20, msd_every=5, w_msd=w_msd)
    o.set_target_rdf(revsim.perturbed_target(o.rdf(box216.coords)[0]))
    xN, vN, Lf = o.forward(box216.coords, box216.velocities, 13, 40)
    tf = o.loss_terms()
    op = _orc(w_rdf=0.0)  # intermediate states of the same trajectory (the loss does not steer it)
    lr, le = [], []
    for tb in (20, 40):
        x = op.forward(box216.coords, box216.velocities, 13, tb)[0]
        lr.append(w_rdf * o.rdf(x)[2])
        U = o.energy(x)[0]
        le.append(w_dh * (3.69 - U / 216 + RT300 - 46.92) ** 2)
    frames = np.concatenate([op.forward(box216.coords, box216.velocities, 13, t)[0].reshape(-1, 9)[:, :3].reshape(-1)
                             for t in (25, 30, 35, 40)])
    D, lm, _ = o.msd(frames)
    assert tf["l_rdf"] == pytest.approx(np.mean(lr), rel=1e-12)
    assert tf["l_ev"] == pytest.approx(np.mean(le), rel=1e-12)
    assert tf["l_msd"] == pytest.approx(lm, rel=1e-12)
    assert tf["D"] == pytest.approx(D * 1e-6, rel=1e-12)
    assert Lf == pytest.approx(tf["l_rdf"] + tf["l_ev"] + tf["l_msd"], rel=1e-12)
    o.forward(box216.coords, box216.velocities, 13, 40)
    Lr, _, _, _ = o.reverse(xN, vN, 13, 40)
    tr = o.loss_terms()
    for k in ("l_rdf", "l_ev", "l_msd", "D"):
        assert tr[k] == pytest.approx(tf[k], rel=1e-9), k
    assert Lr == pytest.approx(tr["l_rdf"] + tr["l_ev"] + tr["l_msd"], rel=1e-12)
