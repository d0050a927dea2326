"""GPU tests of the boundary's host-facing behaviour: the reference's forward_sim / reverse_sim pair
through the Python mirror (water/water.jl:1662-1669) against Simulator.value_and_grad, the per-term
losses against the oracle's, a parameter change that moves the list's cell grid, and the rerun
counters."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1
P0 = np.array([0.315, 0.636, -0.834, 462750.4, 0.09572, 836.8, 1.82421813418])


@pytest.mark.parametrize("loss_type", ["rdf", "rdf_msd"])
def test_forward_then_reverse_sim_equals_value_and_grad(gpu_lib, box216, loss_type):
    """forward_sim followed by reverse_sim (with truncation and a clip threshold passed to the
    reverse call only) gives value_and_grad's loss and gradient on the same configuration, and the
    loss terms sum to the loss."""
    import revsim
    from revsim import interface as I

    ctx = I.Context(216, 1.862, cutoff=0.8, precision=FP64)
    tgt = revsim.perturbed_target(ctx.simulator().rdf(box216.coords)[0])
    n, nl, nt, clip = 200, 100, 50, 1e4
    sys_f, Lf, _ = I.forward_sim(box216.coords, box216.velocities, None, 13, 1.0, n, nl, P0, 0.001,
                                 loss_type=loss_type, context=ctx, target_rdf=tgt)
    out = I.reverse_sim(sys_f.coords, sys_f.velocities, None, 13, 1.0, n, nl, nt, P0, 0.001, 1, sys_f.loggers,
                        clip, loss_type=loss_type)
    L, g = out[1], out[2]
    kw = I._loss_cfg(loss_type, n, nl, n_steps_trunc=nt, clip=clip)
    sim = revsim.Simulator(revsim.make_config(216, 1.862, 0.8, precision=FP64, **kw), device=0)
    sim.set_theta(I.params_to_theta(P0))
    sim.set_target_rdf(tgt)
    Lv, gv = sim.value_and_grad(box216.coords, box216.velocities, 13, n)
    assert L == pytest.approx(Lv, rel=1e-12)
    assert np.array_equal(g, I.theta_grad_to_params(gv))
    # the reconstructed start state
    d = sys_f.coords.reshape(-1)  # final state of the forward pass, not the start
    assert out[0].coords.shape == (648, 3) and not np.allclose(out[0].coords.reshape(-1), d)
    if loss_type == "rdf_msd":
        D, l_msd, l_rdf, l_ev = out[3:]
        assert D > 0 and l_msd > 0 and l_ev == 0.0
    else:
        l_rdf, l_ev = out[3:]
        l_msd = 0.0
        assert l_ev == 0.0
    assert l_rdf > 0 and L == pytest.approx(l_rdf + l_ev + l_msd, rel=1e-12)
    # reverse_sim refuses parameters other than the forward trajectory's
    with pytest.raises(ValueError):
        I.reverse_sim(sys_f.coords, sys_f.velocities, None, 13, 1.0, n, nl, nt, P0 * 1.01, 0.001, 1,
                      sys_f.loggers, clip, loss_type=loss_type)


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-10), (FP32, 1e-5)])
def test_loss_terms_match_oracle(gpu_lib, box216, prec, tol):
    """RDF, enthalpy-of-vaporization and MSD terms of a two-block trajectory (the extra return
    values of reverse_sim, water/water.jl:1549-1557; diffusion/diffusion.jl:1616) against the
    oracle's, after the forward and after the full gradient call."""
    import oracle as O
    import revsim

    kw = dict(n_steps_loss=20, w_dhvap=0.5, msd_window_steps=20, msd_every=5, w_msd=1e-3)
    sim = revsim.Simulator(revsim.make_config(216, 1.862, 0.8, precision=prec, **kw), device=0)
    orc = O.Oracle(revsim.make_config(216, 1.862, 0.8, precision=prec, **kw))
    for o in (sim, orc):
        o.set_theta(revsim.DEFAULT_THETA)
    tgt = revsim.perturbed_target(orc.rdf(box216.coords)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    for call in ("forward", "value_and_grad"):
        getattr(sim, call)(box216.coords, box216.velocities, 13, 40)
        getattr(orc, call)(box216.coords, box216.velocities, 13, 40)
        t, to = sim.loss_terms(), orc.loss_terms()
        for k in ("l_rdf", "l_ev", "l_msd", "D"):
            assert t[k] == pytest.approx(to[k], rel=tol), (call, k, t[k], to[k])


def test_set_theta_that_refines_the_cell_grid(gpu_lib, box3000):
    """A smaller O-H rest length shortens the list radius, so the 3,000-water box gets a finer cell
    grid (216 -> 343 cells) than the context was created with: the per-cell arrays must follow.
    The gradient at r0 = 0.093 nm against the oracle's."""
    import oracle as O
    import revsim

    th = revsim.DEFAULT_THETA.copy()
    th[4] = 0.093
    sim = revsim.Simulator(revsim.make_config(3000, 4.48, 1.0, precision=FP64), device=0)
    sim.set_theta(revsim.DEFAULT_THETA)
    r_default = sim.list_radius()
    sim.set_theta(th)
    assert sim.list_radius() < r_default
    orc = O.Oracle(revsim.make_config(3000, 4.48, 1.0, precision=FP64))
    orc.set_theta(th)
    tgt = revsim.perturbed_target(orc.rdf(box3000.coords)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(box3000.coords, box3000.velocities, 13, 20)
    Lo, go = orc.value_and_grad(box3000.coords, box3000.velocities, 13, 20)
    assert L == pytest.approx(Lo, rel=1e-10)
    assert rel_err(g, go) < 1e-10, (g, go)


def test_rerun_counters(gpu_lib, box3000):
    """A molecule that outruns the skin makes the call rerun with a wider list; rs_stats reports
    the rerun for that call and none for the next one (whose list is already wide enough)."""
    import revsim

    sim = revsim.Simulator(revsim.make_config(3000, 4.48, 1.0, precision=FP64, w_rdf=0.0), device=0)
    sim.set_theta(revsim.DEFAULT_THETA)
    v = box3000.velocities.copy().reshape(-1, 3)
    v[0:3, 0] += 4.0
    v = np.ascontiguousarray(v.reshape(-1))
    sim.forward(box3000.coords, v, 21, 20)
    st = sim.stats()
    assert st["reruns_call"] >= 1 and st["reruns_total"] == st["reruns_call"]
    assert st["list_builds_call"] >= 2 * 3  # each attempt builds at steps 0, 10, 20
    sim.forward(box3000.coords, box3000.velocities, 21, 20)
    st2 = sim.stats()
    assert st2["reruns_call"] == 0 and st2["reruns_total"] == st["reruns_total"]
    assert st2["list_builds_call"] == 3
