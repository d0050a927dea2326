"""GPU parity of the potential-energy kernel, the enthalpy-of-vaporization block loss and the
ensemble-reweighting estimator (water/water.jl:738-754,1107-1210) against the FP64 oracle. The
energy is FP64 in both precisions; trajectories in the FP32 path differ from the oracle's by the
pair arithmetic's rounding, hence the looser FP32 bounds."""
import numpy as np
import pytest

from conftest import per_atom_rel, rel_err

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


def _pair(prec, n=216, box=1.862, rc=0.8, **kw):
    import oracle as O
    import revsim

    sim = revsim.Simulator(revsim.make_config(n, box, rc, precision=prec, **kw), device=0)
    sim.set_theta(revsim.DEFAULT_THETA)
    orc = O.Oracle(revsim.make_config(n, box, rc, precision=prec, **kw))
    orc.set_theta(revsim.DEFAULT_THETA)
    return sim, orc


@pytest.mark.parametrize("prec", [FP64, FP32])
@pytest.mark.parametrize("n,box,rc", [(216, 1.862, 0.8), (3000, 4.48, 1.0)])
def test_energy_matches_oracle(gpu_lib, prec, n, box, rc):
    import revsim

    wb = revsim.water_box(n, box)
    rng = np.random.default_rng(7)
    x = np.ascontiguousarray(wb.coords + 0.01 * rng.standard_normal(wb.coords.size))
    sim, orc = _pair(prec, n=n, box=box, rc=rc)
    U, dU, F = sim.energy(x)
    Uo, dUo, Fo = orc.energy(x)
    assert U == pytest.approx(Uo, rel=1e-11)
    assert rel_err(dU, dUo) < 1e-10
    assert per_atom_rel(F, Fo) < 1e-10
    # the same forces as the force kernel's FP64 path
    if prec == FP64:
        assert per_atom_rel(sim.forces(x), F) < 1e-10
    # deterministic: fixed-order reductions
    U2, dU2, F2 = sim.energy(x)
    assert U2 == U and np.array_equal(dU2, dU) and np.array_equal(F2, F)


def test_energy_needs_the_classical_force_field(gpu_lib, box216):
    import revsim

    sim = revsim.Simulator(revsim.make_config(216, 1.862, 0.8, use_mlp=1), device=0)
    sim.set_mlp(revsim.system.mlp_init())
    with pytest.raises(revsim.RevSimError) as e:
        sim.energy(box216.coords)
    assert e.value.code == -1


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-9), (FP32, 1e-5)])
def test_dhvap_trajectory_gradient(gpu_lib, box216, prec, tol):
    """RDF + enthalpy-of-vaporization loss over two 10-step blocks: the enthalpy term seeds
    dL/dx = c' F and adds its direct -c' dU/dθ at each block end."""
    sim, orc = _pair(prec, w_dhvap=1.0, n_steps_loss=10)
    import revsim

    tgt = revsim.perturbed_target(orc.rdf(box216.coords)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    xN, vN, Lf = sim.forward(box216.coords, box216.velocities, 13, 20)
    xo, vo, Lfo = orc.forward(box216.coords, box216.velocities, 13, 20)
    assert abs(Lf - Lfo) <= tol * abs(Lfo)
    L, g = sim.value_and_grad(box216.coords, box216.velocities, 13, 20)
    Lo, go = orc.value_and_grad(box216.coords, box216.velocities, 13, 20)
    assert abs(L - Lo) <= tol * abs(Lo)
    assert rel_err(g, go) < tol, (g, go)


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-9), (FP32, 1e-5)])
def test_reweight_matches_oracle(gpu_lib, box216, prec, tol):
    """Three 6-step blocks, RDF + enthalpy losses."""
    import revsim

    sim, orc = _pair(prec, w_dhvap=0.5, n_steps_loss=6)
    tgt = revsim.perturbed_target(orc.rdf(box216.coords)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g, xN, _ = sim.reweight(box216.coords, box216.velocities, 13, 18)
    Lo, go, xo, _ = orc.reweight(box216.coords, box216.velocities, 13, 18)
    assert abs(L - Lo) <= tol * abs(Lo)
    assert rel_err(g, go) < tol, (g, go)
    d = xN - xo
    d -= 1.862 * np.round(d / 1.862)
    assert np.abs(d).max() < (1e-10 if prec == FP64 else 1e-5)


def test_reweight_rejects_msd_loss(gpu_lib, box216):
    import revsim

    sim = revsim.Simulator(revsim.make_config(216, 1.862, 0.8, msd_window_steps=10, msd_every=5,
                                              w_msd=1.0), device=0)
    with pytest.raises(revsim.RevSimError):
        sim.reweight(box216.coords, box216.velocities, 13, 10)


def test_barostat_matches_oracle(gpu_lib, box216):
    """FP64 forward trajectory with a Monte Carlo barostat move every 10 steps (water/water.jl:
    811-865): the same accepted volumes as the oracle (same Philox uniforms, FP64 energies) and
    the same final state; the FP32 path ends in the same box within the pair arithmetic's noise."""
    import revsim

    boxes = {}
    for prec in (FP64, FP32):
        sim, orc = _pair(prec, barostat_every=10, barostat_pressure=1.0, w_rdf=0.0)
        assert np.array_equal(sim.get_box(), [1.862] * 3)
        x, v, _ = sim.forward(box216.coords, box216.velocities, 13, 60)
        xo, vo, _ = orc.forward(box216.coords, box216.velocities, 13, 60)
        boxes[prec] = sim.get_box()
        if prec == FP64:
            assert np.allclose(boxes[prec], orc.get_box(), rtol=1e-13, atol=0)
            d = (x - xo).reshape(-1, 3)
            d -= boxes[prec] * np.round(d / boxes[prec])
            assert np.abs(d).max() < 1e-9
            assert rel_err(v, vo) < 1e-9
        # a trajectory that moved the box has no reverse pass
        with pytest.raises(revsim.RevSimError) as e:
            sim.reverse(x, v, 13, 60)
        assert e.value.code == -6
    assert not np.allclose(boxes[FP64], 1.862)
    assert np.allclose(boxes[FP32], boxes[FP64], rtol=1e-3)


def test_set_box(gpu_lib, jittered216):
    """rs_set_box: the forces at the new box equal the oracle's at that box."""
    import oracle as O
    import revsim

    sim, _ = _pair(FP64)
    sim.set_box([1.9] * 3)
    assert np.array_equal(sim.get_box(), [1.9] * 3)
    x = np.mod(jittered216 * (1.9 / 1.862), 1.9)
    orc = O.Oracle(revsim.make_config(216, 1.9, 0.8, precision=FP64))
    orc.set_theta(revsim.DEFAULT_THETA)
    assert per_atom_rel(sim.forces(x), orc.forces(x)) < 1e-10
    with pytest.raises(revsim.RevSimError):
        sim.set_box([1.5] * 3)  # below twice the cutoff


def test_interface_barostat_then_reweighting(gpu_lib, box216):
    """The reference's epoch shape (water/water.jl:1646-1669): an equilibration forward_sim with the
    barostat, then a trajectory at the box it reached (here reweighting_sim with the :both loss)."""
    import revsim
    from revsim import interface as I

    ctx = I.Context(216, 1.862, cutoff=0.8, precision=FP64)
    p = np.array([0.315, 0.636, -0.834, 462750.4, 0.09572, 836.8, 1.82421813418])
    sys_eq, _, _ = I.forward_sim(box216.coords, box216.velocities, None, 5, 1.0, 60, 60, p, 0.001,
                                 use_barostat=True, calc_loss=False, context=ctx)
    assert not np.allclose(sys_eq.boundary, 1.862)
    tgt = revsim.perturbed_target(ctx.simulator().rdf(sys_eq.coords.reshape(-1))[0])
    sys_r, L, g, _, _ = I.reweighting_sim(sys_eq.coords, sys_eq.velocities, sys_eq.boundary, 7, 1.0, 20,
                                          5, p, 0.001, loss_type="both", context=ctx, target_rdf=tgt)
    assert np.array_equal(sys_r.boundary, sys_eq.boundary)
    assert np.isfinite(L) and np.all(np.isfinite(g)) and g[6] == 0.0
    # the same trajectory in the oracle at that box
    import oracle as O

    orc = O.Oracle(revsim.make_config(216, list(sys_eq.boundary), 0.8, precision=FP64, n_steps_loss=5,
                                      w_dhvap=1.0))
    orc.set_theta(I.params_to_theta(p))
    orc.set_target_rdf(tgt)
    Lo, go, _, _ = orc.reweight(sys_eq.coords.reshape(-1), sys_eq.velocities.reshape(-1), 7, 20)
    assert L == pytest.approx(Lo, rel=1e-9)
    assert rel_err(g, I.theta_grad_to_params(go)) < 1e-8
