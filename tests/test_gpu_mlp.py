"""GPU parity of the radial-basis MLP pair term against the oracle's exact per-pair FP64 network.

FP64 path: the library's FP64 table + cubic Hermite interpolation differs from the exact network
by the interpolation error only (< 1e-8 relative in f, < 1e-6 in f'), so forces agree to 1e-9
relative to the typical force and gradients to 1e-6. FP32 path: the table comes from BF16 MFMA
GEMMs with FP32 accumulation whose activation operand is split into three BF16 terms (the weights
are BF16 values), so the table carries FP32 rounding (~1e-6) and the FP32 gradients and dL/dW meet
1e-4."""
import numpy as np
import pytest

from conftest import per_atom_rel, rel_err

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


def _pair(prec, n=216, box=1.862, rc=0.8, **kw):
    import oracle as O
    import revsim
    from revsim.system import mlp_init

    w = mlp_init()
    sim = revsim.Simulator(revsim.make_config(n, box, rc, precision=prec, use_mlp=1, **kw), device=0)
    sim.set_theta(revsim.DEFAULT_THETA)
    sim.set_mlp(w)
    orc = O.Oracle(revsim.make_config(n, box, rc, precision=prec, use_mlp=1, **kw))
    orc.set_theta(revsim.DEFAULT_THETA)
    orc.set_mlp(w)
    return sim, orc


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-12), (FP32, 1e-5)])
def test_mlp_table(gpu_lib, prec, tol):
    sim, orc = _pair(prec)
    r, f, df = sim.mlp_table()
    _, u1, u2 = orc.mlp_eval(r)
    assert np.max(np.abs(f + u1)) <= tol * np.max(np.abs(u1))
    assert np.max(np.abs(df + u2)) <= tol * np.max(np.abs(u2))


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-9), (FP32, 1e-5)])
def test_mlp_forces(gpu_lib, prec, tol, jittered216):
    sim, orc = _pair(prec)
    F, Fo = sim.forces(jittered216), orc.forces(jittered216)
    assert per_atom_rel(F, Fo) < tol, per_atom_rel(F, Fo)


@pytest.mark.parametrize("prec,tol,tol_w", [(FP64, 1e-7, 1e-6), (FP32, 1e-5, 1e-4)])
def test_mlp_adjoint_terms_and_weight_vjp(gpu_lib, prec, tol, tol_w, jittered216):
    sim, orc = _pair(prec)
    rng = np.random.default_rng(5)
    c = rng.standard_normal(jittered216.size)
    w = rng.standard_normal(jittered216.size)
    F, h, v = sim.adjoint_terms(jittered216, c, w)
    dW = sim.mlp_grad()
    Fo, ho, vo, dWo = orc.adjoint_terms_mlp(jittered216, c, w)
    assert per_atom_rel(F, Fo) < tol
    assert per_atom_rel(h, ho) < tol, per_atom_rel(h, ho)
    for k in range(6):
        assert abs(v[k] - vo[k]) <= tol * 10 * (abs(vo[k]) + 1e-30), (k, v[k], vo[k])
    assert rel_err(dW, dWo) < tol_w, rel_err(dW, dWo)


@pytest.mark.parametrize("prec,tol,tol_w", [(FP64, 1e-6, 1e-5), (FP32, 1e-4, 1e-4)])
def test_mlp_value_and_grad(gpu_lib, prec, tol, tol_w, box216):
    import revsim

    sim, orc = _pair(prec)
    g0, _, _, _ = orc.rdf(box216.coords)
    tgt = revsim.perturbed_target(g0)
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(box216.coords, box216.velocities, 13, 30)
    dW = sim.mlp_grad()
    Lo, go = orc.value_and_grad(box216.coords, box216.velocities, 13, 30)
    dWo = orc.mlp_grad()
    assert abs(L - Lo) <= tol * abs(Lo)
    assert rel_err(g, go) < tol, (g, go)
    assert rel_err(dW, dWo) < tol_w, rel_err(dW, dWo)


def test_mlp_gradient_is_deterministic(gpu_lib, box216):
    """Fixed-point table-gradient sums: two identical runs give identical bits."""
    import revsim

    sim, orc = _pair(FP32)
    g0, _, _, _ = orc.rdf(box216.coords)
    sim.set_target_rdf(revsim.perturbed_target(g0))
    L1, g1 = sim.value_and_grad(box216.coords, box216.velocities, 13, 20)
    d1 = sim.mlp_grad()
    L2, g2 = sim.value_and_grad(box216.coords, box216.velocities, 13, 20)
    d2 = sim.mlp_grad()
    assert L1 == L2 and np.array_equal(g1, g2) and np.array_equal(d1, d2)


_ORACLE_CACHE = {}


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("prec,tol", [(FP64, 1e-8), (FP32, 5e-5)])
def test_98k_atoms_mlp_first_100_steps(gpu_lib, prec, tol):
    """Learned correction at scale: the 98,304-atom box with the MLP term and the RDF + MSD loss,
    against the FP64 oracle (exact per-pair network) on the first 100 steps. The FP64 path's table
    is within ~1e-11 of the network, which chaos grows ~100x over 100 steps; the FP32 path's
    rounding grows the same way (1.3e-5 at 100 steps without the MLP, 7e-8 at 20), hence the
    bounds. The MLP changes dL/dθ against the plain force field (by ~3e-3 with Xavier weights over
    these 100 steps) by more than ten times either bound."""
    import oracle as O
    import revsim
    from revsim.system import mlp_init

    wb = revsim.water_box(32768, 9.95)
    kw = dict(msd_window_steps=50, msd_every=25, w_msd=1.0)
    sim, orc = _pair(prec, n=32768, box=9.95, rc=1.0, **kw)
    tgt = revsim.perturbed_target(orc.rdf(wb.coords)[0])
    sim.set_target_rdf(tgt)
    if "c4" not in _ORACLE_CACHE:  # one oracle run (~2 minutes) serves both precisions
        orc_f = O.Oracle(revsim.make_config(32768, 9.95, 1.0, precision=FP64, use_mlp=1, **kw))
        orc_f.set_theta(revsim.DEFAULT_THETA)
        orc_f.set_mlp(mlp_init())
        orc_f.set_target_rdf(tgt)
        _ORACLE_CACHE["c4"] = orc_f.value_and_grad(wb.coords, wb.velocities, 13, 100)
    Lo, go = _ORACLE_CACHE["c4"]
    L, g = sim.value_and_grad(wb.coords, wb.velocities, 13, 100)
    assert abs(L - Lo) <= tol * abs(Lo)
    assert rel_err(g, go) < tol, (rel_err(g, go), g, go)
    plain = revsim.Simulator(revsim.make_config(32768, 9.95, 1.0, precision=prec, **kw), device=0)
    plain.set_theta(revsim.DEFAULT_THETA)
    plain.set_target_rdf(tgt)
    _, gp = plain.value_and_grad(wb.coords, wb.velocities, 13, 100)
    assert rel_err(gp, go) > 10 * tol
