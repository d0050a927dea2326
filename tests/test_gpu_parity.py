"""GPU parity: librevsim.so (through its C ABI) against the FP64 CPU oracle on the same inputs.

Tolerances: FP64 pair arithmetic (precision 0) within 1e-10 relative; FP32 pair arithmetic
(precision 1) within 1e-5 relative for forces/adjoint terms; neighbour sets and RDF histograms
bit-exact in both modes.
"""
import numpy as np
import pytest

from conftest import per_atom_rel, rel_err

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


def _pair(prec, n=216, box=1.862, rc=0.8, **kw):
    import oracle as O
    import revsim

    cfg = revsim.make_config(n, box, rc, precision=prec, **kw)
    sim = revsim.Simulator(cfg, device=0)
    sim.set_theta(revsim.DEFAULT_THETA)
    orc = O.Oracle(revsim.make_config(n, box, rc, precision=prec, **kw))
    orc.set_theta(revsim.DEFAULT_THETA)
    return sim, orc


def test_noise_matches_oracle(gpu_lib):
    sim, orc = _pair(FP64)
    for step in (1, 7, 1000):
        a, b = sim.noise(13, step), orc.noise(13, step)
        assert np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_molecule_list_exact(gpu_lib, prec, jittered216, box3000):
    for x, args in ((jittered216, (216, 1.862, 0.8)), (box3000.coords, (3000, 4.48, 1.0))):
        sim, orc = _pair(prec, *args)
        got = sim.molecule_pairs(x)
        want = orc.molecule_pairs(x, sim.list_radius())
        assert got.shape == want.shape
        assert np.array_equal(got, want)


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-11), (FP32, 1e-5)])
def test_forces(gpu_lib, prec, tol, jittered216, box3000):
    for x, args in ((jittered216, (216, 1.862, 0.8)), (box3000.coords, (3000, 4.48, 1.0))):
        sim, orc = _pair(prec, *args)
        F, Fo = sim.forces(x), orc.forces(x)
        assert per_atom_rel(F, Fo) < tol, per_atom_rel(F, Fo)
        # translation invariance: the net force vanishes up to the pair arithmetic's rounding
        net_tol = (1e-12 if prec == FP64 else 1e-6) * np.abs(Fo).sum()
        assert abs(F.reshape(-1, 3).sum(0)).max() < net_tol


def test_noncubic_box_fp64(gpu_lib, box3000):
    """An orthorhombic box large enough for cells (FP64 path; the FP32 path needs a cubic box):
    the list build decides membership and distance shell by the FP64 rule for every candidate.
    Pair set, forces, adjoint terms and a 10-step gradient against the oracle."""
    import revsim

    box = np.array([4.3, 4.48, 4.7])
    x = (box3000.coords.reshape(-1, 3) * (box / 4.48)).reshape(-1)
    x = np.ascontiguousarray(np.mod(x.reshape(-1, 3), box).reshape(-1))
    sim, orc = _pair(FP64, 3000, list(box), 1.0)
    assert np.array_equal(sim.molecule_pairs(x), orc.molecule_pairs(x, sim.list_radius()))
    assert per_atom_rel(sim.forces(x), orc.forces(x)) < 1e-11
    rng = np.random.default_rng(6)
    c, w = rng.standard_normal(x.size), rng.standard_normal(x.size)
    F, h, vj = sim.adjoint_terms(x, c, w)
    Fo, ho, vjo = orc.adjoint_terms(x, c, w)
    assert per_atom_rel(h, ho) < 1e-10
    assert rel_err(vj, vjo) < 1e-10
    tgt = revsim.perturbed_target(orc.rdf(x)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(x, box3000.velocities, 13, 10)
    Lo, go = orc.value_and_grad(x, box3000.velocities, 13, 10)
    assert abs(L - Lo) <= 1e-10 * abs(Lo)
    assert rel_err(g, go) < 1e-10, (g, go)


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-10), (FP32, 1e-5)])
def test_adjoint_terms(gpu_lib, prec, tol, jittered216):
    sim, orc = _pair(prec)
    rng = np.random.default_rng(5)
    c = rng.standard_normal(jittered216.size)
    w = rng.standard_normal(jittered216.size)
    F, h, v = sim.adjoint_terms(jittered216, c, w)
    Fo, ho, vo = orc.adjoint_terms(jittered216, c, w)
    assert per_atom_rel(F, Fo) < tol
    assert per_atom_rel(h, ho) < tol, per_atom_rel(h, ho)
    # the bonded VJPs are FP64 in both modes; pair VJPs follow the pair precision
    for k in range(6):
        assert abs(v[k] - vo[k]) <= tol * 10 * (abs(vo[k]) + 1e-30), (k, v[k], vo[k])


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_rdf_histogram_bit_exact(gpu_lib, prec, jittered216):
    import revsim

    sim, orc = _pair(prec)
    g0, H0, _, _ = orc.rdf(jittered216)
    tgt = revsim.perturbed_target(g0)
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    g, H, L, d = sim.rdf(jittered216)
    go, Ho, Lo, do = orc.rdf(jittered216)
    assert np.array_equal(H, Ho)
    assert np.array_equal(g, go)
    assert L == Lo
    assert rel_err(d, do) < 1e-12


def test_msd(gpu_lib, box216):
    sim, orc = _pair(FP64, msd_window_steps=40, msd_every=5, w_msd=1e-3)
    rng = np.random.default_rng(3)
    base = box216.coords.reshape(-1, 9)[:, :3].reshape(-1)
    fr = np.concatenate([base + 0.02 * k * rng.standard_normal(base.size) for k in range(8)])
    D, L, d = sim.msd(fr)
    Do, Lo, do = orc.msd(fr)
    assert abs(D - Do) <= 1e-10 * abs(Do)
    assert abs(L - Lo) <= 1e-9 * abs(Lo)
    assert rel_err(d, do) < 1e-9


@pytest.mark.parametrize("n,box,rc,H", [(216, 1.862, 0.8, 20), (3000, 4.48, 1.0, 20)])
def test_fp32_gradient_against_fp64_path(gpu_lib, n, box, rc, H):
    """dL/dθ of the FP32 path against the FP64 path of the library (itself matched to the oracle at
    1e-10) on a short trajectory, where chaos has not yet amplified rounding: this checks the FP32
    pair arithmetic itself, which the FP32-mode oracle, sharing the FP32 membership rule, cannot."""
    import revsim

    wb = revsim.water_box(n, box)
    g = {}
    for prec in (FP64, FP32):
        sim, orc = _pair(prec, n=n, box=box, rc=rc)
        if prec == FP64:
            tgt = revsim.perturbed_target(orc.rdf(wb.coords)[0])
        sim.set_target_rdf(tgt)
        g[prec] = sim.value_and_grad(wb.coords, wb.velocities, 13, H)[1]
        sim.close()
    assert rel_err(g[FP32], g[FP64]) < 1e-5, (g[FP32], g[FP64])


def test_fp32_forces_3000_waters(gpu_lib, box3000):
    """FP32 forces and adjoint terms of the 3,000-water box (cell-list path) against the oracle."""
    sim, orc = _pair(FP32, n=3000, box=4.48, rc=1.0)
    rng = np.random.default_rng(2)
    x = box3000.coords + 0.01 * rng.standard_normal(box3000.coords.size)
    assert per_atom_rel(sim.forces(x), orc.forces(x)) < 1e-5
    c = rng.standard_normal(x.size)
    w = rng.standard_normal(x.size)
    F, h, v = sim.adjoint_terms(x, c, w)
    Fo, ho, vo = orc.adjoint_terms(x, c, w)
    assert per_atom_rel(h, ho) < 1e-5
    assert rel_err(v, vo) < 1e-5


def test_forward_trajectory_fp64(gpu_lib, box216):
    sim, orc = _pair(FP64)
    x, v, _ = sim.forward(box216.coords, box216.velocities, 13, 50)
    xo, vo, _ = orc.forward(box216.coords, box216.velocities, 13, 50)
    dx = x - xo
    dx -= 1.862 * np.round(dx / 1.862)
    assert np.max(np.abs(dx)) < 1e-10
    assert rel_err(v, vo) < 1e-10


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-10), (FP32, 1e-5)])
def test_value_and_grad(gpu_lib, prec, tol, box216):
    import revsim

    sim, orc = _pair(prec)
    g0, _, _, _ = orc.rdf(box216.coords)
    tgt = revsim.perturbed_target(g0)
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(box216.coords, box216.velocities, 13, 50)
    Lo, go = orc.value_and_grad(box216.coords, box216.velocities, 13, 50)
    assert abs(L - Lo) <= tol * abs(Lo)
    assert rel_err(g, go) < tol, (g, go)


@pytest.mark.parametrize("prec,chunks,tol", [(FP64, 8, 1e-10), (FP32, 8, 1e-5)])
def test_eight_lanes_per_molecule(gpu_lib, prec, chunks, tol, box216):
    """pair_chunks = 8 (the default is 16): the 8-lane pair kernels, whose forward kernel runs the
    integrator step on two waves per block (32 molecules; 216 waters leave the last block partly
    filled). The forward trajectory and the loss gradient against the oracle, to the 16-lane
    tolerances."""
    import revsim

    sim, orc = _pair(prec, pair_chunks=chunks)
    x, v, _ = sim.forward(box216.coords, box216.velocities, 13, 50)
    xo, vo, _ = orc.forward(box216.coords, box216.velocities, 13, 50)
    dx = x - xo
    dx -= 1.862 * np.round(dx / 1.862)
    assert np.max(np.abs(dx)) < (1e-10 if prec == FP64 else 1e-5)
    g0, _, _, _ = orc.rdf(box216.coords)
    tgt = revsim.perturbed_target(g0)
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(box216.coords, box216.velocities, 13, 50)
    Lo, go = orc.value_and_grad(box216.coords, box216.velocities, 13, 50)
    assert abs(L - Lo) <= tol * abs(Lo)
    assert rel_err(g, go) < tol, (g, go)


def test_reversibility_fp64(gpu_lib, box216):
    import revsim

    sim, orc = _pair(FP64)
    g0, _, _, _ = orc.rdf(box216.coords)
    sim.set_target_rdf(revsim.perturbed_target(g0))
    xN, vN, lf = sim.forward(box216.coords, box216.velocities, 13, 100)
    lr, g, x0, v0 = sim.reverse(xN, vN, 13, 100)
    d = x0 - np.mod(box216.coords, 1.862)
    d -= 1.862 * np.round(d / 1.862)
    assert np.max(np.abs(d)) < 1e-9  # the reference's own bound, water/water.jl:1613
    assert abs(lr - lf) < 1e-7  # water/water.jl:1614


def test_reversibility_fp64_98k_atoms(gpu_lib):
    """The reference's own reversibility bounds (water/water.jl:1613-1614) at the 98k-atom size,
    200 + 200 steps, FP64 path: x0 recovered within 1e-9 nm, |L_rev - L_fwd| < 1e-7."""
    import revsim

    wb = revsim.water_box(32768, 9.95)
    sim = revsim.Simulator(revsim.make_config(32768, 9.95, 1.0, precision=FP64), device=0)
    sim.set_theta(revsim.DEFAULT_THETA)
    sim.set_target_rdf(revsim.perturbed_target(sim.rdf(wb.coords)[0]))
    xN, vN, lf = sim.forward(wb.coords, wb.velocities, 13, 200)
    lr, g, x0, v0 = sim.reverse(xN, vN, 13, 200)
    d = x0 - np.mod(wb.coords, 9.95)
    d -= 9.95 * np.round(d / 9.95)
    assert np.max(np.abs(d)) < 1e-9
    assert abs(lr - lf) < 1e-7
    assert np.all(np.isfinite(g))


@pytest.mark.timeout(600)
def test_98k_atoms_fp32_rdf_msd(gpu_lib):
    """The FP32 path at 32,768 waters (98,304 atoms), 9.95 nm, r_c 1.0 nm, list every 10
    steps, RDF + MSD loss (MSD over the last 20 steps), 20 + 20 steps, against the oracle with the
    same FP32 pair-membership rule: loss and dL/dθ within 1e-5 (measured 3e-8 and 7e-8), the
    per-term losses within 1e-5, and no list rerun."""
    import revsim

    wb = revsim.water_box(32768, 9.95)
    kw = dict(msd_window_steps=20, msd_every=5, w_msd=1.0)
    sim, orc = _pair(FP32, n=32768, box=9.95, rc=1.0, **kw)
    tgt = revsim.perturbed_target(orc.rdf(wb.coords)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(wb.coords, wb.velocities, 13, 20)
    Lo, go = orc.value_and_grad(wb.coords, wb.velocities, 13, 20)
    print(f"98k FP32 20+20 RDF+MSD: loss rel {abs(L - Lo) / abs(Lo):.3e}, dL/dθ rel {rel_err(g, go):.3e}")
    assert abs(L - Lo) <= 1e-5 * abs(Lo)
    assert rel_err(g, go) < 1e-5, (g, go)
    t, to = sim.loss_terms(), orc.loss_terms()
    for k in ("l_rdf", "l_msd", "D"):
        assert abs(t[k] - to[k]) <= 1e-5 * abs(to[k]) + 1e-30, (k, t[k], to[k])
    assert sim.stats()["reruns_call"] == 0
