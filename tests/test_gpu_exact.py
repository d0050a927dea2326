"""FP64 path against the FP64 oracle, bit for bit: the Langevin noise (Philox bits, Box-Muller by
the library's own log and sincos), the forces (pair terms summed as exact 64-bit fixed-point
integers, bonded terms in one shared operation order), and therefore whole trajectories. With the
trajectory exact, the reverse pass's adjoint is a linear recurrence whose relative rounding does
not grow with chaos, so dL/dθ matches the oracle at full benchmark lengths, where two FP64 runs
that merely sum in different orders decorrelate (about 5e-2 after 1,000 steps of the 3,000-water
box, measured with the oracle at 4 against 8 threads before its sums were made exact)."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

FP64 = 0


def _pair(n, box, rc, **kw):
    import oracle as O
    import revsim

    sim = revsim.Simulator(revsim.make_config(n, box, rc, precision=FP64, **kw), device=0)
    orc = O.Oracle(revsim.make_config(n, box, rc, precision=FP64, **kw))
    for o in (sim, orc):
        o.set_theta(revsim.DEFAULT_THETA)
    return sim, orc


def test_noise_bit_exact(gpu_lib):
    sim, orc = _pair(216, 1.862, 0.8)
    for seed, step in ((13, 1), (13, 7), (99, 1000), (2 ** 63 + 5, 123456)):
        assert np.array_equal(sim.noise(seed, step), orc.noise(seed, step)), (seed, step)


@pytest.mark.parametrize("n,box,rc", [(216, 1.862, 0.8), (3000, 4.48, 1.0)])
def test_forces_bit_exact(gpu_lib, n, box, rc):
    """Forces and the adjoint pass's forces equal the oracle's bit for bit (perturbed lattice)."""
    import revsim

    sim, orc = _pair(n, box, rc)
    wb = revsim.water_box(n, box)
    rng = np.random.default_rng(11)
    x = np.ascontiguousarray(wb.coords + 0.01 * rng.standard_normal(wb.coords.size))
    F, Fo = sim.forces(x), orc.forces(x)
    assert np.array_equal(F, Fo), np.abs(F - Fo).max()
    c, w = rng.standard_normal(x.size), rng.standard_normal(x.size)
    Fa, h, vj = sim.adjoint_terms(x, c, w)
    Fao, ho, vjo = orc.adjoint_terms(x, c, w)
    assert np.array_equal(Fa, Fo)
    assert rel_err(h, ho) < 1e-13 and rel_err(vj, vjo) < 1e-12


@pytest.mark.parametrize("n,box,rc,steps", [(216, 1.862, 0.8, 200), (3000, 4.48, 1.0, 300)])
def test_trajectory_bit_exact(gpu_lib, n, box, rc, steps):
    import revsim

    sim, orc = _pair(n, box, rc, w_rdf=0.0)
    wb = revsim.water_box(n, box)
    x, v, _ = sim.forward(wb.coords, wb.velocities, 13, steps)
    xo, vo, _ = orc.forward(wb.coords, wb.velocities, 13, steps)
    assert np.array_equal(x, xo) and np.array_equal(v, vo), (np.abs(x - xo).max(), np.abs(v - vo).max())


def test_216_waters_full_length(gpu_lib, box216):
    """216 waters, 1.862 nm, r_c 0.8 nm, 200 + 200 steps, RDF loss: the loss equals the oracle's
    bit for bit and dL/dθ to 1e-12 (was 4e-11 with FP64 sums in different orders)."""
    import revsim

    sim, orc = _pair(216, 1.862, 0.8)
    tgt = revsim.perturbed_target(orc.rdf(box216.coords)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(box216.coords, box216.velocities, 13, 200)
    Lo, go = orc.value_and_grad(box216.coords, box216.velocities, 13, 200)
    assert L == Lo
    assert rel_err(g, go) < 1e-12, (g, go)


@pytest.mark.timeout(900)
def test_3000_waters_full_length(gpu_lib, box3000):
    """3,000 waters, 4.48 nm, r_c 1.0 nm, 1,000 + 1,000 steps, RDF loss, FP64 path: loss bit-exact,
    dL/dθ within 1e-10 of the oracle at the full length, and the reconstructed start state too."""
    import revsim

    sim, orc = _pair(3000, 4.48, 1.0)
    tgt = revsim.perturbed_target(orc.rdf(box3000.coords)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(box3000.coords, box3000.velocities, 13, 1000)
    Lo, go = orc.value_and_grad(box3000.coords, box3000.velocities, 13, 1000)
    assert L == Lo
    assert rel_err(g, go) < 1e-10, (rel_err(g, go), g, go)


@pytest.mark.timeout(900)
def test_98k_atoms_rdf_msd(gpu_lib):
    """32,768 waters (98,304 atoms), 9.95 nm, r_c 1.0 nm, RDF + MSD (last 50 steps) loss over
    100 + 100 steps, FP64 path: loss bit-exact, dL/dθ within 1e-10; the per-term losses equal."""
    import revsim

    wb = revsim.water_box(32768, 9.95)
    sim, orc = _pair(32768, 9.95, 1.0, msd_window_steps=50, msd_every=25, w_msd=1.0)
    tgt = revsim.perturbed_target(orc.rdf(wb.coords)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(wb.coords, wb.velocities, 13, 100)
    Lo, go = orc.value_and_grad(wb.coords, wb.velocities, 13, 100)
    assert L == Lo
    assert rel_err(g, go) < 1e-10, (rel_err(g, go), g, go)
    t, to = sim.loss_terms(), orc.loss_terms()
    assert t["l_rdf"] == to["l_rdf"] and t["l_msd"] == to["l_msd"] and t["D"] == to["D"]


@pytest.mark.timeout(600)
def test_lists_beyond_16_bit_indices(gpu_lib):
    """65,600 waters (196,800 atoms) in 12.54 nm: the sorted molecule indices no longer fit 16 bits,
    so the reverse pass's lists keep 32-bit entries, as the forward pass's always do. FP64 path
    over 10 + 10 steps (list builds in both passes) against the oracle: loss bit-exact, dL/dθ
    within 1e-10."""
    import revsim

    wb = revsim.water_box(65600, 12.54)
    sim, orc = _pair(65600, 12.54, 1.0)
    tgt = revsim.perturbed_target(orc.rdf(wb.coords)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(wb.coords, wb.velocities, 13, 10)
    Lo, go = orc.value_and_grad(wb.coords, wb.velocities, 13, 10)
    assert L == Lo
    assert rel_err(g, go) < 1e-10, (rel_err(g, go), g, go)


@pytest.mark.timeout(1200)
def test_98k_atoms_full_length(gpu_lib):
    """The headline accuracy target: 98,304 atoms, RDF + MSD loss (last 500 fs), 1,000 + 1,000
    steps, FP64 path against the FP64 oracle — the loss bit-exact and dL/dθ within 1e-10, at the
    horizon where two FP64 runs with differently ordered sums would already disagree at O(1e-2)."""
    import revsim

    wb = revsim.water_box(32768, 9.95)
    sim, orc = _pair(32768, 9.95, 1.0, msd_window_steps=500, msd_every=50, w_msd=1.0)
    tgt = revsim.perturbed_target(orc.rdf(wb.coords)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(wb.coords, wb.velocities, 13, 1000)
    Lo, go = orc.value_and_grad(wb.coords, wb.velocities, 13, 1000)
    assert L == Lo
    assert rel_err(g, go) < 1e-10, (rel_err(g, go), g, go)
