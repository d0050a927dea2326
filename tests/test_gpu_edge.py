"""Edge cases of the GPU path against the FP64 oracle: a molecule that outruns the list's skin and
an O-H bond beyond the assumed extent (the call is rerun with a wider skin or extent and the
results equal the oracle's, which needs no list), a packing dense enough to overflow the list's
capacity (rebuilt larger), a molecule count that fills no whole wave or block on the cell-list
path, invalid step counts, and a close contact whose FP64 pair force leaves the fixed-point sums'
range."""
import numpy as np
import pytest

from conftest import per_atom_rel, rel_err

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


def _pair(prec, n, box, rc, **kw):
    import oracle as O
    import revsim

    sim = revsim.Simulator(revsim.make_config(n, box, rc, precision=prec, **kw), device=0)
    sim.set_theta(revsim.DEFAULT_THETA)
    orc = O.Oracle(revsim.make_config(n, box, rc, precision=prec, **kw))
    orc.set_theta(revsim.DEFAULT_THETA)
    return sim, orc


def _mic(d, box):
    return d - box * np.round(d / box)


def test_stale_list_rerun(gpu_lib, box3000):
    """Molecule 0 at +4 nm/ps moves ~0.04 nm within a 10-step list interval, beyond skin/2 =
    0.03 nm: the staleness check reruns the call with a wider skin (the list radius grows), and
    the 20-step trajectory equals the oracle's."""
    sim, orc = _pair(FP64, 3000, 4.48, 1.0, w_rdf=0.0)
    r0 = sim.list_radius()
    v = box3000.velocities.copy().reshape(-1, 3)
    v[0:3, 0] += 4.0
    v = np.ascontiguousarray(v.reshape(-1))
    x, vv, _ = sim.forward(box3000.coords, v, 21, 20)
    xo, vo, _ = orc.forward(box3000.coords, v, 21, 20)
    assert sim.list_radius() > r0 + 0.01
    assert np.abs(_mic(x - xo, 4.48)).max() < 1e-10
    assert rel_err(vv, vo) < 1e-10


@pytest.mark.parametrize("prec,tol", [(FP64, 1e-11), (FP32, 1e-5)])
def test_extent_rerun(gpu_lib, box3000, prec, tol):
    """Molecule 0's first O-H bond stretched to 0.12 nm, beyond the assumed extent (0.1121 nm):
    the evaluation is rerun with a 0.01 nm wider extent and the forces equal the oracle's."""
    sim, orc = _pair(prec, 3000, 4.48, 1.0)
    r0 = sim.list_radius()
    x = box3000.coords.copy().reshape(-1, 3)
    oh = _mic(x[1] - x[0], 4.48)
    x[1] = np.mod(x[0] + oh * (0.12 / np.linalg.norm(oh)), 4.48)
    x = np.ascontiguousarray(x.reshape(-1))
    F, Fo = sim.forces(x), orc.forces(x)
    assert sim.list_radius() == pytest.approx(r0 + 0.02, abs=1e-12)
    assert per_atom_rel(F, Fo) < tol


@pytest.mark.parametrize("prec", [FP64, FP32])
def test_list_capacity_rebuild(gpu_lib, box3000, prec):
    """The 3,000 waters packed into one octant of the box (each molecule translated so that its O
    sits at half its coordinates: 8x the density, ~1,300 molecules within 1.06 nm of a central
    one). The first build overflows the list's capacity (49 slots per lane and shell, 784 entries
    per molecule and shell), the call rebuilds with twice the capacity, and the pair set equals
    the oracle's."""
    sim, orc = _pair(prec, 3000, 4.48, 1.0)
    x = box3000.coords.reshape(-1, 3, 3)
    o = np.mod(x[:, 0, :], 4.48)
    xs = np.ascontiguousarray(np.mod(x + (0.5 * o - x[:, 0, :])[:, None, :], 4.48).reshape(-1))
    got = sim.molecule_pairs(xs)
    assert sim.stats()["reruns_call"] >= 1
    want = orc.molecule_pairs(xs, sim.list_radius())
    assert got.shape == want.shape
    assert np.array_equal(got, want)


@pytest.mark.parametrize("prec,tol,gtol", [(FP64, 1e-11, 1e-10), (FP32, 1e-5, 1e-5)])
def test_ragged_molecule_count(gpu_lib, prec, tol, gtol):
    """2,999 waters (not a multiple of the 4 molecules per pair wave, the 16 per pair block or the
    21 per integrator wave) on the cell-list path: pair set, forces and a 10-step gradient."""
    import revsim

    wb = revsim.water_box(2999, 4.48)
    sim, orc = _pair(prec, 2999, 4.48, 1.0)
    x = wb.coords
    assert np.array_equal(sim.molecule_pairs(x), orc.molecule_pairs(x, sim.list_radius()))
    assert per_atom_rel(sim.forces(x), orc.forces(x)) < tol
    tgt = revsim.perturbed_target(orc.rdf(x)[0])
    sim.set_target_rdf(tgt)
    orc.set_target_rdf(tgt)
    L, g = sim.value_and_grad(x, wb.velocities, 13, 10)
    Lo, go = orc.value_and_grad(x, wb.velocities, 13, 10)
    assert abs(L - Lo) <= gtol * abs(Lo)
    assert rel_err(g, go) < gtol, (g, go)


def test_invalid_step_counts(gpu_lib, box216):
    import revsim

    sim, _ = _pair(FP64, 216, 1.862, 0.8)
    for n in (0, -5):
        with pytest.raises(revsim.RevSimError) as e:
            sim.forward(box216.coords, box216.velocities, 1, n)
        assert e.value.code == -1
        with pytest.raises(revsim.RevSimError):
            sim.value_and_grad(box216.coords, box216.velocities, 1, n)
    # the context stays usable
    x, v, _ = sim.forward(box216.coords, box216.velocities, 1, 3)
    assert np.all(np.isfinite(x)) and np.all(np.isfinite(v))


def test_fixed_point_range_is_an_error(gpu_lib, box216):
    """FP64 pair forces are exact 64-bit fixed-point sums of terms below 2^17 kJ/mol/nm. Two waters
    whose O atoms sit 0.12 nm apart (O-O Lennard-Jones force ~3e7 kJ/mol/nm) must make the call
    fail with RS_E_OVERFLOW rather than return a wrapped sum; the context stays usable and then
    matches the oracle."""
    import revsim

    sim, orc = _pair(FP64, 216, 1.862, 0.8)
    x = box216.coords.copy().reshape(-1, 3, 3)
    d = _mic(x[1, 0] - x[0, 0], 1.862)
    x[1] += (x[0, 0] + 0.12 * d / np.linalg.norm(d)) - x[1, 0]  # the whole molecule moves
    with pytest.raises(revsim.RevSimError) as e:
        sim.forces(x.reshape(-1))
    assert e.value.code == -3 and "fixed-point" in str(e.value)
    # the adjoint kernel (flags once after its loops) as well
    rng = np.random.default_rng(3)
    with pytest.raises(revsim.RevSimError) as e:
        sim.adjoint_terms(x.reshape(-1), rng.standard_normal(x.size), rng.standard_normal(x.size))
    assert e.value.code == -3 and "fixed-point" in str(e.value)
    F = sim.forces(box216.coords)
    Fo = orc.forces(box216.coords)
    assert rel_err(F, Fo) < 1e-12
