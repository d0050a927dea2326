"""The Langevin integrator on the GPU against closed forms (no interactions: Coulomb constant, ε,
bond and angle constants zero; molecules move rigidly, one velocity and one noise vector for
their three atoms, so no bond is needed to keep them whole).
- Zero noise: v_n = e^n v_0 and x_n = x_0 + (dt/2) v_0 (1 + e)(1 - e^n)/(1 - e), both precisions.
- Noise of the molecular scale σ² = k_B T / M from rest: an Ornstein-Uhlenbeck process, whose
  kinetic temperature is T (1 - e^{2n}) (statistical, 3,000 molecules)."""
import numpy as np
import pytest

from test_oracle import _free_config, free_particle_closed_form

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [0, 1])
def test_free_particle_closed_form_gpu(gpu_lib, box3000, prec):
    import revsim

    cfg, th = _free_config(3000, 4.48, 1.0, precision=prec)
    sim = revsim.Simulator(cfg, device=0)
    sim.set_theta(th)
    rng = np.random.default_rng(4)
    v0 = np.repeat(rng.normal(0.0, 0.4, (3000, 1, 3)), 3, axis=1).reshape(-1)
    n = 60
    sim.set_external_noise(np.zeros(n * v0.size), n)
    x, v, _ = sim.forward(box3000.coords, v0, 1, n)
    xc, vc = free_particle_closed_form(box3000.coords, v0, n, cfg.dt, cfg.gamma, 4.48)
    d = x - xc
    d -= 4.48 * np.round(d / 4.48)
    assert np.abs(d).max() < 1e-12
    assert np.abs(v - vc).max() < 1e-13


def test_ornstein_uhlenbeck_temperature(gpu_lib, box3000):
    import revsim
    from revsim.system import KB, MASS_H, MASS_O

    cfg, th = _free_config(3000, 4.48, 1.0, precision=1)
    sim = revsim.Simulator(cfg, device=0)
    sim.set_theta(th)
    M, T = MASS_O + 2 * MASS_H, cfg.temperature
    e = np.exp(-cfg.gamma * cfg.dt)
    rng = np.random.default_rng(8)
    v = np.zeros(9 * 3000)
    x = box3000.coords
    for n in (300, 700):  # consecutive legs: 300 then 1,000 steps from rest
        xi = np.repeat(rng.normal(0.0, np.sqrt(KB * T / M), (n, 3000, 1, 3)), 3, axis=2).reshape(-1)
        sim.set_external_noise(np.ascontiguousarray(xi), n)
        x, v, _ = sim.forward(x, v, 1, n)
        vm = v.reshape(3000, 3, 3)
        assert np.array_equal(vm[:, 0], vm[:, 1]) and np.array_equal(vm[:, 0], vm[:, 2])
        total = 300 if n == 300 else 1000
        want = T * (1.0 - e ** (2 * total))
        got = M * np.mean(vm[:, 0] ** 2) / KB
        # 9,000 velocity components: relative standard error sqrt(2/9000) = 1.5 %
        assert abs(got - want) < 0.06 * want, (total, got, want)
