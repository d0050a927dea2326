"""Simulator: a thin Python owner of one librevsim context (one GPU, one stream).

Every method calls the HIP library through its C ABI; there is no CPU code path behind it. A
missing library raises LibraryMissing, a failing call raises RevSimError with the library's
message.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._abi import RsConfig, dptr, fptr, lptr


class RevSimError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_abi.ERRORS.get(code, code)}: {msg}")
        self.code = code


class Simulator:
    def __init__(self, cfg: RsConfig, device: int = 0, lib=None):
        self.lib = lib or _abi.load_library()
        self.cfg = cfg
        h = ctypes.c_void_p()
        rc = self.lib.rs_create(ctypes.byref(cfg), int(device), ctypes.byref(h))
        if rc:
            raise RevSimError(rc, self.lib.rs_last_error(None).decode())
        self.h = h
        self.M = cfg.n_molecules
        self.N3 = 9 * self.M

    def close(self):
        if getattr(self, "h", None):
            self.lib.rs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc:
            raise RevSimError(rc, self.lib.rs_last_error(self.h).decode())

    def _vec(self, a, n=None):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        if n is not None and a.size != n:
            raise ValueError(f"expected {n} values, got {a.size}")
        return a

    # -- parameters -------------------------------------------------------------------------
    def set_theta(self, theta):
        self._chk(self.lib.rs_set_theta(self.h, dptr(self._vec(theta, 6))))

    def set_reverse_options(self, n_steps_trunc: int, clip_norm: float):
        """reverse_sim's own arguments (truncation, HVP clip threshold); the forward records stay."""
        self._chk(self.lib.rs_set_reverse_options(self.h, int(n_steps_trunc), float(clip_norm)))
        self.cfg.n_steps_trunc = int(n_steps_trunc)
        self.cfg.clip_norm = float(clip_norm)

    def set_target_rdf(self, g):
        g = self._vec(g, self.cfg.rdf_bins)
        self._chk(self.lib.rs_set_target_rdf(self.h, dptr(g), g.size))

    def set_external_noise(self, xi, n_steps: int):
        xi = None if xi is None else self._vec(xi, n_steps * self.N3)
        self._chk(self.lib.rs_set_external_noise(self.h, dptr(xi), int(n_steps)))

    # -- radial-basis MLP on O-O pairs -------------------------------------------------------
    def set_mlp(self, w):
        """Packed weights (W1, b1, W2, b2, w3, b3; 5313 values), rounded to bfloat16 on entry."""
        w = np.ascontiguousarray(w, dtype=np.float32).reshape(-1)
        self._chk(self.lib.rs_set_mlp(self.h, fptr(w), w.size))

    def mlp_grad(self):
        d = np.empty(_abi.MLP_NW)
        self._chk(self.lib.rs_mlp_grad(self.h, dptr(d), d.size))
        return d

    def mlp_table(self):
        n = ctypes.c_int32()
        rlo = ctypes.c_double()
        self._chk(self.lib.rs_mlp_table(self.h, None, None, 0, ctypes.byref(n), ctypes.byref(rlo)))
        f, df = np.empty(n.value), np.empty(n.value)
        self._chk(self.lib.rs_mlp_table(self.h, dptr(f), dptr(df), n.value, ctypes.byref(n), ctypes.byref(rlo)))
        r = rlo.value + np.arange(n.value) * (self.cfg.cutoff - rlo.value) / (n.value - 1)
        return r, f, df

    # -- trajectories -----------------------------------------------------------------------
    def forward(self, x0, v0, seed: int, n_steps: int):
        x0, v0 = self._vec(x0, self.N3), self._vec(v0, self.N3)
        x, v, loss = np.empty(self.N3), np.empty(self.N3), np.zeros(1)
        self._chk(self.lib.rs_forward(self.h, dptr(x0), dptr(v0), seed, n_steps, dptr(x), dptr(v),
                                      dptr(loss)))
        return x, v, float(loss[0])

    def reverse(self, xN, vN, seed: int, n_steps: int):
        xN, vN = self._vec(xN, self.N3), self._vec(vN, self.N3)
        loss, g = np.zeros(1), np.zeros(6)
        x0, v0 = np.empty(self.N3), np.empty(self.N3)
        self._chk(self.lib.rs_reverse(self.h, dptr(xN), dptr(vN), seed, n_steps, dptr(loss), dptr(g),
                                      dptr(x0), dptr(v0)))
        return float(loss[0]), g, x0, v0

    def value_and_grad(self, x0, v0, seed: int, n_steps: int, mlp_grad_size: int = 0):
        x0, v0 = self._vec(x0, self.N3), self._vec(v0, self.N3)
        loss, g = np.zeros(1), np.zeros(6)
        gm = np.zeros(mlp_grad_size, dtype=np.float32) if mlp_grad_size else None
        self._chk(self.lib.rs_value_and_grad(self.h, dptr(x0), dptr(v0), seed, n_steps, dptr(loss),
                                             dptr(g), fptr(gm)))
        return (float(loss[0]), g) if gm is None else (float(loss[0]), g, gm)

    def reweight(self, x0, v0, seed: int, n_steps: int):
        """Ensemble-reweighting estimate of dL/dθ over one forward trajectory (reweighting_sim,
        water/water.jl:1107-1210): returns (mean block loss, dL/dθ, x_N, v_N)."""
        x0, v0 = self._vec(x0, self.N3), self._vec(v0, self.N3)
        loss, g = np.zeros(1), np.zeros(6)
        x, v = np.empty(self.N3), np.empty(self.N3)
        self._chk(self.lib.rs_reweight(self.h, dptr(x0), dptr(v0), seed, n_steps, dptr(loss), dptr(g),
                                       dptr(x), dptr(v)))
        return float(loss[0]), g, x, v

    # -- single evaluations -----------------------------------------------------------------
    def energy(self, x):
        """Potential energy U(x), dU/dθ (6) and F = -dU/dx, FP64 (potential_energy,
        water/water.jl:743,1183)."""
        x = self._vec(x, self.N3)
        U, dU, F = np.zeros(1), np.zeros(6), np.empty(self.N3)
        self._chk(self.lib.rs_energy(self.h, dptr(x), dptr(U), dptr(dU), dptr(F)))
        return float(U[0]), dU, F

    def forces(self, x):
        x = self._vec(x, self.N3)
        F = np.empty(self.N3)
        self._chk(self.lib.rs_forces(self.h, dptr(x), dptr(F)))
        return F

    def adjoint_terms(self, x, c, w):
        x, c, w = self._vec(x, self.N3), self._vec(c, self.N3), self._vec(w, self.N3)
        F, h, v = np.empty(self.N3), np.empty(self.N3), np.zeros(6)
        self._chk(self.lib.rs_adjoint_terms(self.h, dptr(x), dptr(c), dptr(w), dptr(F), dptr(h),
                                            dptr(v)))
        return F, h, v

    def noise(self, seed: int, step: int):
        xi = np.empty(self.N3)
        self._chk(self.lib.rs_noise(self.h, seed, step, dptr(xi)))
        return xi

    def rdf(self, x):
        x = self._vec(x, self.N3)
        nb = self.cfg.rdf_bins
        g, H, loss, d = np.empty(nb), np.empty(nb, dtype=np.int64), np.zeros(1), np.empty(self.N3)
        self._chk(self.lib.rs_rdf(self.h, dptr(x), dptr(g), lptr(H), dptr(loss), dptr(d)))
        return g, H, float(loss[0]), d

    def msd(self, frames):
        frames = self._vec(frames)
        nf = frames.size // (3 * self.M)
        D, loss, d = np.zeros(1), np.zeros(1), np.empty(frames.size)
        self._chk(self.lib.rs_msd(self.h, dptr(frames), nf, dptr(D), dptr(loss), dptr(d)))
        return float(D[0]), float(loss[0]), d

    def molecule_pairs(self, x, capacity: int | None = None):
        x = self._vec(x, self.N3)
        cap = capacity or min(self.M * (self.M - 1) // 2, 8_000_000)
        pairs = np.empty(2 * cap, dtype=np.int32)
        n = ctypes.c_int64()
        self._chk(self.lib.rs_molecule_pairs(self.h, dptr(x), pairs.ctypes.data_as(_abi._ip), cap,
                                             ctypes.byref(n)))
        return pairs[: 2 * n.value].reshape(-1, 2)

    def get_box(self) -> np.ndarray:
        """Current box edges (nm): where the barostat left it after a forward trajectory."""
        b = np.zeros(3)
        self._chk(self.lib.rs_get_box(self.h, dptr(b)))
        return b

    def set_box(self, box) -> None:
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)))
        self._chk(self.lib.rs_set_box(self.h, dptr(b)))

    def list_radius(self) -> float:
        return float(self.lib.rs_list_radius(self.h))

    def loss_terms(self) -> dict:
        """Loss terms of the last trajectory call: l_rdf, l_ev (enthalpy of vaporization), l_msd
        (averaged over loss blocks, weights applied; they sum to the loss) and D (m^2/s)."""
        t = np.zeros(4)
        self._chk(self.lib.rs_loss_terms(self.h, dptr(t)))
        return {"l_rdf": float(t[0]), "l_ev": float(t[1]), "l_msd": float(t[2]), "D": float(t[3])}

    def stats(self) -> dict:
        """List-driven reruns and list builds of the last call that ran device work, and reruns
        since the context was created."""
        st = np.zeros(3, dtype=np.int64)
        self._chk(self.lib.rs_stats(self.h, lptr(st), st.size))
        return {"reruns_call": int(st[0]), "list_builds_call": int(st[1]), "reruns_total": int(st[2])}

    def copy_calibration(self, nbytes: int, reps: int = 5):
        """Launch a plain copy kernel of known traffic (nbytes read + nbytes written per launch)."""
        self._chk(self.lib.rs_copy_calibration(self.h, int(nbytes), int(reps)))

    def enable_timing(self, on: bool = True):
        self._chk(self.lib.rs_enable_timing(self.h, int(on)))

    def timings(self) -> dict:
        """Per-phase device time (ms) and launch counts since the last call, from HIP events."""
        ms = np.zeros(len(_abi.PHASES))
        cnt = np.zeros(len(_abi.PHASES), dtype=np.int64)
        self._chk(self.lib.rs_timing_counts(self.h, lptr(cnt), cnt.size))
        self._chk(self.lib.rs_timings(self.h, dptr(ms), ms.size))
        return {k: {"ms": float(m), "launches": int(n)} for k, m, n in zip(_abi.PHASES, ms, cnt)}
