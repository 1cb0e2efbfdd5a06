"""Device-resident KDK leapfrog over the PM solver (include/gng_integrate.h).

Mirrors, on top of libgng_pm.so:

* :class:`CosmologyParams` -- crates/gadget-ng-core/src/cosmology.rs:37-221
  (``advance_a``, ``drift_kick_factors``; the arithmetic runs in the library);
* :class:`CosmoFactors`    -- crates/gadget-ng-integrators/src/leapfrog.rs:40-64;
* :func:`gravity_coupling_qksl` -- cosmology.rs:354-356;
* :class:`Leapfrog`        -- ``leapfrog_kdk_step`` / ``leapfrog_cosmo_kdk_step``
  (leapfrog.rs:5-95) over a particle state that stays in HBM, with the
  periodic wrap of the cosmological engine
  (crates/gadget-ng-cli/src/engine/stepping/mod.rs:1572-1576).

Everything runs through the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from . import GNG_OK, HipExecutionError, Plan, _ptr, last_error, lib


def _check(rc: int, what: str) -> None:
    if rc != GNG_OK:
        raise HipExecutionError(what, rc, last_error())


class _Cosmo(ctypes.Structure):
    _fields_ = [("omega_m", ctypes.c_double), ("omega_lambda", ctypes.c_double), ("h0", ctypes.c_double),
                ("w0", ctypes.c_double), ("wa", ctypes.c_double), ("omega_nu", ctypes.c_double)]


@dataclass(frozen=True)
class CosmologyParams:
    """Flat background (cosmology.rs:37-44): h0 in internal time units."""
    omega_m: float
    omega_lambda: float
    h0: float
    w0: float = -1.0
    wa: float = 0.0
    omega_nu: float = 0.0

    def _c(self) -> _Cosmo:
        return _Cosmo(self.omega_m, self.omega_lambda, self.h0, self.w0, self.wa, self.omega_nu)

    def advance_a(self, a: float, dt: float) -> float:
        c = self._c()
        return float(lib().gng_cosmo_advance_a(ctypes.byref(c), float(a), float(dt)))

    def hubble(self, a: float) -> float:
        """H(a) (cosmology.rs:509-520)."""
        c = self._c()
        return float(lib().gng_cosmo_hubble(ctypes.byref(c), float(a)))

    def growth_rate(self, a: float) -> float:
        """f(a) = Omega_m(a)^0.55 (cosmology.rs:598-610)."""
        c = self._c()
        return float(lib().gng_cosmo_growth_rate(ctypes.byref(c), float(a)))

    def growth_factor(self, a: float) -> float:
        """D(a), not normalised (cosmology.rs:715-724)."""
        c = self._c()
        return float(lib().gng_cosmo_growth_factor(ctypes.byref(c), float(a)))

    def drift_kick_factors(self, a0: float, dt: float) -> Tuple[float, float, float]:
        c = self._c()
        out = np.zeros(3, dtype=np.float64)
        _check(lib().gng_cosmo_drift_kick_factors(ctypes.byref(c), float(a0), float(dt), _ptr(out)),
               "gng_cosmo_drift_kick_factors")
        return float(out[0]), float(out[1]), float(out[2])


@dataclass(frozen=True)
class CosmoFactors:
    """drift = int dt/a^2, kick_half / kick_half2 = int dt/a over each half step."""
    drift: float
    kick_half: float
    kick_half2: float

    @staticmethod
    def flat(dt: float) -> "CosmoFactors":
        return CosmoFactors(dt, dt * 0.5, dt * 0.5)


def gravity_coupling_qksl(g: float, a: float) -> float:
    return g * a * a * a


class Leapfrog:
    """KDK leapfrog of n particles on the device of `plan` (a PM plan, or the
    long-range plan of a TreePM split with :meth:`set_short_range`)."""

    def __init__(self, plan: Plan, n: int):
        self.plan = plan
        self.n = int(n)
        h = ctypes.c_void_p()
        _check(lib().gng_kdk_create(plan.handle, self.n, ctypes.byref(h)), "gng_kdk_create")
        self._h = h
        self._sr = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().gng_kdk_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_short_range(self, sr, r_split: float, r_cut: Optional[float] = None, eps2: float = 0.0,
                        kernel: int = 0) -> None:
        """Add a learned.ShortRange to every force evaluation (None removes it)."""
        r_cut = 5.0 * r_split if r_cut is None else r_cut
        _check(lib().gng_kdk_set_short_range(self._h, sr._h if sr is not None else None, float(r_split),
                                             float(r_cut), float(eps2), int(kernel)), "gng_kdk_set_short_range")
        self._sr = sr

    def set_state(self, pos, vel, mass) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64).reshape(self.n)
        _check(lib().gng_kdk_set_state(self._h, _ptr(pos), _ptr(vel), _ptr(mass)), "gng_kdk_set_state")

    def set_state_device(self, d_pos: int, d_vel: int, d_mass: Optional[int] = None,
                         uniform_mass: Optional[float] = None, stream: int = 0) -> None:
        """The state from device buffers (n x 3 f64 positions and velocities;
        masses from d_mass, or uniform_mass for all), without a host copy.
        Enqueued on `stream`: order it after steps running on other streams."""
        if not d_mass and uniform_mass is None:
            raise ValueError("give d_mass or uniform_mass")
        uniform_mass = 0.0 if uniform_mass is None else uniform_mass
        _check(lib().gng_kdk_set_state_device(self._h, ctypes.c_void_p(d_pos), ctypes.c_void_p(d_vel),
                                              ctypes.c_void_p(d_mass) if d_mass else None, float(uniform_mass),
                                              ctypes.c_void_p(stream) if stream else None),
               "gng_kdk_set_state_device")

    def state(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(positions, velocities, last accelerations), each (n, 3) float64."""
        pos = np.zeros((self.n, 3), dtype=np.float64)
        vel = np.zeros((self.n, 3), dtype=np.float64)
        acc = np.zeros((self.n, 3), dtype=np.float64)
        _check(lib().gng_kdk_get_state(self._h, _ptr(pos), _ptr(vel), _ptr(acc)), "gng_kdk_get_state")
        return pos, vel, acc

    def step(self, cf: CosmoFactors, g: float, wrap: bool = False, stream: int = 0) -> None:
        """One KDK step (leapfrog_cosmo_kdk_step; CosmoFactors.flat(dt) gives leapfrog_kdk_step)."""
        _check(lib().gng_kdk_step(self._h, float(cf.drift), float(cf.kick_half), float(cf.kick_half2), float(g),
                                  int(bool(wrap)), stream or None), "gng_kdk_step")

    def run(self, nsteps: int, dt: float, g: float, cosmology: Optional[CosmologyParams] = None,
            a: float = 1.0, wrap: bool = True, stream: int = 0) -> float:
        """nsteps steps; with a cosmology the coupling is g a^3 at each step's
        start and a advances along advance_a.  Returns the final a."""
        av = ctypes.c_double(float(a))
        c = cosmology._c() if cosmology is not None else None
        _check(lib().gng_kdk_run(self._h, ctypes.byref(c) if c is not None else None, ctypes.byref(av), float(dt),
                                 float(g), int(nsteps), int(bool(wrap)), stream or None), "gng_kdk_run")
        return float(av.value)
