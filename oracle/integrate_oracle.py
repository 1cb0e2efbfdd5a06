"""FP64 restatement of the reference's KDK leapfrog and cosmological factors.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's ``cpu_baseline``
leg as the checker or as the timed CPU baseline, never by the product path.

Every function mirrors one reference function (paths relative to the
gadget-ng tree):

* :func:`da_dt`, :func:`advance_a`      crates/gadget-ng-core/src/cosmology.rs:141-167
* :func:`drift_kick_factors`            crates/gadget-ng-core/src/cosmology.rs:188-221
* :func:`gravity_coupling_qksl`         crates/gadget-ng-core/src/cosmology.rs:354-356
* :func:`rem_euclid`, :func:`wrap_position`  crates/gadget-ng-core/src/cosmology.rs:759-778
* :func:`kdk_step`                      crates/gadget-ng-integrators/src/leapfrog.rs:5-23, 76-95
* :func:`run`                           the periodic PM loop of
                                        crates/gadget-ng-cli/src/engine/stepping/mod.rs:1076-1576

The arithmetic keeps the reference's operation order; numpy rounds every
product and sum separately, as Rust does.
"""
from __future__ import annotations

import math
from typing import Callable, Optional, Sequence, Tuple

import numpy as np


def da_dt(c: Sequence[float], a: float) -> float:
    """c = (omega_m, omega_lambda, h0, w0, wa, omega_nu)."""
    omega_m, omega_lambda, h0, w0, wa, omega_nu = c
    if abs(w0 + 1.0) < 1e-10 and abs(wa) < 1e-10:
        omega_de = omega_lambda
    else:
        exp_factor = math.exp(3.0 * wa * (a - 1.0))
        power = -3.0 * (1.0 + w0 + wa)
        omega_de = omega_lambda * math.pow(a, power) * exp_factor
    h_sq = (omega_m + omega_nu) / (a * a * a) + omega_de
    return a * h0 * math.sqrt(max(h_sq, 0.0))


def advance_a(c, a: float, dt: float) -> float:
    k1 = da_dt(c, a)
    k2 = da_dt(c, a + 0.5 * dt * k1)
    k3 = da_dt(c, a + 0.5 * dt * k2)
    k4 = da_dt(c, a + dt * k3)
    return a + dt / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4)


def drift_kick_factors(c, a0: float, dt: float) -> Tuple[float, float, float]:
    n_sub = 16
    sub_dt = dt / float(n_sub)
    a = a0
    drift = kick_half = kick_half2 = 0.0
    for sub in range(n_sub):
        a_left = a
        a_mid = advance_a(c, a, sub_dt * 0.5)
        a_right = advance_a(c, a, sub_dt)
        drift += sub_dt / 6.0 * (1.0 / (a_left * a_left) + 4.0 / (a_mid * a_mid) + 1.0 / (a_right * a_right))
        k_sub = sub_dt / 6.0 * (1.0 / a_left + 4.0 / a_mid + 1.0 / a_right)
        if sub < n_sub // 2:
            kick_half += k_sub
        else:
            kick_half2 += k_sub
        a = a_right
    return drift, kick_half, kick_half2


def gravity_coupling_qksl(g: float, a: float) -> float:
    return g * a * a * a


def rem_euclid(x, l: float):
    r = np.fmod(x, l)
    return np.where(r < 0.0, r + abs(l), r)


def wrap_position(pos: np.ndarray, box: float) -> np.ndarray:
    return rem_euclid(np.asarray(pos, dtype=np.float64), box)


def kdk_step(pos: np.ndarray, vel: np.ndarray, factors: Tuple[float, float, float],
             force: Callable[[np.ndarray], np.ndarray]):
    """One KDK step; factors = (drift, kick_half, kick_half2).  Returns the new
    (pos, vel, acc) with acc the second force evaluation."""
    drift, kick_half, kick_half2 = factors
    acc = force(pos)
    vel = vel + acc * kick_half
    pos = pos + vel * drift
    acc = force(pos)
    vel = vel + acc * kick_half2
    return pos, vel, acc


def run(pos, vel, nsteps: int, dt: float, g: float, force_g: Callable[[np.ndarray, float], np.ndarray],
        cosmology: Optional[Sequence[float]] = None, a: float = 1.0, box: Optional[float] = None):
    """nsteps steps; force_g(pos, g) is the force with coupling g.  With a
    cosmology: g a^3 at each step's start, factors from drift_kick_factors(a, dt),
    then a = advance_a(a, dt); flat factors otherwise.  box: wrap after each step."""
    acc = np.zeros_like(np.asarray(pos, dtype=np.float64))
    for _ in range(nsteps):
        if cosmology is not None:
            gk = gravity_coupling_qksl(g, a)
            f = drift_kick_factors(cosmology, a, dt)
            a = advance_a(cosmology, a, dt)
        else:
            gk, f = g, (dt, dt * 0.5, dt * 0.5)
        pos, vel, acc = kdk_step(pos, vel, f, lambda p: force_g(p, gk))
        if box is not None:
            pos = wrap_position(pos, box)
    return pos, vel, acc, a
