"""Power spectrum on the GPU (gng_pm_power_spectrum in include/gng_pm.h).

Mirrors crates/gadget-ng-analysis/src/power_spectrum.rs: :func:`power_spectrum`
takes the reference's arguments (positions, masses, box size, mesh) and returns
:class:`PkBin` values in increasing k.  The deposit, transforms and shell
sums run in libgng_pm.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import GNG_OK, HipExecutionError, Plan, _ptr, last_error, lib


@dataclass(frozen=True)
class PkBin:
    """k (2 pi / L times the shell's upper integer), mean power, modes (power_spectrum.rs:14-23)."""
    k: float
    pk: float
    n_modes: int


def _bins(k, pk, nm, nb) -> List[PkBin]:
    return [PkBin(float(k[i]), float(pk[i]), int(nm[i])) for i in range(nb)]


def power_spectrum_arrays(plan: Plan, positions, masses):
    """(k, pk, n_modes) arrays of the non-empty shells, from host f64 input."""
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    m = np.ascontiguousarray(masses, dtype=np.float64)
    cap = plan.grid_size // 2
    k = np.zeros(cap)
    pk = np.zeros(cap)
    nm = np.zeros(cap, dtype=np.uint64)
    nb = ctypes.c_int32(0)
    rc = lib().gng_pm_power_spectrum_f64(plan.handle, _ptr(pos), _ptr(m), pos.shape[0], cap, _ptr(k), _ptr(pk),
                                         _ptr(nm), ctypes.byref(nb))
    if rc != GNG_OK:
        raise HipExecutionError("gng_pm_power_spectrum_f64", rc, last_error())
    return k[:nb.value], pk[:nb.value], nm[:nb.value]


def power_spectrum_device(plan: Plan, x: int, y: int, z: int, m: int, n: int, stream: int = 0):
    """Same on device f32 SoA pointers (synchronises the stream)."""
    cap = plan.grid_size // 2
    k = np.zeros(cap)
    pk = np.zeros(cap)
    nm = np.zeros(cap, dtype=np.uint64)
    nb = ctypes.c_int32(0)
    rc = lib().gng_pm_power_spectrum(plan.handle, x, y, z, m, int(n), cap, _ptr(k), _ptr(pk), _ptr(nm),
                                     ctypes.byref(nb), stream or None)
    if rc != GNG_OK:
        raise HipExecutionError("gng_pm_power_spectrum", rc, last_error())
    return k[:nb.value], pk[:nb.value], nm[:nb.value]


def power_spectrum(positions, masses, box_size: float, mesh: int, device: int = -1,
                   plan: Optional[Plan] = None) -> List[PkBin]:
    """The reference's power_spectrum(positions, masses, box_size, mesh)."""
    own = plan is None
    if own:
        plan = Plan(mesh, box_size, device=device)
    try:
        k, pk, nm = power_spectrum_arrays(plan, positions, masses)
        return _bins(k, pk, nm, len(k))
    finally:
        if own:
            plan.close()
