"""ctypes front end of the FP64 CPU oracle (``pm_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
bench.py's ``cpu_baseline`` leg, always as the checker or as the timed CPU
baseline, never as part of the product path.

Every function mirrors one reference function (paths relative to the
gadget-ng tree):

* :func:`cic_assign`          crates/gadget-ng-pm/src/cic.rs:37-93 (and :355-401)
* :func:`cic_interpolate`     crates/gadget-ng-pm/src/cic.rs:100-170
* :func:`fft3d`               crates/gadget-ng-pm/src/fft_poisson.rs:733-813
* :func:`solve_forces`        crates/gadget-ng-pm/src/fft_poisson.rs:55-122, 306-476
* :func:`pm_accelerations`    crates/gadget-ng-pm/src/solver.rs:74-140 and the
                              long-range half of crates/gadget-ng-treepm/src/solver.rs:65-90
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

MODE_FAITHFUL = 0
MODE_LEAN = 1

DEPOSIT_SERIAL = 0
DEPOSIT_PRIVATE_GRIDS = 1
DEPOSIT_PLANES = 2

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        d = ctypes.c_double
        i32 = ctypes.c_int
        L.gngo_set_threads.argtypes = [i32]
        L.gngo_get_threads.restype = i32
        L.gngo_fft1d.argtypes = [P, i32, i64, i32]
        L.gngo_fft3d.argtypes = [P, i32, i32, i32]
        L.gngo_cic_assign.argtypes = [P, P, i64, d, i32, i32, P]
        L.gngo_cic_interpolate.argtypes = [P, P, P, P, i64, d, i32, P]
        L.gngo_solve_forces.argtypes = [P, d, i32, d, d, d, i32, i32, i32, P, P, P]
        L.gngo_pm_accelerations.argtypes = [P, P, i64, P, i64, d, i32, d, d, d, i32, i32, i32, i32, P, P]
        for f in (L.gngo_fft1d, L.gngo_fft3d, L.gngo_cic_assign, L.gngo_cic_interpolate,
                  L.gngo_solve_forces, L.gngo_pm_accelerations):
            f.restype = i32
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n: int) -> None:
    lib().gngo_set_threads(int(n))


def get_threads() -> int:
    return int(lib().gngo_get_threads())


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def fft1d(a: np.ndarray, inverse: bool = False) -> np.ndarray:
    """Unnormalised 1-D transform along the last axis (forward sign e^{-2 pi i})."""
    out = np.ascontiguousarray(a, dtype=np.complex128).copy()
    n = out.shape[-1]
    _check(lib().gngo_fft1d(_p(out), n, out.size // n, int(inverse)), "fft1d")
    return out


def fft3d(a: np.ndarray, inverse: bool = False, threads: int = 0) -> np.ndarray:
    """Full complex nm^3 transform, indices [iz, iy, ix]; unnormalised."""
    out = np.ascontiguousarray(a, dtype=np.complex128).copy()
    nm = out.shape[0]
    assert out.shape == (nm, nm, nm)
    _check(lib().gngo_fft3d(_p(out), nm, int(inverse), threads), "fft3d")
    return out


def cic_assign(pos: np.ndarray, mass: np.ndarray, box: float, nm: int,
               deposit_mode: int = DEPOSIT_SERIAL) -> np.ndarray:
    """Mass per cell, flat index iz*nm^2 + iy*nm + ix (returned as [nm,nm,nm])."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    assert mass.shape[0] == pos.shape[0]
    rho = np.empty((nm, nm, nm), dtype=np.float64)
    _check(lib().gngo_cic_assign(_p(pos), _p(mass), pos.shape[0], float(box), nm, deposit_mode, _p(rho)),
           "cic_assign")
    return rho


def cic_interpolate(fx, fy, fz, pos, box: float, nm: int) -> np.ndarray:
    fx, fy, fz = (np.ascontiguousarray(f, dtype=np.float64) for f in (fx, fy, fz))
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    acc = np.empty_like(pos)
    _check(lib().gngo_cic_interpolate(_p(fx), _p(fy), _p(fz), _p(pos), pos.shape[0], float(box), nm,
                                      _p(acc)), "cic_interpolate")
    return acc


def solve_forces(density, g: float, nm: int, box: float, r_split: float = 0.0,
                 plummer_eps: float = 0.0, deconvolve: bool = False, mode: int = MODE_LEAN,
                 fft_threads: int = 0):
    density = np.ascontiguousarray(density, dtype=np.float64)
    assert density.size == nm ** 3
    f = np.empty((3, nm, nm, nm), dtype=np.float64)
    _check(lib().gngo_solve_forces(_p(density), float(g), nm, float(box), float(r_split), float(plummer_eps),
                                   int(deconvolve), mode, fft_threads, _p(f[0]), _p(f[1]), _p(f[2])),
           "solve_forces")
    return f[0], f[1], f[2]


def pm_accelerations(pos, mass, box: float, nm: int, g: float = 1.0, idx=None, r_split: float = 0.0,
                     plummer_eps: float = 0.0, deconvolve: bool = False, mode: int = MODE_LEAN,
                     deposit_mode: int = DEPOSIT_PLANES, fft_threads: int = 0, want_density: bool = False):
    """Accelerations (K x 3 float64) for pos[idx] (all N when idx is None)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    n = pos.shape[0]
    if idx is None:
        k = n
        idx_p = None
    else:
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        k = idx.shape[0]
        idx_p = _p(idx)
    out = np.empty((k, 3), dtype=np.float64)
    dens = np.empty((nm, nm, nm), dtype=np.float64) if want_density else None
    _check(lib().gngo_pm_accelerations(_p(pos), _p(mass), n, idx_p, k, float(box), nm, float(g),
                                       float(r_split), float(plummer_eps), int(deconvolve), mode,
                                       deposit_mode, fft_threads, _p(out),
                                       _p(dens) if dens is not None else None), "pm_accelerations")
    return (out, dens) if want_density else out


def treepm_long_range(pos, mass, box: float, nm: int, g: float = 1.0, r_split: float = 0.0, idx=None,
                      mode: int = MODE_LEAN, **kw):
    """Long-range half of TreePmSolver: r_s = r_split, or 2.5 cells when r_split <= 0
    (crates/gadget-ng-treepm/src/solver.rs:56-62)."""
    r_s = r_split if r_split > 0.0 else 2.5 * box / nm
    return pm_accelerations(pos, mass, box, nm, g=g, idx=idx, r_split=r_s, mode=mode, **kw)
