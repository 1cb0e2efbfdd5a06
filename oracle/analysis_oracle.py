"""FP64 restatement of the reference's power-spectrum estimator.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's ``cpu_baseline``
leg as the checker or as the timed CPU baseline, never by the product path.

Mirrors crates/gadget-ng-analysis/src/power_spectrum.rs (paths relative to the
gadget-ng tree):

* :func:`cic_grid`        ``cic_assign`` (:148-170): mass to 8 cells, flat index
                          jx*n^2 + jy*n + jz, every index wrapped;
* :func:`pk_from_mass`    steps 1b-3 of ``power_spectrum`` (:54-127): delta =
                          rho/(rho_bar V_cell) - 1, the full complex 3-D DFT,
                          division by (wx wy wz)^2, shells floor(|n| - 1/2);
* :func:`power_spectrum`  the whole function.

The DFT is numpy's (pocketfft); the reference's is rustfft (Cargo.lock pins
rustfft 6.4.1) -- both compute the unnormalised forward DFT, and only
|delta_k|^2 enters.
"""
from __future__ import annotations

import math

import numpy as np


def freq(i: np.ndarray, n: int) -> np.ndarray:
    """Signed frequency of index i: i for i <= n/2, else i - n (:130-135)."""
    i = np.asarray(i)
    return np.where(i <= n // 2, i, i - n)


def sinc(x: np.ndarray) -> np.ndarray:
    """sin(pi x)/(pi x), 1 for |x| < 1e-12 (:138-145)."""
    x = np.asarray(x, dtype=np.float64)
    px = math.pi * x
    safe = np.where(np.abs(x) < 1e-12, 1.0, px)
    return np.where(np.abs(x) < 1e-12, 1.0, np.sin(px) / safe)


def cic_grid(pos: np.ndarray, mass: np.ndarray, box: float, n: int) -> np.ndarray:
    """Mass per cell, shape (n, n, n) indexed [jx, jy, jz]."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    mass = np.asarray(mass, dtype=np.float64)
    cell = box / n
    f = pos / cell
    i0 = np.floor(f)
    t = f - i0
    i0 = i0.astype(np.int64)
    # the 8 (cell, mass) pairs of every particle, particle-major and corner
    # order (dx, dy, dz), so the per-cell sums run in the reference's order
    idx = np.empty((len(mass), 8), dtype=np.int64)
    val = np.empty((len(mass), 8))
    c = 0
    for dx in (0, 1):
        wx = t[:, 0] if dx else 1.0 - t[:, 0]
        jx = np.mod(i0[:, 0] + dx, n)
        for dy in (0, 1):
            wy = t[:, 1] if dy else 1.0 - t[:, 1]
            jy = np.mod(i0[:, 1] + dy, n)
            for dz in (0, 1):
                wz = t[:, 2] if dz else 1.0 - t[:, 2]
                jz = np.mod(i0[:, 2] + dz, n)
                idx[:, c] = (jx * n + jy) * n + jz
                val[:, c] = mass * wx * wy * wz
                c += 1
    grid = np.bincount(idx.reshape(-1), weights=val.reshape(-1), minlength=n * n * n)
    return grid.reshape(n, n, n)


def shells(n: int):
    """Per mode of the full n^3 spectrum: shell index (-1 when excluded) and the
    deconvolution weight (wx wy wz)^2, arrays of shape (n, n, n)."""
    f = freq(np.arange(n), n).astype(np.float64)
    kx, ky, kz = np.meshgrid(f, f, f, indexing="ij")
    k2 = kx * kx + ky * ky + kz * kz
    bf = np.sqrt(k2) - 0.5
    nb = n // 2
    ok = (k2 != 0.0) & (bf >= 0.0) & (bf < nb)
    shell = np.where(ok, np.floor(np.where(ok, bf, 0.0)), -1).astype(np.int64)
    w = sinc(f / n)
    wx, wy, wz = np.meshgrid(w, w, w, indexing="ij")
    return shell, (wx * wy * wz) ** 2


def pk_from_mass(grid: np.ndarray, box: float, total_mass=None):
    """(k, pk, n_modes) of the non-empty shells from a mass-per-cell grid;
    total_mass defaults to the grid's sum (the reference sums the particles)."""
    n = grid.shape[0]
    total = float(grid.sum()) if total_mass is None else float(total_mass)
    mean_rho = total / (box * box * box)
    cell = box / n
    delta = grid / (mean_rho * (cell * cell * cell)) - 1.0
    dk = np.fft.fftn(delta.astype(np.complex128))
    shell, w2 = shells(n)
    norm = (box ** 3 / n ** 3) ** 2
    d2 = (dk.real ** 2 + dk.imag ** 2) / w2 * norm
    nb = n // 2
    sel = shell >= 0
    sums = np.bincount(shell[sel], weights=d2[sel], minlength=nb)
    counts = np.bincount(shell[sel], minlength=nb)
    k_fund = 2.0 * math.pi / box
    keep = counts > 0
    b = np.nonzero(keep)[0]
    return (b + 1.0) * k_fund, sums[keep] / counts[keep], counts[keep].astype(np.uint64)


def power_spectrum(pos, mass, box: float, mesh: int):
    """The reference's power_spectrum: (k, pk, n_modes) arrays."""
    mass = np.asarray(mass, dtype=np.float64)
    return pk_from_mass(cic_grid(pos, mass, box, mesh), box, total_mass=float(mass.sum()))
