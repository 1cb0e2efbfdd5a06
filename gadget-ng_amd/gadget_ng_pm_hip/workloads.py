"""Seeded synthetic particle sets for the PM solver's workloads.

Every generator is deterministic for a seed: random numbers come from
numpy's counter-based Philox generator keyed by the seed, and Gaussian fields
are synthesised in FP64 with scipy's FFT.  None of the reference's IC code or
RNG is reproduced; the shapes follow the workloads the PM path is measured on:

* :func:`uniform_box`         N uniform positions in [0, L)^3, m = 1/N
                              (32^3 particles, L = 100, 64^3 mesh: the FP64
                              CPU parity case).
* :func:`zeldovich_lattice`   a node-aligned lattice displaced by the
                              Zel'dovich field of a Gaussian density with
                              P(k) ~ k^n up to the mesh Nyquist (256^3
                              particles on a 512^3 mesh, L = 256: the headline
                              case).
* :func:`ap08_curve`          the input of the reference's GPU-vs-CPU PM bench
                              (crates/gadget-ng-cuda/benches/cuda_vs_simd.rs:17-31):
                              positions (sin t, cos t, sin t/2) * 10 with
                              t = 0.37 i, box 20, m = 1/N.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SEED_UNIFORM = 0x6E670001
SEED_ZELDOVICH = 0x6E670002


@dataclass
class Particles:
    x: np.ndarray  # float32 SoA
    y: np.ndarray
    z: np.ndarray
    m: np.ndarray
    box: float
    name: str

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def pos64(self) -> np.ndarray:
        """(N, 3) float64 positions holding exactly the float32 values."""
        return np.stack([self.x, self.y, self.z], axis=1).astype(np.float64)

    def mass64(self) -> np.ndarray:
        return self.m.astype(np.float64)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=int(seed)))


def uniform_box(n: int = 32 ** 3, box: float = 100.0, seed: int = SEED_UNIFORM) -> Particles:
    g = _rng(seed)
    pos = g.uniform(0.0, box, (n, 3))
    p = pos.astype(np.float32)
    m = np.full(n, 1.0 / n, dtype=np.float32)
    return Particles(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), m, float(box), f"uniform N={n} L={box}")


def ap08_curve(grid: int) -> Particles:
    n = grid * grid
    t = 0.37 * np.arange(n, dtype=np.float64)
    p = (np.stack([np.sin(t), np.cos(t), np.sin(0.5 * t)], axis=1) * 10.0).astype(np.float32)
    m = np.full(n, 1.0 / n, dtype=np.float32)
    return Particles(p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), m, 20.0, f"ap08 grid={grid} N={n}")


def zeldovich_displacement_field(nm: int, box: float, spectral_index: float, seed: int, workers: int = -1):
    """Zel'dovich displacement Psi (3, nm, nm, nm) float64 of a Gaussian field.

    delta-hat is white noise shaped by sqrt(P(k)), P ~ k^n for 0 < |k| <= k_Nyq
    (spherical cut at the mesh Nyquist), the DC mode and every Nyquist plane
    zeroed; Psi-hat = i k delta-hat / k^2 (so div Psi = -delta).  The
    amplitude is arbitrary (callers rescale).
    """
    import scipy.fft as sf

    g = _rng(seed)
    white = g.standard_normal((nm, nm, nm))
    dk = sf.rfftn(white, workers=workers)
    del white
    n1 = np.fft.fftfreq(nm, 1.0 / nm)
    nh = np.arange(nm // 2 + 1, dtype=np.float64)
    kf = 2.0 * np.pi / box
    kz = (kf * n1)[:, None, None]
    ky = (kf * n1)[None, :, None]
    kx = (kf * nh)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    k2[0, 0, 0] = 1.0
    kmax = kf * (nm // 2)
    amp = np.where(k2 <= kmax * kmax, k2 ** (0.25 * spectral_index), 0.0)
    amp[0, 0, 0] = 0.0
    # zero the Nyquist planes (self-conjugate modes)
    amp[nm // 2, :, :] = 0.0
    amp[:, nm // 2, :] = 0.0
    amp[:, :, nm // 2] = 0.0
    dk *= amp
    del amp
    psi = np.empty((3, nm, nm, nm), dtype=np.float64)
    for a, ka in enumerate((kx, ky, kz)):
        psi[a] = sf.irfftn(1j * ka * dk / k2, s=(nm, nm, nm), workers=workers)
    return psi


def zeldovich_lattice(n_side: int = 256, nm: int = 512, box: float = 256.0, spectral_index: float = -2.0,
                      rms_disp_cells: float = 0.5, seed: int = SEED_ZELDOVICH, mass: float = 1.0,
                      workers: int = -1) -> Particles:
    """Lattice q = i * box/n_side (node-aligned on the nm mesh) displaced by Psi(q),
    Psi scaled to a 3-D rms |Psi| of rms_disp_cells mesh cells; positions wrapped
    into [0, box) and rounded to float32; equal masses."""
    if nm % n_side != 0:
        raise ValueError("lattice must be node-aligned: nm must be a multiple of n_side")
    step = nm // n_side
    psi = zeldovich_displacement_field(nm, box, spectral_index, seed, workers)
    sub = psi[:, ::step, ::step, ::step]  # [comp, iz, iy, ix] at the lattice nodes
    cell = box / nm
    rms = float(np.sqrt((sub ** 2).sum(axis=0).mean()))
    scale = rms_disp_cells * cell / rms if rms > 0 else 0.0
    q = np.arange(n_side, dtype=np.float64) * (box / n_side)
    out = []
    for a in range(3):
        # axis a of the position: x varies fastest (flat index iz*n^2 + iy*n + ix)
        shape = [1, 1, 1]
        shape[2 - a] = n_side
        qa = np.broadcast_to(q.reshape(shape), (n_side, n_side, n_side))
        xa = np.mod(qa + scale * sub[a], box)
        out.append(xa.astype(np.float32).reshape(-1))
    del psi, sub
    # float32 rounding can land exactly on box; wrap those to 0
    for xa in out:
        xa[xa >= np.float32(box)] = 0.0
    n = n_side ** 3
    m = np.full(n, mass, dtype=np.float32)
    return Particles(out[0], out[1], out[2], m, float(box),
                     f"zeldovich lattice {n_side}^3 on {nm}^3 L={box} n_s={spectral_index}")


def algorithmic_bytes(n: int, nm: int) -> int:
    """B = 40 N + 32 M + 64 M_h (FP32 path): particles 16 B in + 12 B re-read +
    12 B out; mesh deposit write 4M, R2C 4M + 8M_h, k-space 8M_h + 24M_h,
    3 C2R 24M_h + 12M, interpolation read 12M."""
    m = nm ** 3
    mh = nm * nm * (nm // 2 + 1)
    return 40 * n + 32 * m + 64 * mh
