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
* :func:`clustered_box`       70 % of the particles on a Zel'dovich lattice, 30 %
                              in Plummer spheres (512^3 particles, 1024^3 mesh,
                              L = 512: the TreePM long-range case at a
                              clustered shape).
* :func:`ensemble`            64 independent boxes of uneven size (32^3 to 128^3
                              particles, meshes 64^3 to 256^3) solved together.
* :func:`ap08_curve`          the input of the reference's GPU-vs-CPU PM bench
                              (crates/gadget-ng-cuda/benches/cuda_vs_simd.rs:17-31):
                              positions (sin t, cos t, sin t/2) * 10 with
                              t = 0.37 i, box 20, m = 1/N.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

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
    p64: Optional[np.ndarray] = None  # (N, 3) float64 positions before the float32 rounding, when kept

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
                      workers: int = -1, keep_f64: bool = False) -> Particles:
    """Lattice q = i * box/n_side (node-aligned on the nm mesh) displaced by Psi(q),
    Psi scaled to a 3-D rms |Psi| of rms_disp_cells mesh cells; positions wrapped
    into [0, box) and rounded to float32; equal masses.  With keep_f64 the
    unrounded float64 positions are kept as well (``p64``)."""
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
    p64 = np.empty((n_side ** 3, 3), dtype=np.float64) if keep_f64 else None
    for a in range(3):
        # axis a of the position: x varies fastest (flat index iz*n^2 + iy*n + ix)
        shape = [1, 1, 1]
        shape[2 - a] = n_side
        qa = np.broadcast_to(q.reshape(shape), (n_side, n_side, n_side))
        xa = np.mod(qa + scale * sub[a], box)
        if keep_f64:
            p64[:, a] = xa.reshape(-1)
        out.append(xa.astype(np.float32).reshape(-1))
    del psi, sub
    # float32 rounding can land exactly on box; wrap those to 0
    for xa in out:
        xa[xa >= np.float32(box)] = 0.0
    n = n_side ** 3
    m = np.full(n, mass, dtype=np.float32)
    return Particles(out[0], out[1], out[2], m, float(box),
                     f"zeldovich lattice {n_side}^3 on {nm}^3 L={box} n_s={spectral_index}", p64)


SEED_CLUSTERED = 0x6E670003


def zeldovich_offgrid_lattice(n_side: int, box: float, spectral_index: float, rms_disp: float, seed: int,
                              workers: int = -1):
    """(3, n_side^3) float64 positions: lattice q = i * box/n_side displaced by a
    Zel'dovich field synthesised on the lattice's own grid (P ~ k^n up to that
    grid's Nyquist, DC and Nyquist planes zeroed), 3-D rms displacement rms_disp,
    wrapped into [0, box)."""
    psi = zeldovich_displacement_field(n_side, box, spectral_index, seed, workers)
    rms = float(np.sqrt((psi ** 2).sum(axis=0).mean()))
    scale = rms_disp / rms if rms > 0 else 0.0
    q = np.arange(n_side, dtype=np.float64) * (box / n_side)
    out = np.empty((3, n_side ** 3), dtype=np.float64)
    for a in range(3):
        shape = [1, 1, 1]
        shape[2 - a] = n_side
        out[a] = np.mod(np.broadcast_to(q.reshape(shape), (n_side,) * 3) + scale * psi[a], box).reshape(-1)
    return out


def plummer_spheres(n_total: int, n_spheres: int, box: float, a: float, seed: int, trunc: float = 20.0):
    """(3, n_total) float64 positions in n_spheres Plummer spheres of scale radius a:
    uniform centres, log-normal (sigma = 1) member counts rescaled to sum exactly
    to n_total, radii from the inverse of M(<r)/M = r^3 / (r^2 + a^2)^{3/2}
    truncated at trunc * a, isotropic directions, wrapped into [0, box)."""
    g = _rng(seed)
    centres = g.uniform(0.0, box, (n_spheres, 3))
    w = g.lognormal(0.0, 1.0, n_spheres)
    counts = np.floor(w / w.sum() * n_total).astype(np.int64)
    short = n_total - int(counts.sum())
    counts[np.argsort(-(w / w.sum() * n_total - counts))[:short]] += 1  # largest remainders
    assert int(counts.sum()) == n_total
    umax = trunc ** 3 / (trunc ** 2 + 1.0) ** 1.5
    owner = np.repeat(np.arange(n_spheres), counts)
    u = g.uniform(0.0, umax, n_total)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    cz = g.uniform(-1.0, 1.0, n_total)
    ph = g.uniform(0.0, 2.0 * np.pi, n_total)
    sz = np.sqrt(1.0 - cz * cz)
    out = np.empty((3, n_total), dtype=np.float64)
    out[0] = np.mod(centres[owner, 0] + r * sz * np.cos(ph), box)
    out[1] = np.mod(centres[owner, 1] + r * sz * np.sin(ph), box)
    out[2] = np.mod(centres[owner, 2] + r * cz, box)
    return out, counts


def clustered_box(n_total: int = 512 ** 3, lattice_side: int = 454, n_spheres: int = 2000, box: float = 512.0,
                  nm: int = 1024, a: float = 0.2, spectral_index: float = -2.0, rms_disp_cells: float = 0.5,
                  seed: int = SEED_CLUSTERED, workers: int = -1, keep_f64: bool = False) -> Particles:
    """Clustered TreePM long-range workload: a Zel'dovich lattice of lattice_side^3
    particles (rms displacement rms_disp_cells mesh cells) plus the remaining
    particles in Plummer spheres; equal unit masses, float32 positions (and,
    with keep_f64, the unrounded float64 ones as ``p64``)."""
    n_lat = lattice_side ** 3
    if n_lat > n_total:
        raise ValueError("lattice larger than the particle total")
    lat = zeldovich_offgrid_lattice(lattice_side, box, spectral_index, rms_disp_cells * box / nm, seed, workers)
    sph, _ = plummer_spheres(n_total - n_lat, n_spheres, box, a, seed + 1)
    xs = []
    p64 = np.empty((n_total, 3), dtype=np.float64) if keep_f64 else None
    for c in range(3):
        v64 = np.concatenate([lat[c], sph[c]])
        if keep_f64:
            p64[:, c] = v64
        v = v64.astype(np.float32)
        del v64
        v[v >= np.float32(box)] = 0.0
        xs.append(v)
    del lat, sph
    m = np.ones(n_total, dtype=np.float32)
    return Particles(xs[0], xs[1], xs[2], m, float(box),
                     f"clustered {lattice_side}^3 lattice + {n_total - n_lat} in {n_spheres} Plummer spheres, "
                     f"L={box} a={a}", p64)


SEED_ENSEMBLE = 0x6E670004


def ensemble(n_boxes: int = 64, sides=(32, 48, 64, 96, 128), seed: int = SEED_ENSEMBLE, workers: int = -1):
    """Independent periodic boxes of uneven size: for box b, N_side is drawn from
    `sides`, the mesh is 2 N_side, L_b = N_side (1 Mpc/h spacing), the spectral
    index n_s ~ U[-2.5, -1.5], and the particles follow the Zel'dovich recipe of
    zeldovich_lattice with seed + b.  Returns (boxes, offsets, x, y, z, m):
    per-box dicts (grid_size, box_size, n_side, n_s, particles) and the
    concatenated float32 arrays with box b at [offsets[b], offsets[b+1])."""
    g = _rng(seed)
    picks = g.integers(0, len(sides), n_boxes)
    ns = g.uniform(-2.5, -1.5, n_boxes)
    boxes = []
    for b in range(n_boxes):
        side = int(sides[picks[b]])
        p = zeldovich_lattice(n_side=side, nm=2 * side, box=float(side), spectral_index=float(ns[b]),
                              rms_disp_cells=0.5, seed=seed + b, workers=workers)
        boxes.append({"grid_size": 2 * side, "box_size": float(side), "n_side": side, "n_s": float(ns[b]),
                      "particles": p})
    offsets = np.zeros(n_boxes + 1, dtype=np.int64)
    offsets[1:] = np.cumsum([bx["particles"].n for bx in boxes])
    cat = [np.concatenate([getattr(bx["particles"], a) for bx in boxes]) for a in ("x", "y", "z", "m")]
    return boxes, offsets, cat[0], cat[1], cat[2], cat[3]


SEED_LEARNED = 0x6E670005


def learned_box(seed: int, cells: int = 256, workers: int = -1):
    """Learned-correction box: the clustered recipe at one particle per cell,
    L = cells (1 Mpc/h cells, mesh = cells^3), N = cells^3 particles, of which
    a lattice of round(227 cells/256)^3 is Zel'dovich-displaced (rms 0.5 cell,
    P ~ k^-2) and the rest sit in round(250 (cells/256)^3) Plummer spheres of
    a = 0.2 Mpc/h.  cells = 256 is the full 256^3 workload; smaller boxes keep
    the same density, resolution and sphere statistics.  Returns
    (Particles, n_lattice): particles [0, n_lattice) are the lattice."""
    lat = int(round(227 * cells / 256))
    nsph = max(1, int(round(250 * (cells / 256) ** 3)))
    p = clustered_box(n_total=cells ** 3, lattice_side=lat, n_spheres=nsph, box=float(cells), nm=cells,
                      a=0.2, seed=SEED_LEARNED + int(seed) * 7919, workers=workers)
    return p, lat ** 3


def algorithmic_bytes(n: int, nm: int) -> int:
    """B = 40 N + 32 M + 64 M_h (FP32 path): particles 16 B in + 12 B re-read +
    12 B out; mesh deposit write 4M, R2C 4M + 8M_h, k-space 8M_h + 24M_h,
    3 C2R 24M_h + 12M, interpolation read 12M."""
    m = nm ** 3
    mh = nm * nm * (nm // 2 + 1)
    return 40 * n + 32 * m + 64 * mh
