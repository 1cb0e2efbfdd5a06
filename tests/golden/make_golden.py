"""Writes the PM golden vectors under tests/golden/ (run once; output committed).

The vectors come from an independent numpy restatement of the reference CPU PM
(written separately from the C oracle): deposit with np.add.at following
crates/gadget-ng-pm/src/cic.rs:37-93, numpy.fft.fftn / ifftn (pocketfft, the
same forward-sign convention as rustfft) for fft_poisson.rs:306-476, the
spectral kernel of fft_poisson.rs:509-549 and interpolation following
cic.rs:100-170.  The reference itself ships no numeric vectors for this path.

Each case stores inputs (positions, masses, box, mesh, G, filters) and
outputs (accelerations; the mass-per-cell mesh for nm <= 16) in one .npz
(loaded with allow_pickle=False) and lists its SHA-256 in golden.json.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = [
    # name, nm, n, box, G, r_split, eps, deconv, seed, lo, hi
    ("nm4_plain", 4, 16, 1.0, 1.0, 0.0, 0.0, False, 101, 0.0, 1.0),
    ("nm8_rsplit", 8, 200, 10.0, 2.0, 1.0, 0.0, False, 102, -5.0, 15.0),
    ("nm12_rsplit_radix3", 12, 500, 3.0, 1.0, 0.4, 0.0, False, 103, 0.0, 3.0),
    ("nm16_eps_deconv", 16, 1000, 7.5, 0.5, 0.0, 0.3, True, 104, 0.0, 7.5),
    ("nm32_plain", 32, 4096, 100.0, 1.0, 0.0, 0.0, False, 105, 0.0, 100.0),
]


def cell(pos, box, nm):
    c = np.mod(pos * (nm / box), float(nm))
    i0 = np.floor(c)
    d = c - i0
    i0 = i0.astype(np.int64) % nm
    return i0, d


def assign(pos, mass, box, nm):
    i0, d = cell(pos, box, nm)
    rho = np.zeros((nm, nm, nm))
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = (np.where(c, d[:, 0], 1 - d[:, 0]) * np.where(b, d[:, 1], 1 - d[:, 1])
                     * np.where(a, d[:, 2], 1 - d[:, 2]))
                np.add.at(rho, ((i0[:, 2] + a) % nm, (i0[:, 1] + b) % nm, (i0[:, 0] + c) % nm), mass * w)
    return rho


def forces(rho, g, nm, box, r_split, eps, deconv):
    rk = np.fft.fftn(rho / (box / nm) ** 3)
    n = np.fft.fftfreq(nm, 1.0 / nm)
    n[nm // 2] = nm // 2  # index nm/2 maps to +nm/2 (fft_poisson.rs:724-730)
    kz, ky, kx = np.meshgrid(2 * np.pi / box * n, 2 * np.pi / box * n, 2 * np.pi / box * n, indexing="ij")
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    filt = np.ones_like(k2)
    if r_split > 0:
        filt *= np.exp(-0.5 * k2 * r_split ** 2)
    if eps > 0:
        filt *= np.exp(-k2 * eps ** 2)
    if deconv:
        s = np.sinc(n / nm)
        w = (s[:, None, None] * s[None, :, None] * s[None, None, :]) ** 2
        filt /= w ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        green = np.where(k2 < 1e-30, 0.0, -4 * np.pi * g * filt / np.where(k2 < 1e-30, 1.0, k2))
    phi = rk * green
    return [np.real(np.fft.ifftn(-1j * ka * phi)) for ka in (kx, ky, kz)]


def interpolate(f, pos, box, nm):
    i0, d = cell(pos, box, nm)
    acc = np.zeros((pos.shape[0], 3))
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                w = (np.where(c, d[:, 0], 1 - d[:, 0]) * np.where(b, d[:, 1], 1 - d[:, 1])
                     * np.where(a, d[:, 2], 1 - d[:, 2]))
                ix = ((i0[:, 2] + a) % nm, (i0[:, 1] + b) % nm, (i0[:, 0] + c) % nm)
                for comp in range(3):
                    acc[:, comp] += f[comp][ix] * w
    return acc


def main():
    meta = {}
    for name, nm, n, box, g, rs, eps, dc, seed, lo, hi in CASES:
        rng = np.random.default_rng(seed)
        pos = rng.uniform(lo, hi, (n, 3))
        mass = rng.uniform(0.5, 2.0, n)
        rho = assign(pos, mass, box, nm)
        f = forces(rho, g, nm, box, rs, eps, dc)
        acc = interpolate(f, pos, box, nm)
        arrays = dict(pos=pos, mass=mass, acc=acc,
                      params=np.array([nm, n, box, g, rs, eps, float(dc)], dtype=np.float64))
        if nm <= 16:
            arrays["density"] = rho
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **arrays)
        meta[name] = {"nm": nm, "n": n, "box": box, "g": g, "r_split": rs, "plummer_eps": eps, "deconvolve": dc,
                      "sha256": hashlib.sha256(open(path, "rb").read()).hexdigest()}
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(meta, fh, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
