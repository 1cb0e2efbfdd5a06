"""Pins of the FP64 power-spectrum oracle (oracle/analysis_oracle.py).  No GPU.

* the reference's own tests (crates/gadget-ng-analysis/src/power_spectrum.rs:
  210-250): a uniform lattice gives P ~ 0, one particle gives finite P;
* closed forms: a particle on a mesh node has |delta_k|^2 = n^6 for every
  k != 0, so each shell holds V^2 <1/W^2> (enumerated mode by mode here, with
  no FFT); a single cosine mode of amplitude A fills one shell with
  2 |A n^3 / 2|^2 (V/n^3)^2 / W^2 / count;
* a brute-force DFT on a 4^3 mesh, the CIC weights, and the invariance of P
  under a mass rescaling.
"""
import itertools
import math

import numpy as np

import analysis_oracle as ao


def test_uniform_lattice_gives_zero_power():
    n, box = 4, 1.0
    cell = box / n
    g = (np.arange(n) + 0.5) * cell
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    k, pk, nm = ao.power_spectrum(pos, np.ones(len(pos)), box, n)
    assert len(k) > 0 and (pk < 1e-20).all()


def test_single_particle_is_finite():
    k, pk, nm = ao.power_spectrum(np.array([[0.5, 0.5, 0.5]]), np.ones(1), 1.0, 8)
    assert np.isfinite(pk).all() and len(pk) == 4


def _sinc(x):
    return 1.0 if abs(x) < 1e-12 else math.sin(math.pi * x) / (math.pi * x)


def test_particle_on_a_node_closed_form():
    n, box = 8, 3.0
    cell = box / n
    pos = np.array([[2 * cell, 5 * cell, 7 * cell]])
    k, pk, nm = ao.power_spectrum(pos, np.array([2.5]), box, n)
    acc, cnt = {}, {}
    for ix, iy, iz in itertools.product(range(n), repeat=3):
        f = [i if i <= n // 2 else i - n for i in (ix, iy, iz)]
        k2 = f[0] ** 2 + f[1] ** 2 + f[2] ** 2
        if k2 == 0:
            continue
        bf = math.sqrt(k2) - 0.5
        if bf >= n // 2:
            continue
        b = int(bf)
        w = _sinc(f[0] / n) * _sinc(f[1] / n) * _sinc(f[2] / n)
        acc[b] = acc.get(b, 0.0) + box ** 6 / w ** 2
        cnt[b] = cnt.get(b, 0) + 1
    bins = sorted(cnt)
    assert list(nm) == [cnt[b] for b in bins]
    np.testing.assert_allclose(k, [(b + 1) * 2 * math.pi / box for b in bins], rtol=1e-15)
    np.testing.assert_allclose(pk, [acc[b] / cnt[b] for b in bins], rtol=1e-12)


def test_single_cosine_mode_fills_one_shell():
    n, box, amp = 16, 5.0, 0.3
    p = (2, 1, 0)
    i = np.arange(n)
    X, Y, Z = np.meshgrid(i, i, i, indexing="ij")
    grid = 1.7 * (1.0 + amp * np.cos(2 * math.pi * (p[0] * X + p[1] * Y + p[2] * Z) / n))
    k, pk, nm = ao.pk_from_mass(grid, box)
    kp = math.sqrt(sum(c * c for c in p))
    b = int(kp - 0.5)
    w = _sinc(p[0] / n) * _sinc(p[1] / n) * _sinc(p[2] / n)
    shell_k = (b + 1) * 2 * math.pi / box
    j = int(np.argmin(np.abs(k - shell_k)))
    expect = 2 * (amp * n ** 3 / 2) ** 2 * (box ** 3 / n ** 3) ** 2 / w ** 2 / nm[j]
    assert abs(pk[j] - expect) / expect < 1e-12
    others = np.delete(pk, j)
    assert (others < 1e-20 * expect).all()


def test_dft_against_brute_force_on_4_cubed():
    rng = np.random.default_rng(1)
    n, box = 4, 2.0
    grid = rng.uniform(0.5, 1.5, (n, n, n))
    k, pk, nm = ao.pk_from_mass(grid, box)
    delta = grid / (grid.sum() / box ** 3 * (box / n) ** 3) - 1.0
    shell, w2 = ao.shells(n)
    sums = np.zeros(n // 2)
    for a, b_, c in itertools.product(range(n), repeat=3):
        dk = sum(delta[x, y, z] * np.exp(-2j * math.pi * (a * x + b_ * y + c * z) / n)
                 for x, y, z in itertools.product(range(n), repeat=3))
        if shell[a, b_, c] >= 0:
            sums[shell[a, b_, c]] += abs(dk) ** 2 / w2[a, b_, c] * (box ** 3 / n ** 3) ** 2
    counts = np.bincount(shell[shell >= 0], minlength=n // 2)
    keep = counts > 0
    np.testing.assert_allclose(pk, sums[keep] / counts[keep], rtol=1e-10)


def test_cic_weights():
    n, box = 4, 4.0
    g = ao.cic_grid(np.array([[0.25, 3.5, 1.0]]), np.array([2.0]), box, n)
    assert abs(g.sum() - 2.0) < 1e-15
    # x: 0.75 to cell 0, 0.25 to cell 1; y: 0.5 to cells 3 and 0 (wrapped); z: all in cell 1
    assert abs(g[0, 3, 1] - 2.0 * 0.75 * 0.5) < 1e-15 and abs(g[1, 0, 1] - 2.0 * 0.25 * 0.5) < 1e-15
    assert np.count_nonzero(g) == 4


def test_power_is_invariant_under_a_mass_rescaling():
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 10.0, (500, 3))
    m = rng.uniform(0.5, 1.5, 500)
    a = ao.power_spectrum(pos, m, 10.0, 16)
    b = ao.power_spectrum(pos, 7.0 * m, 10.0, 16)
    np.testing.assert_allclose(a[1], b[1], rtol=1e-12)
    assert np.array_equal(a[2], b[2])
