"""Pins of the FP64 oracle that do not depend on retyping its formulas.

* closed forms: a single Fourier mode has an exact discrete solution, which
  fixes the FFT normalisation, the sign, the Green's function, the filters,
  the deconvolution power and the x-fastest axis layout;
* the Nyquist rule (the n_alpha = nm/2 mode gives zero force);
* brute force: a naive long-double DFT at nm <= 8 and numpy's pocketfft;
* the two solve modes (full complex Z2Z + Re(), R2C with the Nyquist rule)
  agree to rounding;
* the CIC deposit equals an explicit evaluation of the trilinear kernel;
* exact invariants: translation by whole cells, particle permutation,
  linearity in mass and G;
* the committed golden vectors (tests/golden/, independent numpy code).
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def single_mode_mass(nm, box, q, axis, amp, kind="sin"):
    """Mass per cell V_cell * A * sin(2 pi q j / nm) along `axis` (0 = x)."""
    j = np.arange(nm)
    f = np.sin if kind == "sin" else np.cos
    wave = f(2 * np.pi * q * j / nm)
    shape = [1, 1, 1]
    shape[2 - axis] = nm
    vcell = (box / nm) ** 3
    return np.broadcast_to(vcell * amp * wave.reshape(shape), (nm, nm, nm)).copy()


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("q,rs,eps,dc", [(1, 0.0, 0.0, False), (3, 0.0, 0.0, False), (5, 0.7, 0.0, False),
                                          (2, 0.0, 0.4, False), (6, 0.0, 0.0, True), (7, 0.5, 0.2, True)])
def test_single_mode_closed_form(oracle_mod, mode, axis, q, rs, eps, dc):
    """rho = A sin(k x), k = 2 pi q / L: F_alpha = (4 pi G A / k) cos(k x) W(k) [/ W_CIC^2]."""
    nm, box, g, amp = 16, 3.7, 1.3, 0.8
    mass = single_mode_mass(nm, box, q, axis, amp)
    f = oracle_mod.solve_forces(mass, g, nm, box, r_split=rs, plummer_eps=eps, deconvolve=dc, mode=mode)
    k = 2 * np.pi * q / box
    w = 1.0
    if rs > 0:
        w *= np.exp(-0.5 * k * k * rs * rs)
    if eps > 0:
        w *= np.exp(-k * k * eps * eps)
    if dc:
        w /= np.sinc(q / nm) ** 4  # W_CIC^2 with W_CIC = sinc^2 along the one axis of the mode
    x = np.arange(nm) * box / nm
    expect = (4 * np.pi * g * amp / k) * np.cos(k * x) * w
    shape = [1, 1, 1]
    shape[2 - axis] = nm
    expect = np.broadcast_to(expect.reshape(shape), (nm, nm, nm))
    # FFT rounding scales with the unfiltered input, so bound by the larger amplitude
    scale = max(np.abs(expect).max(), 4 * np.pi * g * amp / k)
    assert np.abs(f[axis] - expect).max() < 1e-12 * scale
    for other in range(3):
        if other != axis:
            assert np.abs(f[other]).max() < 1e-12 * scale


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_nyquist_mode_gives_zero_force(oracle_mod, mode, axis):
    nm, box = 8, 2.0
    mass = single_mode_mass(nm, box, nm // 2, axis, 1.0, kind="cos")
    f = oracle_mod.solve_forces(mass, 1.0, nm, box, mode=mode)
    assert max(np.abs(c).max() for c in f) < 1e-12


def naive_dft(a, inverse=False):
    """O(M^2) 3-D DFT in long double (forward sign e^{-2 pi i})."""
    a = np.asarray(a, dtype=np.clongdouble)
    nm = a.shape[0]
    j = np.arange(nm, dtype=np.longdouble)
    sign = 1 if inverse else -1
    w = np.exp(sign * 2j * np.pi * np.outer(j, j) / nm).astype(np.clongdouble)
    # sum over (jz, jy, jx) of a[jz,jy,jx] w[kz,jz] w[ky,jy] w[kx,jx], explicitly
    out = np.zeros_like(a)
    for kz in range(nm):
        for ky in range(nm):
            for kx in range(nm):
                out[kz, ky, kx] = np.sum(a * w[kz][:, None, None] * w[ky][None, :, None] * w[kx][None, None, :])
    return out


@pytest.mark.parametrize("nm", [2, 3, 4, 6, 8])
def test_fft3d_vs_naive_long_double_dft(oracle_mod, nm):
    rng = np.random.default_rng(nm)
    a = rng.standard_normal((nm, nm, nm)) + 1j * rng.standard_normal((nm, nm, nm))
    ref = naive_dft(a)
    got = oracle_mod.fft3d(a)
    assert np.abs(got - ref.astype(np.complex128)).max() < 1e-13 * np.abs(ref).max()
    refi = naive_dft(a, inverse=True)
    goti = oracle_mod.fft3d(a, inverse=True)
    assert np.abs(goti - refi.astype(np.complex128)).max() < 1e-13 * np.abs(refi).max()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 24, 25, 48, 96, 128, 192, 512])
def test_fft1d_vs_pocketfft(oracle_mod, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    assert np.abs(oracle_mod.fft1d(a) - np.fft.fft(a, axis=-1)).max() < 1e-13 * max(1, np.log2(max(n, 2))) * np.sqrt(n)
    assert np.abs(oracle_mod.fft1d(a, inverse=True) - n * np.fft.ifft(a, axis=-1)).max() < 1e-13 * np.sqrt(n) * 10


@pytest.mark.parametrize("nm", [8, 12, 16, 24])
@pytest.mark.parametrize("rs,eps,dc", [(0.0, 0.0, False), (1.0, 0.0, False), (0.0, 0.3, True)])
def test_faithful_and_lean_modes_agree(oracle_mod, nm, rs, eps, dc):
    rng = np.random.default_rng(nm)
    pos = rng.uniform(-3, 13, (800, 3))
    m = rng.uniform(0.2, 3.0, 800)
    a = oracle_mod.pm_accelerations(pos, m, 10.0, nm, 1.0, r_split=rs, plummer_eps=eps, deconvolve=dc,
                                    mode=oracle_mod.MODE_FAITHFUL, deposit_mode=0)
    b = oracle_mod.pm_accelerations(pos, m, 10.0, nm, 1.0, r_split=rs, plummer_eps=eps, deconvolve=dc,
                                    mode=oracle_mod.MODE_LEAN)
    assert np.abs(a - b).max() < 1e-13 * np.abs(a).max()


@pytest.mark.parametrize("deposit_mode", [0, 1, 2])
def test_deposit_equals_explicit_cic_kernel(oracle_mod, deposit_mode):
    """Mass at node n = sum_p m_p prod_a max(0, 1 - |periodic dist(x_p,a - n_a h)| / h)."""
    rng = np.random.default_rng(4)
    nm, box = 6, 2.4
    h = box / nm
    pos = rng.uniform(-box, 2 * box, (40, 3))
    m = rng.uniform(0.1, 1.0, 40)
    rho = oracle_mod.cic_assign(pos, m, box, nm, deposit_mode=deposit_mode)
    nodes = np.arange(nm) * h
    expect = np.zeros((nm, nm, nm))
    for p in range(40):
        w = []
        for a in range(3):
            d = np.abs(((pos[p, a] - nodes) + box / 2) % box - box / 2)
            w.append(np.maximum(0.0, 1.0 - d / h))
        expect += m[p] * w[2][:, None, None] * w[1][None, :, None] * w[0][None, None, :]
    assert np.abs(rho - expect).max() < 1e-13


def test_translation_by_whole_cells_permutes_meshes(oracle_mod):
    rng = np.random.default_rng(5)
    nm, box = 16, 8.0
    h = box / nm
    pos = rng.uniform(0, box, (500, 3))
    m = rng.uniform(0.5, 1.5, 500)
    a0 = oracle_mod.pm_accelerations(pos, m, box, nm, 1.0)
    shift = np.array([3, -5, 7]) * h
    a1 = oracle_mod.pm_accelerations(pos + shift, m, box, nm, 1.0)
    assert np.abs(a1 - a0).max() < 1e-11 * np.abs(a0).max()


def test_particle_permutation_invariance(oracle_mod):
    rng = np.random.default_rng(6)
    pos = rng.uniform(0, 5.0, (700, 3))
    m = rng.uniform(0.5, 1.5, 700)
    perm = rng.permutation(700)
    a0 = oracle_mod.pm_accelerations(pos, m, 5.0, 12, 1.0)
    a1 = oracle_mod.pm_accelerations(pos[perm], m[perm], 5.0, 12, 1.0)
    assert np.abs(a1 - a0[perm]).max() < 1e-12 * np.abs(a0).max()


def test_linearity_in_mass_and_g(oracle_mod):
    rng = np.random.default_rng(8)
    pos = rng.uniform(0, 5.0, (300, 3))
    m1 = rng.uniform(0.5, 1.5, 300)
    m2 = rng.uniform(0.5, 1.5, 300)
    a1 = oracle_mod.pm_accelerations(pos, m1, 5.0, 8, 1.0)
    a2 = oracle_mod.pm_accelerations(pos, m2, 5.0, 8, 1.0)
    a12 = oracle_mod.pm_accelerations(pos, 2.0 * m1 + m2, 5.0, 8, 0.5)
    assert np.abs(a12 - (a1 + 0.5 * a2)).max() < 1e-12 * np.abs(a12).max()


GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))


@pytest.mark.parametrize("name", sorted(GOLDEN))
@pytest.mark.parametrize("mode", [0, 1])
def test_oracle_matches_golden_vectors(oracle_mod, name, mode):
    d = np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False)
    nm, n, box, g, rs, eps, dc = d["params"]
    nm = int(nm)
    acc, dens = oracle_mod.pm_accelerations(d["pos"], d["mass"], box, nm, g, r_split=rs, plummer_eps=eps,
                                            deconvolve=bool(dc), mode=mode, deposit_mode=0, want_density=True)
    assert np.abs(acc - d["acc"]).max() <= 1e-12 * np.abs(d["acc"]).max()
    if "density" in d:
        assert np.abs(dens - d["density"]).max() <= 1e-12 * np.abs(d["density"]).max()


def test_golden_files_match_recorded_hashes():
    import hashlib
    for name, meta in GOLDEN.items():
        blob = open(os.path.join(HERE, "golden", f"{name}.npz"), "rb").read()
        assert hashlib.sha256(blob).hexdigest() == meta["sha256"], name
