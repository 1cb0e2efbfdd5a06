"""Pins of the Zel'dovich initial-condition oracle (oracle/ic_oracle.py) and
of the library's host-side growth functions against it.  No GPU needed.

* The generator: the ChaCha block function against RFC 8439 sec. 2.3.2 and
  the all-zero ChaCha20 keystream (RFC 8439 A.1 #1); ChaCha12 with next_u64
  against the value-stability vector of rand's StdRng (rand src/rngs/std.rs,
  test_stdrng_construction: seed bytes 1,0,0,0, 23,0,0,0, 200,1,0,0,
  210,30,0,0, then zeros; a second generator seeded from the next 32 bytes);
  the PCG32 step and output against the pcg32 reference demo (seed 42,
  sequence 54).  The Gaussians by their moments.
* The field: the reference's own tests (DC zero, Hermitian symmetry, real
  displacement, zero mean displacement; ic_zeldovich.rs:1019-1105), every
  Nyquist mode zero, and |delta|^2 / (2 sigma^2) averaging to 1.
* The displacement: a planted single mode against its closed form.
* Growth: Einstein-de Sitter closed forms and the reference's own tests
  (cosmology.rs:925-976); the ODE branch against scipy's integration of the
  growth equation.
"""
import math
import struct

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import ic_oracle as ico
from gadget_ng_pm_hip import integrate as gi

C = np.array(ico.CHACHA_CONST, dtype=np.uint32)


def _block(words, rounds):
    return [int(w) for w in ico.chacha_block(np.array(words, dtype=np.uint32).reshape(16, 1), rounds)[:, 0]]


def test_chacha20_rfc8439_block_vector():
    key = list(struct.unpack("<8I", bytes(range(32))))
    inp = list(C) + key + [1, 0x09000000, 0x4A000000, 0]
    out = _block(inp, 20)
    assert out == [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7, 0x0368C033, 0x9AAA2204, 0x4E6CD4C3,
                   0x466482D2, 0x09AA9F07, 0x05D7C214, 0xA2028BD9, 0xD19C12B5, 0xB94E16DE, 0xE883D0CB, 0x4E3C50A2]


def test_chacha20_zero_key_keystream():
    out = _block(list(C) + [0] * 12, 20)
    ks = struct.pack("<16I", *out)
    assert ks[:16].hex() == "76b8e0ada0f13d90405d6ae55386bd28"
    assert ks[16:32].hex() == "bdd219b8a08ded1aa836efcc8b770dc7"


def test_stdrng_value_stability_vector():
    seed = bytes([1, 0, 0, 0, 23, 0, 0, 0, 200, 1, 0, 0, 210, 30, 0, 0] + [0] * 16)
    key = np.array(struct.unpack("<8I", seed), dtype=np.uint32).reshape(8, 1)
    u0, _ = ico.stdrng_first_u64s(key)
    assert int(u0[0]) == 10719222850664546238
    # from_rng: the next 32 bytes of the stream (words 2..9 of block 0) key a second generator
    full = _block(list(C) + [int(k) for k in key[:, 0]] + [0, 0, 0, 0], 12)
    key2 = np.array(full[2:10], dtype=np.uint32).reshape(8, 1)
    v0, _ = ico.stdrng_first_u64s(key2)
    assert int(v0[0]) == 14064965282130556830


def test_pcg32_reference_demo_vector():
    inc = (54 << 1) | 1
    st = np.uint64(0)
    st = ico.pcg32_advance(st, inc=inc)
    with np.errstate(over="ignore"):
        st = np.uint64(st) + np.uint64(42)
    st = ico.pcg32_advance(st, inc=inc)
    outs = []
    for _ in range(6):
        outs.append(int(ico.pcg32_output(st)))
        st = ico.pcg32_advance(st, inc=inc)
    assert outs == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_seed_expansion_changes_every_key_word():
    k = ico.seed_from_u64_key(np.array([0, 1, 2, 3], dtype=np.uint64))
    assert k.shape == (8, 4)
    assert len({tuple(k[:, j]) for j in range(4)}) == 4
    # seed 0 still gives a key of about half set bits (the LCG is advanced first)
    bits = sum(bin(int(w)).count("1") for w in k[:, 0])
    assert 96 <= bits <= 160


def test_gaussian_moments():
    g = ico.gaussian_std(ico.hash_u64(np.arange(200000, dtype=np.uint64)))
    assert abs(g.mean()) < 5 / math.sqrt(g.size)
    assert abs(g.var() - 1.0) < 5 * math.sqrt(2.0 / g.size)
    assert abs((g ** 4).mean() - 3.0) < 0.05


def test_hash_is_a_bijection_on_samples():
    x = np.arange(1 << 16, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    assert np.unique(ico.hash_u64(x)).size == x.size
    assert int(ico.hash_u64(np.uint64(0))) == 0


@pytest.mark.parametrize("n,seed", [(8, 42), (8, 99), (12, 5), (16, 1234)])
def test_field_dc_nyquist_and_hermitian(n, seed):
    d = ico.delta_kspace(n, seed, ico.power_law_sigma(n, 1e-4, -2.0))
    assert d[0, 0, 0] == 0
    h = n // 2
    assert np.all(d[h, :, :] == 0) and np.all(d[:, h, :] == 0) and np.all(d[:, :, h] == 0)
    neg = d[(-np.arange(n)) % n][:, (-np.arange(n)) % n][:, :, (-np.arange(n)) % n]
    assert np.array_equal(d, np.conj(neg))
    assert np.count_nonzero(d) == (n - 1) ** 3 - 1


def test_field_variance_matches_the_spectrum():
    n = 32
    sig = ico.power_law_sigma(n, 1.0, -1.0)
    d = ico.delta_kspace(n, 7, sig)
    f = ico.mode_int(np.arange(n), n)
    n2 = f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2
    live = d != 0
    r = np.abs(d[live]) ** 2 / (2.0 * sig(n2[live]) ** 2)
    assert abs(r.mean() - 1.0) < 5 / math.sqrt(r.size / 2)


def test_displacement_real_and_zero_mean():
    n = 16
    d = ico.delta_kspace(n, 1234, ico.power_law_sigma(n, 1e-4, -2.0))
    assert np.abs(np.fft.ifftn(d).imag).max() < 1e-10
    for p in ico.displacement(d, 1.0):
        assert abs(p.mean()) < 1e-12


@pytest.mark.parametrize("axis,m", [(0, 1), (1, 3), (2, 5)])
def test_planted_mode_closed_form(axis, m):
    """delta-hat(k0) = c, delta-hat(-k0) = conj c with k0 = m e_axis gives
    Psi_axis(j) = -2 Im(c e^{2 pi i m j / n}) d / (m n^3), other components 0."""
    n, box, c = 16, 3.0, 0.7 - 0.4j
    d = np.zeros((n, n, n), dtype=np.complex128)
    k0 = [0, 0, 0]
    k0[axis] = m
    d[tuple(k0)] = c
    k0[axis] = n - m
    d[tuple(k0)] = np.conj(c)
    psi = ico.displacement(d, box)
    j = np.arange(n)
    prof = -2.0 * np.imag(c * np.exp(2j * np.pi * m * j / n)) * (box / n) / (m * n ** 3)
    shape = [1, 1, 1]
    shape[axis] = n
    want = np.broadcast_to(prof.reshape(shape), (n, n, n))
    for a in range(3):
        np.testing.assert_allclose(psi[a], want if a == axis else 0.0, rtol=0, atol=1e-13 * np.abs(prof).max())


def test_lattice_particles_follow_the_displacement():
    n, box = 8, 10.0
    d = ico.delta_kspace(n, 3, ico.power_law_sigma(n, 30.0, -2.0))
    psi = ico.displacement(d, box)
    pos, vel = ico.zeldovich(n, box, psi, vel_factor=2.0)
    assert pos.shape == (n ** 3, 3)
    g = 37  # ix = 0, iy = 4, iz = 5
    q = np.array([0.5, 4.5, 5.5]) * box / n
    p = np.array([psi[0][0, 4, 5], psi[1][0, 4, 5], psi[2][0, 4, 5]])
    np.testing.assert_array_equal(vel[g], p * 2.0)
    np.testing.assert_array_equal(pos[g], ico.rem_euclid(q + p, box))
    assert pos.min() >= 0.0 and pos.max() < box
    sub_p, sub_v = ico.zeldovich(n, box, psi, 2.0, lo=30, hi=40)
    np.testing.assert_array_equal(sub_p, pos[30:40])


EDS = (1.0, 0.0, 0.7, -1.0, 0.0, 0.0)
LCDM = (0.3, 0.7, 1.0, -1.0, 0.0, 0.0)


def test_growth_eds_closed_forms():
    for a in (0.01, 0.1, 0.5, 1.0, 2.0):
        assert abs(ico.growth_factor_d(EDS, a) - a) / a < 1e-12
        assert ico.growth_rate_f(EDS, a) == pytest.approx(1.0, abs=1e-14)
        assert ico.hubble_param(EDS, a) == pytest.approx(0.7 * a ** -1.5, rel=1e-14)
    assert ico.growth_factor_d(LCDM, 0.0) == 0.0


def test_growth_reference_ratio_checks():
    for a in (0.01, 0.02, 0.05, 0.10, 0.5):
        s = ico.growth_factor_d(LCDM, a) / ico.growth_factor_d(LCDM, 1.0)
        assert 0.0 < s < 1.0
        assert ico.growth_factor_d(LCDM, min(1.2 * a, 1.0)) / ico.growth_factor_d(LCDM, 1.0) >= s - 1e-12
    s = ico.growth_factor_d(LCDM, 0.02) / ico.growth_factor_d(LCDM, 1.0)
    assert 0.02 < s < 0.05


@pytest.mark.parametrize("c", [(0.3, 0.7, 1.0, -0.9, 0.1, 0.0), (0.3, 0.7, 1.0, -1.0 + 1e-6, 0.0, 0.0)])
def test_growth_ode_against_scipy(c):
    """D'' + (2 + dlnH/dlna) D' - 3/2 Omega_m(a) D = 0 in ln a, from D = D' =
    a at a = 1e-5, by scipy's RK45 with the derivative of ln H evaluated
    analytically from the dark-energy density."""
    om, ol, h0, w0, wa, onu = c

    def e2(a):
        return (om + onu) / a ** 3 + ol * a ** (-3.0 * (1 + w0 + wa)) * math.exp(3.0 * wa * (a - 1.0))

    def dlnh(a):
        de = ol * a ** (-3.0 * (1 + w0 + wa)) * math.exp(3.0 * wa * (a - 1.0))
        d_de = de * (-3.0 * (1 + w0 + wa) + 3.0 * wa * a)
        return 0.5 * (-3.0 * (om + onu) / a ** 3 + d_de) / e2(a)

    def rhs(la, y):
        a = math.exp(la)
        return [y[1], -(2.0 + dlnh(a)) * y[1] + 1.5 * (om + onu) / a ** 3 / e2(a) * y[0]]
    for a in (0.1, 0.5, 1.0):
        sol = solve_ivp(rhs, (math.log(1e-5), math.log(a)), [1e-5, 1e-5], rtol=1e-10, atol=1e-16)
        # 2e-4: the restated method differences H with a step of 1e-7, 1 % of
        # a at the start (a = 1e-5); a wrong term would be off by percents
        assert ico.growth_factor_d(c, a) == pytest.approx(sol.y[0, -1], rel=2e-4)


@pytest.mark.parametrize("c", [EDS, LCDM, (0.3, 0.7, 0.1, -0.8, 0.3, 0.0), (0.28, 0.7, 0.2, -0.9, -0.2, 0.02)])
def test_library_growth_equals_the_oracle_bit_for_bit(c):
    cp = gi.CosmologyParams(*c)
    for a in (0.02, 0.3, 1.0, 1.7):
        assert cp.hubble(a) == ico.hubble_param(c, a)
        assert cp.growth_rate(a) == ico.growth_rate_f(c, a)
        assert cp.growth_factor(a) == ico.growth_factor_d(c, a)


def test_2lpt_planted_modes_closed_form():
    """delta-hat with one cosine mode along axis 0 (m1) and one along axis 1
    (m2): phi,xx = -delta1(x), phi,yy = -delta2(y), every other derivative 0,
    so S = delta1 delta2 = A cos(t1) cos(t2), A = 4 c1 c2 / n^6, and Psi2_a =
    d (A/2) sum over m = (m1, +-m2, 0) of m_a / |m|^2 sin(t1 +- t2)."""
    n, box, m1, m2, c1, c2 = 16, 2.0, 2, 3, 0.8, -0.5
    d = np.zeros((n, n, n), dtype=np.complex128)
    d[m1, 0, 0] = d[n - m1, 0, 0] = c1
    d[0, m2, 0] = d[0, n - m2, 0] = c2
    phi = ico.phi_second_derivatives(d)
    j = np.arange(n)
    t1 = (2 * np.pi * m1 * j / n)[:, None, None]
    t2 = (2 * np.pi * m2 * j / n)[None, :, None]
    amp = 4 * c1 * c2 / n ** 6
    s = ico.source_2lpt(phi)
    np.testing.assert_allclose(s, np.broadcast_to(amp * np.cos(t1) * np.cos(t2), s.shape), rtol=0,
                               atol=1e-12 * abs(amp))
    psi2 = ico.source_to_psi2(s, box)
    want = [np.zeros((n, n, n)) for _ in range(3)]
    for sgn in (1, -1):
        m = np.array([m1, sgn * m2, 0.0])
        mm = (m * m).sum()
        for a in range(3):
            want[a] = want[a] + (box / n) * 0.5 * amp * m[a] / mm * np.sin(t1 + sgn * t2)
    for a in range(3):
        np.testing.assert_allclose(psi2[a], np.broadcast_to(want[a], (n, n, n)), rtol=0,
                                   atol=1e-12 * abs(amp) * box)


def test_2lpt_source_is_quadratic_and_psi2_real():
    n = 16
    d = ico.delta_kspace(n, 42, ico.power_law_sigma(n, 1e-2, -4.0))
    s1 = ico.source_2lpt(ico.phi_second_derivatives(d))
    s2 = ico.source_2lpt(ico.phi_second_derivatives(2.0 * d))
    np.testing.assert_allclose(s2, 4.0 * s1, rtol=1e-12, atol=1e-30)
    p1 = ico.source_to_psi2(s1, 1.0)
    p2 = ico.source_to_psi2(s2, 1.0)
    for a in range(3):
        assert np.all(np.isfinite(p1[a]))
        np.testing.assert_allclose(p2[a], 4.0 * p1[a], rtol=1e-10, atol=1e-12 * np.abs(p2[a]).max())


def test_2lpt_factors():
    v1, v2, d2 = ico.factors_2lpt(EDS, 0.5, True)
    assert d2 == -3.0 / 7.0
    assert v1 == pytest.approx(0.25 * 0.7 * 0.5 ** -1.5, rel=1e-14)
    assert v2 == pytest.approx(0.25 * 2.0 * 0.7 * 0.5 ** -1.5 * (-3.0 / 7.0), rel=1e-14)
    v1, v2, d2 = ico.factors_2lpt(LCDM, 0.02, True)
    assert -0.44 < d2 < -0.42  # Bouchet et al.: about -0.428 at Omega_m(a) near 1
    from gadget_ng_pm_hip import ic
    cp = gi.CosmologyParams(*LCDM[:3])
    assert ic.factors_2lpt(cp, 0.02, 0.1, False)[:3] == (v1, v2, d2)
    assert ic.factors_2lpt(None, 0.3, 0.7, False)[:3] == ico.factors_2lpt((1.0, 0.0, 0.7), 0.3, False)
