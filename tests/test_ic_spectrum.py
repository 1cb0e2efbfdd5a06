"""Host-side spectrum builders of the initial conditions
(gadget_ng_pm_hip.ic): the Eisenstein-Hu branch of build_spectrum_fn and the
sigma_8 normalisation (crates/gadget-ng-core/src/transfer_fn.rs).  No GPU.

Pinned by the reference's own tests (transfer_fn.rs:288-440: T -> 1 at low
k, its range at k = 0.1 h/Mpc, monotonic decrease, the top-hat window, the
amplitude reproducing sigma_8) and by scipy's adaptive quadrature of the
sigma^2 integral, independent of the restated trapezoid rule."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from gadget_ng_pm_hip import ic

P18 = ic.EisensteinHuParams()


def test_transfer_reference_checks():
    assert abs(ic.transfer_eh_nowiggle(1e-4, P18) - 1.0) < 0.01
    assert ic.transfer_eh_nowiggle(1e-3, P18) > 0.90
    t = ic.transfer_eh_nowiggle(0.1, P18)
    assert 0.05 < t < 0.30
    assert ic.transfer_eh_nowiggle(0.01, P18) > t > ic.transfer_eh_nowiggle(1.0, P18)
    for e in range(-4, 4):
        assert 0.0 <= ic.transfer_eh_nowiggle(10.0 ** e, P18) <= 1.0
    assert ic.transfer_eh_nowiggle(0.0, P18) == 1.0


def test_tophat_window():
    assert abs(ic.tophat_window(0.0) - 1.0) < 1e-12
    assert ic.tophat_window(1.0) < ic.tophat_window(0.1)
    assert ic.tophat_window(10.0) < ic.tophat_window(1.0)
    # the series branch joins the closed form
    assert ic.tophat_window(0.99e-4) == pytest.approx(3 * (math.sin(1.01e-4) - 1.01e-4 * math.cos(1.01e-4))
                                                      / 1.01e-4 ** 3, abs=1e-8)


@pytest.mark.parametrize("r,ns", [(8.0, 0.965), (4.0, 1.0), (16.0, 0.9)])
def test_sigma_integral_against_quadrature(r, ns):
    def f(lk):
        k = math.exp(lk)
        t = ic.transfer_eh_nowiggle(k, P18)
        w = ic.tophat_window(k * r)
        return k ** (ns + 3.0) * t * t * w * w
    ref = quad(f, math.log(1e-5), math.log(5e2), limit=500, epsrel=1e-10)[0] / (2 * math.pi ** 2)
    # the restated trapezoid rule (8192 steps in ln k) is good to ~1e-4 once the
    # window oscillates (larger R); a wrong power or factor is off by far more
    assert ic.sigma_sq_unit(r, ns, P18) == pytest.approx(ref, rel=2e-4)


def test_amplitude_reproduces_sigma8():
    a = ic.amplitude_for_sigma8(0.8, 0.965, P18)
    assert a * math.sqrt(ic.sigma_sq_unit(8.0, 0.965, P18)) == pytest.approx(0.8, rel=1e-12)


def test_eisenstein_hu_spectrum_closure():
    n, box = 64, 200.0
    f = ic.eisenstein_hu_sigma(n, 0.965, 1.0, P18, sigma8=0.8, box_size_mpc_h=box)
    a = ic.amplitude_for_sigma8(0.8, 0.965, P18)
    assert f(0.0) == 0.0
    k = 2 * math.pi * P18.h / box * 5.0
    assert f(5.0) == pytest.approx(a * k ** (0.965 / 2) * ic.transfer_eh_nowiggle(k, P18) / math.sqrt(n ** 3),
                                   rel=1e-14)
    tab = ic.sigma_table(n, f)
    assert tab.shape == (ic.table_len(n),) and tab[0] == 0.0 and np.all(tab[1:] > 0.0)
    # without sigma8 the amplitude is used as given; the box defaults to 100 Mpc/h
    g = ic.eisenstein_hu_sigma(n, 1.0, 2.0, P18)
    k1 = 2 * math.pi * P18.h / 100.0
    assert g(1.0) == pytest.approx(2.0 * k1 ** 0.5 * ic.transfer_eh_nowiggle(k1, P18) / math.sqrt(n ** 3), rel=1e-14)


def _table_rows():
    ks = [1e-3, 1.5e-3, 2e-3, 3e-3, 5e-3, 7e-3, 1e-2, 1.5e-2, 2e-2, 3e-2, 5e-2, 7e-2, 0.1, 0.15, 0.2, 0.3, 0.5,
          0.7, 1.0]
    tk = lambda k: (1.0 + 0.02 * math.sin(1.5 * math.log(k))) * (1.0 + 20.0 * k) ** -0.8
    return ks, tk


def test_tabulated_transfer_knots_and_midpoints(tmp_path):
    """The reference's own test (ic_zeldovich.rs tabulated_transfer_reconstructs_knots_and_midpoints):
    knots to 1e-7 through the spectrum closure, midpoints to 0.1 %."""
    ks, tk = _table_rows()
    path = tmp_path / "tf.dat"
    path.write_text("# k[h/Mpc] T(k)\n" + "".join(f"{k:.8e} {tk(k):.8e}\n" for k in ks))
    tab = ic.TabulatedTransfer.from_file(str(path))
    n, h = 32, 0.674
    spec = ic.tabulated_sigma(n, 0.0, 1.0, tab, h, box_size_mpc_h=1.0)
    inv = 1.0 / math.sqrt(n ** 3)
    for k in ks:
        est = spec(k / (2 * math.pi * h)) / inv
        assert abs(est - tk(k)) / tk(k) < 1e-7
    for a, b in zip(ks[:-1], ks[1:]):
        km = math.sqrt(a * b)
        assert abs(tab(km) - tk(km)) / tk(km) < 1e-3


def test_tabulated_transfer_against_scipy_pchip():
    """scipy's PchipInterpolator uses the same Fritsch-Carlson interior slopes
    and shape-preserving end slopes: an independent implementation inside the
    table; outside, the extrapolation is linear in log space."""
    from scipy.interpolate import PchipInterpolator
    g = np.random.default_rng(5)
    k = np.sort(g.uniform(-3, 1, 25))
    t = np.cumsum(g.normal(0, 0.3, 25))  # non-monotone ln T: exercises the zero-slope rule
    tab = ic.TabulatedTransfer(zip(np.exp(k), np.exp(t)))
    ref = PchipInterpolator(k, t)
    for xq in np.linspace(k[0], k[-1], 400):
        assert math.log(tab(math.exp(xq))) == pytest.approx(float(ref(xq)), abs=1e-9)
    lo = k[0] - 0.5
    assert math.log(tab(math.exp(lo))) == pytest.approx(t[0] + tab.m[0] * (lo - k[0]), abs=1e-12)
    assert tab(0.0) == 0.0 and tab(-1.0) == 0.0
    with pytest.raises(ValueError):
        ic.TabulatedTransfer([(1.0, 1.0)])


@pytest.mark.parametrize("m", [0.5, 2.0, 7.0])
def test_dark_matter_half_modes(m):
    """At its half-mode wavenumber each suppression halves the power (T^2 = 1/2),
    it is 1 at k -> 0 and falls monotonically."""
    om, h = 0.3, 0.7
    kw = ic.wdm_half_mode_k(m, om, h)
    assert ic.wdm_transfer_suppression(kw, m, om, h) ** 2 == pytest.approx(0.5, rel=1e-12)
    kf = ic.fdm_half_mode_k(m, om, h)
    assert ic.fdm_transfer_suppression(kf, m, om, h) ** 2 == pytest.approx(0.5, rel=1e-12)
    for fn, k0 in ((ic.wdm_transfer_suppression, kw), (ic.fdm_transfer_suppression, kf)):
        assert fn(1e-6 * k0, m, om, h) == pytest.approx(1.0, abs=1e-6)
        vals = [fn(k0 * f, m, om, h) for f in (0.1, 0.5, 1.0, 2.0, 10.0)]
        assert all(a > b for a, b in zip(vals, vals[1:]))
    assert ic.wdm_transfer_suppression(1.0, 0.0, om, h) == 1.0 and ic.fdm_half_mode_k(0.0, om, h) == math.inf


def test_dark_matter_cutoff_wraps_the_spectrum():
    n, box = 64, 50.0
    base = ic.power_law_sigma(n, 1.0, -2.0)
    f = ic.with_dark_matter_cutoff(base, "warm", 1.0, 0.3, 0.7, box_size_mpc_h=box)
    k = 2 * math.pi * 0.7 / box * 7.0
    assert f(7.0) == pytest.approx(base(7.0) * ic.wdm_transfer_suppression(k, 1.0, 0.3, 0.7), rel=1e-15)
    assert ic.with_dark_matter_cutoff(base, "cold", 1.0, 0.3, 0.7) is base
    with pytest.raises(ValueError):
        ic.with_dark_matter_cutoff(base, "hot", 1.0, 0.3, 0.7)
