"""Zel'dovich initial conditions on the GPU (include/gng_ic.h).

Mirrors crates/gadget-ng-core/src/ic_zeldovich.rs on top of libgng_pm.so:

* :func:`power_law_sigma`  -- ``build_spectrum_fn`` with ``TransferKind::PowerLaw``
  (:404-415), tabulated at sqrt(n2) for the integer |n|^2 of every mode;
* :func:`eisenstein_hu_sigma` -- the ``EisensteinHu`` branch (:417-449) with
  crates/gadget-ng-core/src/transfer_fn.rs (no-wiggle transfer function,
  sigma_8 normalisation);
* :func:`tabulated_sigma`, :class:`TabulatedTransfer` -- the ``Tabulated``
  branch (:450-460) and its PCHIP table (:118-260);
* :func:`with_dark_matter_cutoff` -- (:468-489) with the warm and fuzzy
  suppressions of crates/gadget-ng-core/src/dark_matter.rs;
* :func:`delta_kspace`     -- ``generate_delta_kspace`` (:287-369), FP64 modes
  iz = 0 .. n/2;
* :func:`zeldovich_ics`    -- ``zeldovich_ics`` (:621-766): Gaussian field,
  displacement, lattice particles, canonical momenta a^2 f(a) H(a) Psi, the
  optional rescaling D(a_init) / D(1), the neutrino suppression of the
  spectrum (:679-684, :697-709);
* :func:`zeldovich_ics_power_law` -- (:953-980);
* :func:`zeldovich_2lpt_ics`  -- ``zeldovich_2lpt_ics`` of
  crates/gadget-ng-core/src/ic_2lpt.rs:348-553 (second-order displacement
  from the six second derivatives of the first-order potential).

The plan's grid_size is the lattice side and its box_size the box; the
particles are the global ids lo .. hi - 1 of the lattice.  The spectrum is
any callable of |n| (the reference's spectrum closure); it is evaluated on
the host once per integer |n|^2 and the table goes to the device.  Everything
else runs in the HIP library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Callable, Optional, Tuple

import numpy as np

from . import GNG_OK, HipExecutionError, Plan, _ptr, last_error, lib
from .integrate import CosmologyParams


def _check(rc: int, what: str) -> None:
    if rc != GNG_OK:
        raise HipExecutionError(what, rc, last_error())


def table_len(n: int) -> int:
    """Entries the library needs: every |n|^2 of a mode without a Nyquist
    component, 0 .. 3 (n/2 - 1)^2."""
    return 3 * (n // 2 - 1) ** 2 + 1


def sigma_table(n: int, spectrum_fn: Callable[[float], float]) -> np.ndarray:
    """sigma[n2] = spectrum_fn(sqrt(n2)), as the reference evaluates its
    closure per mode at (n2 as f64).sqrt() (ic_zeldovich.rs:315-316)."""
    return np.array([spectrum_fn(math.sqrt(float(v))) for v in range(table_len(n))], dtype=np.float64)


def power_law_sigma(n: int, amplitude: float, spectral_index: float) -> Callable[[float], float]:
    """The PowerLaw spectrum closure: A |n|^(n_s/4) / sqrt(n^3), 0 at |n| = 0."""
    inv_sqrt_n3 = 1.0 / math.sqrt(float(n * n * n))

    def f(n_abs: float) -> float:
        if n_abs <= 0.0:
            return 0.0
        return amplitude * math.pow(n_abs, spectral_index / 4.0) * inv_sqrt_n3
    return f


@dataclass(frozen=True)
class EisensteinHuParams:
    """crates/gadget-ng-core/src/transfer_fn.rs:39-60 (Planck 2018 defaults)."""
    omega_m: float = 0.315
    omega_b: float = 0.049
    h: float = 0.674
    t_cmb: float = 2.7255


def transfer_eh_nowiggle(k_hmpc: float, p: EisensteinHuParams) -> float:
    """The reference's no-wiggle Eisenstein-Hu transfer function
    (transfer_fn.rs:89-114), term by term: sound horizon s, alpha_Gamma,
    Gamma_eff with k s in Mpc, q = k / Gamma_eff (k in h/Mpc; the reference
    leaves out the CMB-temperature factor of the paper), T = L0 / (L0 + C0 q^2),
    clamped to [0, 1]."""
    if k_hmpc <= 0.0:
        return 1.0
    om_h2 = p.omega_m * p.h * p.h
    ob_h2 = p.omega_b * p.h * p.h
    fb = p.omega_b / p.omega_m
    s = 44.5 * math.log(9.83 / om_h2) / math.sqrt(1.0 + 10.0 * math.pow(ob_h2, 0.75))
    alpha = 1.0 - 0.328 * math.log(431.0 * om_h2) * fb + 0.38 * math.log(22.3 * om_h2) * fb * fb
    ks = 0.43 * (k_hmpc * p.h) * s
    gamma_eff = p.omega_m * p.h * (alpha + (1.0 - alpha) / (1.0 + ks ** 4))
    q = k_hmpc / max(gamma_eff, 1e-30)
    l0 = math.log(math.e * 2.0 + 1.8 * q)
    c0 = 14.2 + 731.0 / (1.0 + 62.5 * q)
    t = l0 / (l0 + c0 * q * q)
    return min(max(t, 0.0), 1.0)


def tophat_window(x: float) -> float:
    """3 (sin x - x cos x) / x^3, 1 - x^2/10 below 1e-4 (transfer_fn.rs)."""
    if abs(x) < 1e-4:
        return 1.0 - x * x / 10.0
    return 3.0 * (math.sin(x) - x * math.cos(x)) / (x * x * x)


def sigma_sq_unit(r_mpc_h: float, n_s: float, p: EisensteinHuParams) -> float:
    """sigma^2(R) for unit amplitude: trapezoid rule in ln k over 8192 steps
    from 1e-5 to 500 of k^(n_s+3) T^2 W^2 / (2 pi^2) (transfer_fn.rs:150-183)."""
    lo, hi, steps = math.log(1e-5), math.log(5e2), 8192
    dl = (hi - lo) / steps

    def f(k):
        t = transfer_eh_nowiggle(k, p)
        w = tophat_window(k * r_mpc_h)
        return math.pow(k, n_s + 3.0) * t * t * w * w
    total = 0.0
    fp = f(1e-5)
    for i in range(1, steps + 1):
        fk = f(math.exp(lo + i * dl))
        total += 0.5 * (fp + fk) * dl
        fp = fk
    return total / (2.0 * math.pi * math.pi)


def amplitude_for_sigma8(sigma8: float, n_s: float, p: EisensteinHuParams) -> float:
    """A with sigma(8 Mpc/h) = sigma8 (transfer_fn.rs:222-228)."""
    s2 = sigma_sq_unit(8.0, n_s, p)
    return sigma8 if s2 <= 0.0 else sigma8 / math.sqrt(s2)


def eisenstein_hu_sigma(n: int, spectral_index: float, amplitude: float, p: EisensteinHuParams,
                        sigma8: Optional[float] = None,
                        box_size_mpc_h: Optional[float] = None) -> Callable[[float], float]:
    """build_spectrum_fn, EisensteinHu branch (ic_zeldovich.rs:417-449):
    A k^(n_s/2) T(k) / sqrt(n^3) with k = 2 pi h |n| / L (L defaults to
    100 Mpc/h) and A from sigma8 when given."""
    inv_sqrt_n3 = 1.0 / math.sqrt(float(n * n * n))
    bsm = 100.0 if box_size_mpc_h is None else box_size_mpc_h
    amp = amplitude_for_sigma8(sigma8, spectral_index, p) if sigma8 is not None else amplitude
    c = 2.0 * math.pi * p.h / bsm

    def f(n_abs: float) -> float:
        if n_abs <= 0.0:
            return 0.0
        k = c * n_abs
        return amp * math.pow(k, spectral_index / 2.0) * transfer_eh_nowiggle(k, p) * inv_sqrt_n3
    return f


class TabulatedTransfer:
    """A transfer function read from a text table of (k [h/Mpc], T) rows,
    interpolated by monotone piecewise-cubic Hermite (PCHIP, Fritsch-Carlson
    slopes) in (ln k, ln T), extrapolated linearly in log space with the end
    slopes (ic_zeldovich.rs:118-260).  Rows: the first two numbers of a line,
    separated by spaces, commas or semicolons; '#' lines and rows without
    k > 0 and T > 0 are skipped; duplicate k keep the first row."""

    def __init__(self, rows):
        rows = sorted(((float(k), float(t)) for k, t in rows), key=lambda r: r[0])
        kept = []
        for k, t in rows:
            if not kept or abs(k - kept[-1][0]) > np.finfo(np.float64).eps:
                kept.append((k, t))
        if len(kept) < 2:
            raise ValueError("a tabulated transfer function needs at least 2 valid rows (k > 0, T > 0)")
        self.x = [math.log(k) for k, _ in kept]
        self.y = [math.log(t) for _, t in kept]
        self.m = self._slopes(self.x, self.y)

    @classmethod
    def from_file(cls, path: str) -> "TabulatedTransfer":
        rows = []
        with open(path) as f:
            for raw in f:
                line = raw.strip()
                if not line or line.startswith("#"):
                    continue
                vals = []
                for tok in line.replace(",", " ").replace(";", " ").split():
                    try:
                        vals.append(float(tok))
                    except ValueError:
                        continue
                    if len(vals) >= 2:
                        break
                if len(vals) >= 2 and all(math.isfinite(v) and v > 0.0 for v in vals[:2]):
                    rows.append((vals[0], vals[1]))
        return cls(rows)

    @staticmethod
    def _slopes(x, y):
        n = len(x)
        h = [x[i + 1] - x[i] for i in range(n - 1)]
        d = [(y[i + 1] - y[i]) / h[i] for i in range(n - 1)]
        if n == 2:
            return [d[0], d[0]]
        sgn = lambda v: math.copysign(1.0, v)
        m = [0.0] * n
        m[0] = ((2.0 * h[0] + h[1]) * d[0] - h[0] * d[1]) / (h[0] + h[1])
        if sgn(m[0]) != sgn(d[0]):
            m[0] = 0.0
        elif sgn(d[0]) != sgn(d[1]) and abs(m[0]) > 3.0 * abs(d[0]):
            m[0] = 3.0 * d[0]
        m[-1] = ((2.0 * h[-1] + h[-2]) * d[-1] - h[-1] * d[-2]) / (h[-1] + h[-2])
        if sgn(m[-1]) != sgn(d[-1]):
            m[-1] = 0.0
        elif sgn(d[-1]) != sgn(d[-2]) and abs(m[-1]) > 3.0 * abs(d[-1]):
            m[-1] = 3.0 * d[-1]
        for i in range(1, n - 1):
            if d[i - 1] == 0.0 or d[i] == 0.0 or sgn(d[i - 1]) != sgn(d[i]):
                m[i] = 0.0
            else:
                w1, w2 = 2.0 * h[i] + h[i - 1], h[i] + 2.0 * h[i - 1]
                m[i] = (w1 + w2) / (w1 / d[i - 1] + w2 / d[i])
        return m

    def __call__(self, k: float) -> float:
        if k <= 0.0 or not math.isfinite(k):
            return 0.0
        x, y, m = self.x, self.y, self.m
        xq = math.log(k)
        if xq <= x[0]:
            return math.exp(y[0] + m[0] * (xq - x[0]))
        if xq >= x[-1]:
            return math.exp(y[-1] + m[-1] * (xq - x[-1]))
        lo, hi = 0, len(x) - 1
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if x[mid] <= xq:
                lo = mid
            else:
                hi = mid
        h = x[lo + 1] - x[lo]
        t = (xq - x[lo]) / h
        t2, t3 = t * t, t * t * t
        yq = ((2.0 * t3 - 3.0 * t2 + 1.0) * y[lo] + (t3 - 2.0 * t2 + t) * h * m[lo]
              + (-2.0 * t3 + 3.0 * t2) * y[lo + 1] + (t3 - t2) * h * m[lo + 1])
        return math.exp(yq)


def tabulated_sigma(n: int, spectral_index: float, amplitude: float, transfer: TabulatedTransfer, h: float,
                    box_size_mpc_h: Optional[float] = None) -> Callable[[float], float]:
    """build_spectrum_fn, Tabulated branch (ic_zeldovich.rs:450-460):
    A k^(n_s/2) T(k) / sqrt(n^3) with k = 2 pi h |n| / L (L defaults to 100)."""
    inv_sqrt_n3 = 1.0 / math.sqrt(float(n * n * n))

    def f(n_abs: float) -> float:
        if n_abs <= 0.0:
            return 0.0
        bsm = 100.0 if box_size_mpc_h is None else box_size_mpc_h
        k = 2.0 * math.pi * h / bsm * n_abs
        return amplitude * math.pow(k, spectral_index / 2.0) * transfer(k) * inv_sqrt_n3
    return f


def wdm_transfer_suppression(k_hmpc: float, m_wdm_kev: float, omega_m: float, h: float) -> float:
    """Warm dark matter: (1 + (alpha k)^(2 nu))^(-5/nu), nu = 1.12 (Viel et al.
    2005 fit as in crates/gadget-ng-core/src/dark_matter.rs)."""
    if k_hmpc <= 0.0 or m_wdm_kev <= 0.0:
        return 1.0
    nu = 1.12
    alpha = (0.049 * math.pow(m_wdm_kev, -1.11) * math.pow(max(omega_m, 1e-30) / 0.25, 0.11)
             * math.pow(max(h, 1e-30) / 0.7, 1.22))
    return math.pow(1.0 + math.pow(alpha * k_hmpc, 2.0 * nu), -5.0 / nu)


def wdm_half_mode_k(m_wdm_kev: float, omega_m: float, h: float) -> float:
    """k where the WDM suppression of the power is 1/2."""
    if m_wdm_kev <= 0.0:
        return math.inf
    nu = 1.12
    alpha = (0.049 * math.pow(m_wdm_kev, -1.11) * math.pow(max(omega_m, 1e-30) / 0.25, 0.11)
             * math.pow(max(h, 1e-30) / 0.7, 1.22))
    return math.pow(math.pow(2.0, nu / 10.0) - 1.0, 1.0 / (2.0 * nu)) / alpha


def fdm_half_mode_k(m_fdm_22: float, omega_m: float, h: float) -> float:
    """Fuzzy dark matter half-mode wavenumber (dark_matter.rs)."""
    if m_fdm_22 <= 0.0:
        return math.inf
    return 9.1 * math.pow(m_fdm_22, 4.0 / 9.0) * math.pow(max(omega_m, 1e-30) / 0.27, 0.11) * math.pow(h / 0.7, 1.22)


def fdm_transfer_suppression(k_hmpc: float, m_fdm_22: float, omega_m: float, h: float) -> float:
    """Fuzzy dark matter: (1 + (k / k_half)^8)^(-1/2)."""
    if k_hmpc <= 0.0 or m_fdm_22 <= 0.0:
        return 1.0
    r = k_hmpc / fdm_half_mode_k(m_fdm_22, omega_m, h)
    return math.pow(1.0 + r ** 8, -0.5)


def with_dark_matter_cutoff(spectrum_fn: Callable[[float], float], model: str, mass: float, omega_m: float,
                            h: float, box_size_mpc_h: Optional[float] = None) -> Callable[[float], float]:
    """with_dark_matter_cutoff (ic_zeldovich.rs:468-489): the spectrum times
    the suppression at k = 2 pi h |n| / L; model 'cold' (no change), 'warm'
    (mass = m_wdm in keV) or 'fuzzy' (mass = m_fdm in 1e-22 eV)."""
    if model == "cold":
        return spectrum_fn
    if model not in ("warm", "fuzzy"):
        raise ValueError("model must be 'cold', 'warm' or 'fuzzy'")
    bsm = 100.0 if box_size_mpc_h is None else box_size_mpc_h
    sup = wdm_transfer_suppression if model == "warm" else fdm_transfer_suppression

    def f(n_abs: float) -> float:
        if n_abs <= 0.0:
            return 0.0
        k = 2.0 * math.pi * h / bsm * n_abs
        return spectrum_fn(n_abs) * sup(k, mass, omega_m, h)
    return f


def delta_kspace(plan: Plan, seed: int, sigma: np.ndarray) -> np.ndarray:
    """delta-hat[ix, iy, iz] for iz = 0 .. n/2 (complex128, shape (n, n, n/2 + 1))."""
    n = plan.grid_size
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    out = np.empty((n, n, n // 2 + 1, 2), dtype=np.float64)
    _check(lib().gng_ic_delta_kspace(plan._h, ctypes.c_uint64(int(seed) & (2 ** 64 - 1)), _ptr(sig), sig.size,
                                     _ptr(out)), "gng_ic_delta_kspace")
    return out[..., 0] + 1j * out[..., 1]


def momentum_factors(cosmology: Optional[CosmologyParams], a_init: float, h0: float,
                     rescale_to_a_init: bool, m_nu_ev: float = 0.0) -> Tuple[float, float, float]:
    """(vel_factor, scale, sqrt of the neutrino suppression) of zeldovich_ics
    (ic_zeldovich.rs:639-709): without cosmology a = 1 and an Einstein-de
    Sitter background with h0, so vel_factor = h0 and scale = 1."""
    if cosmology is None:
        a = 1.0
        cp = CosmologyParams(1.0, 0.0, h0)
        enabled = False
    else:
        a = a_init
        om_nu = m_nu_ev / (93.14 * (cosmology.h0 * 10.0) ** 2) if m_nu_ev > 0.0 else 0.0
        cp = CosmologyParams(cosmology.omega_m, cosmology.omega_lambda, cosmology.h0, cosmology.w0, cosmology.wa,
                             om_nu)
        enabled = True
    sup = 1.0
    if enabled and m_nu_ev > 0.0 and cosmology.omega_m > 0.0:
        f_nu = cp.omega_nu / cosmology.omega_m
        sup = min(max(1.0 - 8.0 * f_nu, 0.0), 1.0)
    vel_factor = a * a * cp.growth_rate(a) * cp.hubble(a)
    scale = 1.0
    if rescale_to_a_init and enabled:
        d_den = cp.growth_factor(1.0)
        scale = 1.0 if d_den == 0.0 else cp.growth_factor(a) / d_den
    return vel_factor, scale, math.sqrt(sup) if sup < 1.0 else 1.0


def zeldovich_ics(plan: Plan, seed: int, spectrum_fn: Callable[[float], float], *,
                  cosmology: Optional[CosmologyParams] = None, a_init: float = 1.0, h0: float = 0.1,
                  rescale_to_a_init: bool = False, m_nu_ev: float = 0.0, particle_count: Optional[int] = None,
                  lo: int = 0, hi: Optional[int] = None):
    """Particles lo .. hi - 1 as (pos, vel, mass): (k, 3) f64 each and the
    uniform mass 1 / particle_count (particle_count defaults to n^3; the
    reference stops at it, ic_zeldovich.rs:735-737)."""
    n = plan.grid_size
    n_part = n ** 3 if particle_count is None else int(particle_count)
    if n_part > n ** 3:
        raise ValueError("particle_count exceeds the lattice (grid_size^3)")
    hi = n_part if hi is None else min(int(hi), n_part)
    lo = min(int(lo), hi)
    if m_nu_ev > 0.0 and cosmology is None:
        raise ValueError("m_nu_ev needs the cosmology (its omega_m and h0)")
    vel_factor, scale, sqrt_sup = momentum_factors(cosmology, a_init, h0, rescale_to_a_init, m_nu_ev)
    fn = spectrum_fn if sqrt_sup == 1.0 else (lambda x: spectrum_fn(x) * sqrt_sup)
    sig = sigma_table(n, fn)
    pos = np.empty((hi - lo, 3), dtype=np.float64)
    vel = np.empty((hi - lo, 3), dtype=np.float64)
    _check(lib().gng_ic_zeldovich(plan._h, ctypes.c_uint64(int(seed) & (2 ** 64 - 1)), _ptr(sig), sig.size,
                                  float(vel_factor), float(scale), lo, hi, _ptr(pos), _ptr(vel)), "gng_ic_zeldovich")
    return pos, vel, np.full(hi - lo, 1.0 / n_part)


def zeldovich_ics_device(plan: Plan, seed: int, sigma: np.ndarray, vel_factor: float, scale: float,
                         d_pos: int, d_vel: int, lo: int = 0, hi: Optional[int] = None, stream: int = 0) -> None:
    """Same into device buffers (raw pointers of (hi - lo) x 3 f64 each),
    enqueued on `stream`; sigma is the table of :func:`sigma_table`."""
    n = plan.grid_size
    hi = n ** 3 if hi is None else int(hi)
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    _check(lib().gng_ic_zeldovich_device(plan._h, ctypes.c_uint64(int(seed) & (2 ** 64 - 1)), _ptr(sig), sig.size,
                                         float(vel_factor), float(scale), int(lo), hi, ctypes.c_void_p(d_pos),
                                         ctypes.c_void_p(d_vel), ctypes.c_void_p(stream)),
           "gng_ic_zeldovich_device")


def zeldovich_ics_power_law(plan: Plan, seed: int, amplitude: float, spectral_index: float, **kw):
    """zeldovich_ics with the PowerLaw spectrum and no rescaling (:953-980)."""
    kw.setdefault("rescale_to_a_init", False)
    return zeldovich_ics(plan, seed, power_law_sigma(plan.grid_size, amplitude, spectral_index), **kw)


def factors_2lpt(cosmology: Optional[CosmologyParams], a_init: float, h0: float,
                 rescale_to_a_init: bool) -> Tuple[float, float, float, float]:
    """(vel1, vel2, d2_over_d1sq, scale) of zeldovich_2lpt_ics
    (ic_2lpt.rs:414-453): its background is (omega_m, omega_lambda, h0) only;
    Omega_m(a) from H(a), f2 = 2 Omega_m(a)^(6/11), D2/D1^2 =
    -3/7 Omega_m(a)^(-1/143); without cosmology a = 1 and Omega_m(a) = 1."""
    enabled = cosmology is not None
    cp = CosmologyParams(cosmology.omega_m, cosmology.omega_lambda, cosmology.h0) if enabled else \
        CosmologyParams(1.0, 0.0, h0)
    a = a_init if enabled else 1.0
    h_a = cp.hubble(a)
    f1 = cp.growth_rate(a)
    if enabled and cp.h0 > 0.0:
        hr = (h_a / cp.h0) * (h_a / cp.h0)
        om_a = max(cp.omega_m / (a * a * a) / hr, 0.0)
    else:
        om_a = 1.0
    f2 = 2.0 * math.pow(om_a, 6.0 / 11.0)
    d2 = -3.0 / 7.0 * math.pow(om_a, -1.0 / 143.0)
    scale = 1.0
    if rescale_to_a_init and enabled:
        d_den = cp.growth_factor(1.0)
        scale = 1.0 if d_den == 0.0 else cp.growth_factor(a) / d_den
    return a * a * f1 * h_a, a * a * f2 * h_a * d2, d2, scale


def zeldovich_2lpt_ics(plan: Plan, seed: int, spectrum_fn: Callable[[float], float], *,
                       cosmology: Optional[CosmologyParams] = None, a_init: float = 1.0, h0: float = 0.1,
                       rescale_to_a_init: bool = False, particle_count: Optional[int] = None,
                       lo: int = 0, hi: Optional[int] = None):
    """Second-order particles lo .. hi - 1 as (pos, vel, mass), like
    :func:`zeldovich_ics`."""
    n = plan.grid_size
    n_part = n ** 3 if particle_count is None else int(particle_count)
    if n_part > n ** 3:
        raise ValueError("particle_count exceeds the lattice (grid_size^3)")
    hi = n_part if hi is None else min(int(hi), n_part)
    lo = min(int(lo), hi)
    vel1, vel2, d2, scale = factors_2lpt(cosmology, a_init, h0, rescale_to_a_init)
    sig = sigma_table(n, spectrum_fn)
    pos = np.empty((hi - lo, 3), dtype=np.float64)
    vel = np.empty((hi - lo, 3), dtype=np.float64)
    _check(lib().gng_ic_2lpt(plan._h, ctypes.c_uint64(int(seed) & (2 ** 64 - 1)), _ptr(sig), sig.size, float(vel1),
                             float(vel2), float(d2), float(scale), lo, hi, _ptr(pos), _ptr(vel)), "gng_ic_2lpt")
    return pos, vel, np.full(hi - lo, 1.0 / n_part)


def zeldovich_2lpt_ics_device(plan: Plan, seed: int, sigma: np.ndarray, vel1: float, vel2: float, d2: float,
                              scale: float, d_pos: int, d_vel: int, lo: int = 0, hi: Optional[int] = None,
                              stream: int = 0) -> None:
    """Same into device buffers ((hi - lo) x 3 f64 each); returns when done."""
    n = plan.grid_size
    hi = n ** 3 if hi is None else int(hi)
    sig = np.ascontiguousarray(sigma, dtype=np.float64)
    _check(lib().gng_ic_2lpt_device(plan._h, ctypes.c_uint64(int(seed) & (2 ** 64 - 1)), _ptr(sig), sig.size,
                                    float(vel1), float(vel2), float(d2), float(scale), int(lo), hi,
                                    ctypes.c_void_p(d_pos), ctypes.c_void_p(d_vel), ctypes.c_void_p(stream)),
           "gng_ic_2lpt_device")
