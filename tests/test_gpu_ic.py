"""Zel'dovich initial conditions on the GPU (include/gng_ic.h) against the
FP64 initial-condition oracle (oracle/ic_oracle.py).

* The FP64 modes: the integer path (mode seeds, PCG32 key expansion, ChaCha12
  words) must match exactly, which leaves only the last-ulp rounding of log
  and cos: |delta_gpu - delta_oracle| <= 1e-13 max |delta|.  Any slip in the
  integer path would change the Gaussians completely.
* The displacements go through FP32 transforms: tolerance 1e-5 rms relative
  (and 1e-4 of the largest component for the worst particle); positions to
  the same absolute error, momenta with the same relative error.
* The reference's own IC tests (ic_zeldovich.rs:1019-1105) on the GPU output.
"""
import math

import numpy as np
import pytest

import ic_oracle as ico
from helpers import rms_rel

pytestmark = pytest.mark.gpu

LCDM = (0.3, 0.7, 0.1, -1.0, 0.0, 0.0)


def _oracle_field(n, seed, sig):
    return ico.delta_kspace(n, seed, lambda n2: sig[np.asarray(n2)])


@pytest.mark.parametrize("n,seed", [(4, 1), (8, 42), (8, 99), (12, 5), (16, 1234), (24, 77), (32, 2 ** 64 - 1)])
def test_modes_match_the_oracle(pm, n, seed):
    from gadget_ng_pm_hip import ic
    print(f"case ic modes n={n} seed={seed}", flush=True)
    plan = pm.Plan(n, 10.0)
    sig = ic.sigma_table(n, ic.power_law_sigma(n, 1e-4, -2.0))
    got = ic.delta_kspace(plan, seed, sig)
    ref = _oracle_field(n, seed, sig)[:, :, : n // 2 + 1]
    assert got.shape == ref.shape
    assert np.array_equal(got == 0, ref == 0)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()
    # the reference's own tests: DC zero, Hermitian symmetry in the iz = 0 plane
    assert got[0, 0, 0] == 0
    neg = (-np.arange(n)) % n
    np.testing.assert_array_equal(got[:, :, 0], np.conj(got[neg][:, neg, 0]))


def _amplitude_for(n, box, ns, target_cells=0.3):
    """Power-law amplitude whose expected 3-D rms displacement is target_cells
    cells: E|Psi-hat|^2 = 2 sigma^2 / |n|^2 per live mode, Psi = d IFFT."""
    f = ico.mode_int(np.arange(n), n).astype(np.float64)
    n2 = f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2
    h = n // 2
    a = np.abs(f)
    live = (n2 > 0) & (a[:, None, None] != h) & (a[None, :, None] != h) & (a[None, None, :] != h)
    s2 = np.sqrt(n2[live]) ** (ns / 2.0) / n ** 3  # sigma^2 at A = 1
    rms2 = (2.0 * s2 / n2[live]).sum() * (box / n) ** 2 / n ** 6
    return target_cells * (box / n) / math.sqrt(rms2)


@pytest.mark.parametrize("n,box,seed,ns", [(8, 5.0, 3, -2.0), (16, 100.0, 11, -1.0), (24, 48.0, 5, -2.5),
                                           (32, 64.0, 2, -2.0), (64, 256.0, 9, -3.0)])
def test_particles_match_the_oracle(pm, n, box, seed, ns):
    from gadget_ng_pm_hip import ic
    from gadget_ng_pm_hip import integrate as gi
    print(f"case ic particles n={n} box={box} seed={seed} n_s={ns}", flush=True)
    amp = _amplitude_for(n, box, ns)
    plan = pm.Plan(n, box)
    cosmo = gi.CosmologyParams(*LCDM)
    pos, vel, mass = ic.zeldovich_ics_power_law(plan, seed, amp, ns, cosmology=cosmo, a_init=0.02,
                                                rescale_to_a_init=True)
    assert pos.shape == (n ** 3, 3) and vel.shape == (n ** 3, 3)
    assert np.all(mass == 1.0 / n ** 3)
    vf = 0.02 ** 2 * ico.growth_rate_f(LCDM, 0.02) * ico.hubble_param(LCDM, 0.02)
    sc = ico.growth_factor_d(LCDM, 0.02) / ico.growth_factor_d(LCDM, 1.0)
    assert ic.momentum_factors(cosmo, 0.02, 0.1, True) == (vf, sc, 1.0)
    sig = ic.sigma_table(n, ic.power_law_sigma(n, amp, ns))
    psi = ico.displacement(_oracle_field(n, seed, sig), box)
    rp, rv = ico.zeldovich(n, box, psi, vf, sc)
    assert rms_rel(vel, rv) < 1e-5
    top = np.abs(rv).max()
    assert np.abs(vel - rv).max() < 1e-4 * top
    dp = (pos - rp + box / 2) % box - box / 2
    assert np.abs(dp).max() < 1e-4 * top / vf + 4 * np.finfo(np.float64).eps * box
    assert pos.min() >= 0.0 and pos.max() < box
    # the lattice: momenta / vel_factor is the displacement from q
    g = np.arange(n ** 3)
    q = np.stack([g // (n * n), (g // n) % n, g % n], axis=1) * (box / n) + 0.5 * box / n
    moved = (pos - q + box / 2) % box - box / 2
    np.testing.assert_allclose(moved, vel / vf, rtol=0, atol=1e-9 * box)


def test_subranges_and_the_device_entry_are_bit_identical(pm):
    from gadget_ng_pm_hip import ic
    n, box, seed = 32, 32.0, 21
    print(f"case ic subranges n={n}", flush=True)
    plan = pm.Plan(n, box)
    fn = ic.power_law_sigma(n, 0.05, -2.0)
    pos, vel, _ = ic.zeldovich_ics(plan, seed, fn, h0=0.1)
    for lo, hi in ((0, 1), (100, 5000), (n ** 3 - 7, n ** 3), (300, 300)):
        p, v, _ = ic.zeldovich_ics(plan, seed, fn, h0=0.1, lo=lo, hi=hi)
        assert np.array_equal(p, pos[lo:hi]) and np.array_equal(v, vel[lo:hi])
    # fewer particles than the lattice: the reference stops at particle_count
    p, v, m = ic.zeldovich_ics(plan, seed, fn, h0=0.1, particle_count=1000)
    assert p.shape == (1000, 3) and np.array_equal(p, pos[:1000]) and np.all(m == 1e-3)
    dp = pm.DeviceArray((n ** 3, 3), np.float64)
    dv = pm.DeviceArray((n ** 3, 3), np.float64)
    vf, sc, _ = ic.momentum_factors(None, 1.0, 0.1, False)
    assert vf == 0.1 and sc == 1.0
    ic.zeldovich_ics_device(plan, seed, ic.sigma_table(n, fn), vf, sc, dp.ptr, dv.ptr)
    pm.device_synchronize()
    assert np.array_equal(dp.numpy(), pos) and np.array_equal(dv.numpy(), vel)


def test_particles_trace_the_input_field(pm):
    """Small displacements: the particles' CIC density contrast correlates with
    the Gaussian field the displacements came from (delta = -div Psi to first
    order); a sign or axis error would anticorrelate or decorrelate them."""
    import analysis_oracle as ao
    from gadget_ng_pm_hip import ic
    n, box, seed = 64, 64.0, 4
    print(f"case ic density correlation n={n}", flush=True)
    sig_unit = ic.sigma_table(n, ic.power_law_sigma(n, 1.0, -4.0))
    amp = _amplitude_for(n, box, -4.0, target_cells=0.05)
    plan = pm.Plan(n, box)
    pos, _, mass = ic.zeldovich_ics_power_law(plan, seed, amp, -4.0)
    grid = ao.cic_grid(pos, mass, box, n)
    dens = grid / grid.mean() - 1.0
    field = np.fft.ifftn(_oracle_field(n, seed, sig_unit * amp)).real
    # CIC puts lattice point j + 1/2 on nodes j and j + 1: compare with the
    # field averaged onto the nodes the same way
    f_nodes = field
    for ax in range(3):
        f_nodes = 0.5 * (f_nodes + np.roll(f_nodes, 1, axis=ax))
    r = np.corrcoef(dens.ravel(), f_nodes.ravel())[0, 1]
    print(f"density/field correlation {r:.4f}", flush=True)
    assert r > 0.95


def test_full_size_invariants(pm):
    """256^3 particles on a 256^3 field: positions in the box, zero mean
    displacement, momenta equal to the displacement times vel_factor, and the
    power spectrum of the particles following the input spectrum at low k."""
    from gadget_ng_pm_hip import analysis, ic
    from gadget_ng_pm_hip import integrate as gi
    n, box, seed = 256, 256.0, 0x6E670005
    print(f"case ic full size n={n}", flush=True)
    plan = pm.Plan(n, box)
    cosmo = gi.CosmologyParams(*LCDM)
    amp = _amplitude_for(n, box, -2.0, target_cells=0.25)
    pos, vel, mass = ic.zeldovich_ics_power_law(plan, seed, amp, -2.0, cosmology=cosmo, a_init=0.05)
    vf, _, _ = ic.momentum_factors(cosmo, 0.05, 0.1, False)
    assert pos.min() >= 0.0 and pos.max() < box
    psi = vel / vf
    rms = np.sqrt((psi ** 2).sum(axis=1).mean())
    print(f"rms displacement {rms:.4f} cells", flush=True)
    assert 0.2 * box / n < rms < 0.3 * box / n
    assert np.abs(psi.mean(axis=0)).max() < 1e-5 * rms
    g = np.arange(n ** 3, dtype=np.int64)
    q = np.stack([g // (n * n), (g // n) % n, g % n], axis=1) * (box / n) + 0.5 * box / n
    del g
    moved = (pos - q + box / 2) % box - box / 2
    del q
    assert np.abs(moved - psi).max() < 1e-9 * box
    # P(k) of the particles against the input spectrum.  To first order the
    # particles' density contrast is -div Psi = (2 pi / N) IFFT(delta-hat)
    # (the displacement is i n delta-hat / |n|^2 in grid units times d), and
    # E|delta-hat|^2 = 2 sigma^2, so the estimator (|delta_k|^2 (V/N^3)^2)
    # should read 2 sigma^2 (2 pi / N)^2 (V / N^3)^2 up to cosmic variance.
    k, pk, nmodes = analysis.power_spectrum_arrays(plan, pos, mass)
    k_fund = 2.0 * math.pi / box
    sig = ic.power_law_sigma(n, amp, -2.0)
    sel = (k / k_fund >= 4.0) & (k / k_fund <= 20.0)
    want = np.array([2.0 * sig(x) ** 2 * (2.0 * math.pi / n) ** 2 * (box ** 3 / n ** 3) ** 2 for x in k[sel] / k_fund])
    ratios = pk[sel] / want
    print("P ratios, shells 4..20: " + " ".join(f"{r:.3f}" for r in ratios), flush=True)
    assert np.all((ratios > 0.75) & (ratios < 1.25))
    mean = float((ratios * nmodes[sel]).sum() / nmodes[sel].sum())
    assert 0.9 < mean < 1.1


@pytest.mark.parametrize("n,box,seed,ns,resc", [(8, 8.0, 3, -2.0, False), (16, 100.0, 11, -1.5, True),
                                                (24, 48.0, 5, -2.5, True), (32, 64.0, 2, -2.0, False)])
def test_2lpt_particles_match_the_oracle(pm, n, box, seed, ns, resc):
    """Second order: the first-order part as before; the second-order term
    (the difference from the first-order particles) within 1e-4 of its own
    size -- its source is a difference of f32 products."""
    from gadget_ng_pm_hip import ic
    from gadget_ng_pm_hip import integrate as gi
    print(f"case ic 2lpt n={n} box={box} seed={seed} n_s={ns} rescale={resc}", flush=True)
    amp = _amplitude_for(n, box, ns, target_cells=0.8)
    plan = pm.Plan(n, box)
    cosmo = gi.CosmologyParams(*LCDM[:3])
    fn = ic.power_law_sigma(n, amp, ns)
    pos, vel, _ = ic.zeldovich_2lpt_ics(plan, seed, fn, cosmology=cosmo, a_init=0.02, rescale_to_a_init=resc)
    v1, v2, d2, sc = ic.factors_2lpt(cosmo, 0.02, 0.1, resc)
    assert (v1, v2, d2) == ico.factors_2lpt(LCDM, 0.02, True)
    sig = ic.sigma_table(n, fn)
    delta = _oracle_field(n, seed, sig)
    psi1 = ico.displacement(delta, box)
    psi2 = ico.source_to_psi2(ico.source_2lpt(ico.phi_second_derivatives(delta)), box)
    rp, rv = ico.zeldovich_2lpt(n, box, psi1, psi2, v1, v2, d2, sc)
    # split the momenta into the two orders with the first-order oracle
    _, rv1 = ico.zeldovich(n, box, psi1, v1, sc)
    sec_ref = rv - rv1
    sec_got = vel - rv1
    print(f"second-order share of the momenta {np.abs(sec_ref).max() / np.abs(rv).max():.3e}", flush=True)
    assert np.abs(sec_ref).max() > 1e-3 * np.abs(rv).max()
    assert rms_rel(sec_got, sec_ref) < 1e-4
    assert rms_rel(vel, rv) < 2e-5
    dp = (pos - rp + box / 2) % box - box / 2
    assert np.abs(dp).max() < 1e-4 * (np.abs(rv1).max() / v1 + abs(d2) * np.abs(sec_ref).max() / abs(v2))
    assert pos.min() >= 0.0 and pos.max() < box
    # sub-range and device entry
    p2, v2b, _ = ic.zeldovich_2lpt_ics(plan, seed, fn, cosmology=cosmo, a_init=0.02, rescale_to_a_init=resc,
                                       lo=7, hi=n ** 3 - 3)
    assert np.array_equal(p2, pos[7:n ** 3 - 3]) and np.array_equal(v2b, vel[7:n ** 3 - 3])
    dpos = pm.DeviceArray((n ** 3, 3), np.float64)
    dvel = pm.DeviceArray((n ** 3, 3), np.float64)
    ic.zeldovich_2lpt_ics_device(plan, seed, sig, v1, v2, d2, sc, dpos.ptr, dvel.ptr)
    assert np.array_equal(dpos.numpy(), pos) and np.array_equal(dvel.numpy(), vel)


def test_bad_arguments(pm):
    from gadget_ng_pm_hip import ic
    plan = pm.Plan(16, 10.0)
    sig = ic.sigma_table(16, ic.power_law_sigma(16, 1.0, -2.0))
    with pytest.raises(Exception):
        ic.delta_kspace(plan, 1, sig[:-1])  # table too short
    with pytest.raises(Exception):
        ic.zeldovich_ics(plan, 1, ic.power_law_sigma(16, 1.0, -2.0), lo=0, hi=16 ** 3 + 1, particle_count=16 ** 3 + 1)
    with pytest.raises(Exception):
        ic.zeldovich_ics(plan, 1, ic.power_law_sigma(16, 1.0, -2.0), m_nu_ev=0.06)
    p, v, m = ic.zeldovich_ics(plan, 1, ic.power_law_sigma(16, 1.0, -2.0), lo=5, hi=5)
    assert p.shape == (0, 3)
