"""Device-resident KDK leapfrog (include/gng_integrate.h) against the FP64
integrator oracle.

Given the accelerations the GPU force produced, a step must reproduce the
oracle's f64 update bit for bit (the kernels round every product and sum
separately, as the reference does).  The PM force is a deterministic function
of the f64 positions (the integrator solves on them through the SoA f64 entry
point, which gives the host f64 entry point's result bit for bit), so the
oracle can take it from the same plan.  Over
several steps the trajectory is also compared with the oracle driven by the
FP64 CPU PM."""
import numpy as np
import pytest

import integrate_oracle as io
from helpers import f32_round

pytestmark = pytest.mark.gpu

LCDM = (0.3, 0.7, 0.1, -1.0, 0.0, 0.0)


def _box(n_side, box, seed, vscale=0.05):
    rng = np.random.default_rng(seed)
    g = (np.arange(n_side) + 0.5) * (box / n_side)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    pos = pos + rng.normal(scale=0.3 * box / n_side, size=pos.shape)
    pos = np.mod(pos, box)
    vel = rng.normal(scale=vscale * box / n_side, size=pos.shape)
    mass = np.full(pos.shape[0], 1.0 / pos.shape[0])
    return pos, vel, mass


def _gpu_force(plan, mass):
    m = np.asarray(mass, dtype=np.float64)

    def force(p, g):
        return plan.accelerations_f64(np.asarray(p, dtype=np.float64), m, g)
    return force


def test_one_flat_step_is_bit_identical_to_the_oracle(pm):
    from gadget_ng_pm_hip import integrate as gi
    box, nm = 32.0, 32
    pos, vel, mass = _box(16, box, 1)
    plan = pm.Plan(nm, box)
    print(f"case kdk flat step N={pos.shape[0]} nm={nm}", flush=True)
    lf = gi.Leapfrog(plan, pos.shape[0])
    lf.set_state(pos, vel, mass)
    dt, g = 0.3, 2.0
    lf.step(gi.CosmoFactors.flat(dt), g)
    p1, v1, a1 = lf.state()
    force = _gpu_force(plan, mass)
    rp, rv, ra = io.kdk_step(pos, vel, (dt, 0.5 * dt, 0.5 * dt), lambda p: force(p, g))
    assert np.array_equal(p1, rp)
    assert np.array_equal(v1, rv)
    assert np.array_equal(a1, ra)
    assert np.abs(rv - vel).max() > 0.0


def test_cosmological_run_with_wrap_is_bit_identical_to_the_oracle(pm):
    from gadget_ng_pm_hip import integrate as gi
    box, nm = 64.0, 32
    pos, vel, mass = _box(16, box, 2, vscale=0.6)
    # start some particles outside the box: the force wraps them, the step's end wraps the state
    pos[:40, 0] += box
    pos[40:80, 1] -= box
    plan = pm.Plan(nm, box)
    print(f"case kdk lcdm run N={pos.shape[0]} nm={nm}", flush=True)
    lf = gi.Leapfrog(plan, pos.shape[0])
    lf.set_state(pos, vel, mass)
    a_end = lf.run(3, 0.02, 1.5, cosmology=gi.CosmologyParams(*LCDM), a=0.25, wrap=True)
    p3, v3, a3 = lf.state()
    rp, rv, ra, ra_end = io.run(pos, vel, 3, 0.02, 1.5, _gpu_force(plan, mass), cosmology=LCDM, a=0.25, box=box)
    assert a_end == ra_end
    assert np.array_equal(p3, rp)
    assert np.array_equal(v3, rv)
    assert np.array_equal(a3, ra)
    assert p3.min() >= 0.0 and p3.max() < box


def test_run_equals_steps_with_host_factors(pm):
    from gadget_ng_pm_hip import integrate as gi
    box, nm = 48.0, 24
    pos, vel, mass = _box(12, box, 3)
    plan = pm.Plan(nm, box)
    print(f"case kdk run vs steps N={pos.shape[0]} nm={nm} (radix-3 mesh)", flush=True)
    cosmo = gi.CosmologyParams(*LCDM)
    a, b = gi.Leapfrog(plan, pos.shape[0]), gi.Leapfrog(plan, pos.shape[0])
    a.set_state(pos, vel, mass)
    b.set_state(pos, vel, mass)
    a_run = a.run(2, 0.05, 1.0, cosmology=cosmo, a=0.5, wrap=True)
    av = 0.5
    for _ in range(2):
        g = gi.gravity_coupling_qksl(1.0, av)
        cf = gi.CosmoFactors(*cosmo.drift_kick_factors(av, 0.05))
        av = cosmo.advance_a(av, 0.05)
        b.step(cf, g, wrap=True)
    assert a_run == av
    for u, w in zip(a.state(), b.state()):
        assert np.array_equal(u, w)


def test_treepm_step_is_bit_identical_to_the_oracle(pm):
    """TreePM: PM long range plus the short range, added in f32; the long
    range is a deterministic function of the f64 positions, the short range
    of their f32 rounding."""
    from gadget_ng_pm_hip import integrate as gi
    from gadget_ng_pm_hip import learned
    box, nm = 32.0, 32
    r_split = 1.25 * box / nm
    pos, vel, mass = _box(16, box, 5)
    plan = pm.Plan(nm, box, r_split=r_split)
    sr = learned.ShortRange(0)
    print(f"case kdk treepm step N={pos.shape[0]} nm={nm} r_s={r_split}", flush=True)
    lf = gi.Leapfrog(plan, pos.shape[0])
    lf.set_short_range(sr, r_split, eps2=0.01)
    lf.set_state(pos, vel, mass)
    lf.step(gi.CosmoFactors.flat(0.2), 1.0)
    p1, v1, a1 = lf.state()
    pmf = _gpu_force(plan, mass)

    def force(p):
        p32 = f32_round(p)
        lr = np.float32(pmf(p, 1.0))
        srf = np.float32(sr.accelerations(p32, mass, box, r_split, eps2=0.01, g=1.0))
        return (lr + srf).astype(np.float64)
    rp, rv, ra = io.kdk_step(pos, vel, (0.2, 0.1, 0.1), force)
    assert np.array_equal(a1, ra)
    assert np.array_equal(v1, rv)
    assert np.array_equal(p1, rp)


def test_trajectory_follows_the_fp64_pm_oracle(pm, oracle_mod):
    from gadget_ng_pm_hip import integrate as gi
    box, nm = 64.0, 32
    pos, vel, mass = _box(16, box, 6, vscale=0.3)
    plan = pm.Plan(nm, box)
    print(f"case kdk 8 steps vs FP64 oracle PM N={pos.shape[0]} nm={nm}", flush=True)
    lf = gi.Leapfrog(plan, pos.shape[0])
    lf.set_state(pos, vel, mass)
    a_end = lf.run(8, 0.05, 40.0, cosmology=gi.CosmologyParams(*LCDM), a=0.3, wrap=True)
    p8, v8, _ = lf.state()

    def f64_force(p, g):
        return oracle_mod.pm_accelerations(p, f32_round(mass), box, nm, g=g)
    rp, rv, _, _ = io.run(pos, vel, 8, 0.05, 40.0, f64_force, cosmology=LCDM, a=0.3, box=box)
    dp = np.abs((p8 - rp + box / 2) % box - box / 2)
    moved = np.abs((rp - pos + box / 2) % box - box / 2).max()
    print(f"max position difference {dp.max():.3e} of a max displacement {moved:.3e}; "
          f"velocity rms rel {np.sqrt(((v8 - rv) ** 2).sum() / ((rv - vel) ** 2).sum()):.3e}", flush=True)
    assert moved > 0.1
    assert dp.max() < 1e-4 * box / nm
    assert np.sqrt(((v8 - rv) ** 2).sum() / ((rv - vel) ** 2).sum()) < 1e-4


def test_empty_state_and_bad_arguments(pm):
    from gadget_ng_pm_hip import integrate as gi
    plan = pm.Plan(16, 16.0)
    lf = gi.Leapfrog(plan, 0)
    lf.set_state(np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0))
    lf.step(gi.CosmoFactors.flat(0.1), 1.0)
    assert lf.state()[0].shape == (0, 3)
    other = gi.Leapfrog(plan, 8)
    with pytest.raises(Exception):
        other.step(gi.CosmoFactors.flat(0.1), 1.0)  # no state yet
    other.set_state(np.zeros((8, 3)), np.zeros((8, 3)), np.ones(8))
    with pytest.raises(Exception):
        other.step(gi.CosmoFactors(float("nan"), 0.1, 0.1), 1.0)


def test_device_state_load_from_gpu_ics_equals_the_host_load(pm):
    """GPU Zel'dovich particles handed to the integrator in device memory
    give the state, and the step after it, of the host load, bit for bit."""
    from gadget_ng_pm_hip import integrate as gi
    from gadget_ng_pm_hip import ic
    n, box = 16, 32.0
    plan = pm.Plan(n, box)
    N = n ** 3
    sig = ic.sigma_table(n, lambda k: 1e-3 * k ** -1.5 if k > 0 else 0.0)
    print(f"case kdk device load from ICs N={N} nm={n}", flush=True)
    d_pos = pm.DeviceArray((N, 3), np.float64)
    d_vel = pm.DeviceArray((N, 3), np.float64)
    ic.zeldovich_ics_device(plan, 7, sig, 0.05, 1.0, d_pos.ptr, d_vel.ptr)
    pm.device_synchronize()
    pos, vel = d_pos.numpy(), d_vel.numpy()
    mass = np.full(N, 1.0 / N)
    a, b, c = gi.Leapfrog(plan, N), gi.Leapfrog(plan, N), gi.Leapfrog(plan, N)
    a.set_state(pos, vel, mass)
    b.set_state_device(d_pos.ptr, d_vel.ptr, uniform_mass=1.0 / N)
    d_mass = pm.DeviceArray.from_host(mass)
    c.set_state_device(d_pos.ptr, d_vel.ptr, d_mass.ptr)
    pm.device_synchronize()
    sa, sb, sc = a.state(), b.state(), c.state()
    for x, y, z in zip(sa, sb, sc):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    assert np.array_equal(sa[0], pos) and not sa[2].any()
    for lf in (a, b, c):
        lf.step(gi.CosmoFactors.flat(0.2), 1.5, wrap=True)
    sa, sb, sc = a.state(), b.state(), c.state()
    for x, y, z in zip(sa, sb, sc):
        assert np.array_equal(x, y) and np.array_equal(x, z)
    assert np.abs(sa[2]).max() > 0.0
    with pytest.raises(Exception):
        b.set_state_device(0, d_vel.ptr)
