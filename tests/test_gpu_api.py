"""GPU tests of every entry point of the C ABI and of the host mirror, and the
reference's own PM tests run through the HIP path."""
import ctypes
import json
import os

import numpy as np
import pytest

from helpers import f32_round, max_rel_robust, rms_rel, uniform_particles

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_vectors_through_gpu(pm, oracle_mod, name):
    d = np.load(os.path.join(HERE, "golden", f"{name}.npz"), allow_pickle=False)
    nm, n, box, g, rs, eps, dc = d["params"]
    nm = int(nm)
    print(f"case golden {name}", flush=True)
    got = pm.Plan(nm, box, r_split=rs, plummer_eps=eps, deconvolve=bool(dc)).accelerations_f64(d["pos"], d["mass"], g)
    # against the vectors themselves (f64 inputs; the GPU rounds positions to f32)
    assert rms_rel(got, d["acc"]) <= 2e-5
    # against the oracle on the same f32-rounded inputs
    ref = oracle_mod.pm_accelerations(f32_round(d["pos"]), f32_round(d["mass"]), box, nm, g, r_split=rs,
                                      plummer_eps=eps, deconvolve=bool(dc))
    assert rms_rel(got, ref) <= 1e-5


def test_config1_uniform_32cubed_on_64cubed(pm, oracle_mod):
    """N = 32^3 uniform in [0,100)^3, m = 1/N, 64^3 mesh, G = 1; density and forces vs FP64."""
    from gadget_ng_pm_hip import workloads
    p = workloads.uniform_box()
    plan = pm.Plan(64, p.box, keep_density=True)
    got = plan.accelerations_f64(p.pos64(), p.mass64(), 1.0)
    ref, dens_ref = oracle_mod.pm_accelerations(p.pos64(), p.mass64(), p.box, 64, 1.0, want_density=True,
                                                mode=oracle_mod.MODE_FAITHFUL)
    assert rms_rel(got, ref) <= 1e-5
    assert max_rel_robust(got, ref) <= 1e-4
    dens = plan.density()
    assert np.abs(dens - dens_ref).max() <= 1e-6 * dens_ref.max()


def test_uniform_lattice_gives_near_zero_force(pm):
    """crates/gadget-ng-pm/tests/pm_uniform_zero_force.rs:32-53 (FP32: relative to a single particle's pull)."""
    i = np.arange(4)
    z, y, x = np.meshgrid(i, i, i, indexing="ij")
    pos = np.stack([(x.ravel() + 0.5) / 4, (y.ravel() + 0.5) / 4, (z.ravel() + 0.5) / 4], axis=1)
    m = np.ones(64)
    solver = pm.PmSolver(16, 1.0)
    acc = solver.accelerations_for_indices(pos, m, 0.0, 1.0, np.arange(64))
    # one unit mass at a quarter box gives |a| ~ 1/0.25^2 = 16; the lattice cancels to FP32 rounding
    assert np.linalg.norm(acc, axis=1).max() < 1e-4
    sub = solver.accelerations_for_indices(pos, m, 0.0, 1.0, np.arange(4))
    assert sub.shape == (4, 3)


def test_momentum_and_direction(pm):
    """crates/gadget-ng-pm/tests/pm_vs_direct.rs:31-110."""
    pos = np.array([[0.5, 0.5, 0.5], [0.6, 0.5, 0.5]])
    m = np.array([10.0, 0.001])
    a = pm.PmSolver(32, 1.0).accelerations_for_indices(pos, m, 1e-4, 1.0, [1])
    assert a[0, 0] < 0
    acc = pm.PmSolver(16, 1.0).accelerations_for_indices(pos, m, 0.0, 1.0, [0, 1])
    s = (acc * f32_round(m)[:, None]).sum(axis=0)
    # In FP64 the sum vanishes to 1e-8 of max m|a|.  In FP32 the heavy particle
    # feels the rounding noise of its own field, whose scale next to it is
    # G m_heavy / h^2; a wrong-sign Green's function would leave an O(1) residue.
    h = 1.0 / 16
    self_field = 1.0 * m[0] / h ** 2
    assert np.linalg.norm(s) < 1e-6 * m[0] * self_field


def test_g_linearity_is_exact(pm):
    """cosmo_pm.rs:265-296 asks 1e-10; halving G is exact in binary, so the GPU result halves bit for bit."""
    rng = np.random.default_rng(3)
    pos, m = uniform_particles(rng, 2000, 10.0)
    plan = pm.Plan(32, 10.0)
    a1 = plan.accelerations_f64(pos, m, 1.0)
    a2 = plan.accelerations_f64(pos, m, 0.5)
    assert np.array_equal(0.5 * a1, a2)


def test_periodic_force_symmetry(pm):
    """cosmo_pm.rs:416-449."""
    acc = pm.PmSolver(32, 1.0).accelerations_for_indices(np.array([[0.1, 0.5, 0.5], [0.9, 0.5, 0.5]]), np.ones(2),
                                                          0.0, 1.0, [0, 1])
    assert acc[0, 0] / (-acc[1, 0] + 1e-30) > 0.5
    for a in acc:
        assert np.hypot(a[1], a[2]) / abs(a[0]) < 0.15


def test_results_are_bitwise_reproducible(pm):
    rng = np.random.default_rng(21)
    pos, m = uniform_particles(rng, 50000, 20.0)
    plan = pm.Plan(64, 20.0, keep_density=True)
    a1 = plan.accelerations_f64(pos, m, 1.0)
    d1 = plan.density()
    a2 = plan.accelerations_f64(pos, m, 1.0)
    assert np.array_equal(a1, a2) and np.array_equal(d1, plan.density())


def test_legacy_hip_pm_abi(pm, oracle_mod):
    """hip_pm_create / hip_pm_solve / hip_pm_destroy as the reference crate calls them
    (crates/gadget-ng-hip/src/pm_solver.rs:101,190-203), with r_split passed per call."""
    L = pm.lib()
    rng = np.random.default_rng(2)
    n, nm, box = 5000, 32, 50.0
    pos, m = uniform_particles(rng, n, box)
    h = L.hip_pm_create(nm, ctypes.c_float(box))
    assert h
    try:
        x, y, z = (np.ascontiguousarray(pos[:, c], dtype=np.float32) for c in range(3))
        mf = m.astype(np.float32)
        out = np.zeros((3, n), dtype=np.float32)
        for rs in (0.0, 3.0):
            rc = L.hip_pm_solve(h, x.ctypes.data, y.ctypes.data, z.ctypes.data, mf.ctypes.data, out[0].ctypes.data,
                                out[1].ctypes.data, out[2].ctypes.data, n, ctypes.c_float(0.0), ctypes.c_float(1.0),
                                ctypes.c_float(rs))
            assert rc == 0
            ref = oracle_mod.pm_accelerations(pos, m, box, nm, 1.0, r_split=float(np.float32(rs)))
            assert rms_rel(out.T, ref) <= 1e-5
    finally:
        L.hip_pm_destroy(h)


def test_f32_f64_and_device_entry_points_agree(pm):
    rng = np.random.default_rng(4)
    n, nm, box = 30000, 48, 30.0
    pos, m = uniform_particles(rng, n, box)
    plan = pm.Plan(nm, box)
    a64 = plan.accelerations_f64(pos, m, 1.0)
    a32 = plan.accelerations_f32(pos[:, 0], pos[:, 1], pos[:, 2], m, 1.0)
    assert np.array_equal(a32.T.astype(np.float64), a64)
    xs = [pm.DeviceArray.from_host(np.ascontiguousarray(pos[:, c], dtype=np.float32)) for c in range(3)]
    dm = pm.DeviceArray.from_host(m.astype(np.float32))
    sel = rng.choice(n, 999, replace=False).astype(np.int64)
    didx = pm.DeviceArray.from_host(sel)
    out = [pm.DeviceArray((999,), np.float32) for _ in range(3)]
    plan.solve_device(xs[0].ptr, xs[1].ptr, xs[2].ptr, dm.ptr, n, out[0].ptr, out[1].ptr, out[2].ptr, 1.0,
                      idx_ptr=didx.ptr, k=999)
    pm.device_synchronize()
    got = np.stack([o.numpy() for o in out], axis=1).astype(np.float64)
    assert np.array_equal(got, a64[sel])


def test_host_entries_fill_a_caller_output_array(pm):
    """out= writes into the caller's array (the reference fills `out: &mut
    [Vec3]`): same values as a fresh result, the array itself is returned,
    a wrong shape, dtype or layout is refused, and the drop-in solvers fill
    a suitable out in place."""
    rng = np.random.default_rng(14)
    n, nm, box = 20000, 32, 25.0
    pos, m = uniform_particles(rng, n, box)
    plan = pm.Plan(nm, box)
    ref = plan.accelerations_f64(pos, m, 1.0)
    out = np.full((n, 3), np.nan)
    assert plan.accelerations_f64(pos, m, 1.0, out=out) is out
    assert np.array_equal(out, ref)
    sel = rng.choice(n, 777, replace=False).astype(np.int64)
    outk = np.empty((777, 3))
    plan.accelerations_f64(pos, m, 1.0, idx=sel, out=outk)
    assert np.array_equal(outk, ref[sel])
    for bad in (np.empty((n, 3), np.float32), np.empty((n, 4)), np.empty((3, n)).T, np.empty((n - 1, 3))):
        with pytest.raises(ValueError):
            plan.accelerations_f64(pos, m, 1.0, out=bad)
    o32 = np.empty((3, n), np.float32)
    r32 = plan.accelerations_f32(pos[:, 0], pos[:, 1], pos[:, 2], m, 1.0)
    assert plan.accelerations_f32(pos[:, 0], pos[:, 1], pos[:, 2], m, 1.0, out=o32) is o32
    assert np.array_equal(o32, r32)
    with pytest.raises(ValueError):
        plan.accelerations_f32(pos[:, 0], pos[:, 1], pos[:, 2], m, 1.0, out=np.empty((n, 3), np.float32))
    solver = pm.PmSolver(nm, box)
    o = np.empty((777, 3))
    assert solver.accelerations_for_indices(pos, m, 0.0, 1.0, sel, out=o) is o
    assert np.array_equal(o, ref[sel])
    ol = np.empty((777, 3), order="F")  # not C-contiguous: filled through a copy
    solver.accelerations_for_indices(pos, m, 0.0, 1.0, sel, out=ol)
    assert np.array_equal(ol, ref[sel])


def test_hip_pm_solver_mirror(pm, oracle_mod):
    assert pm.HipPmSolver.is_available()
    s = pm.HipPmSolver.try_new_with_r_split(32, 16.0, 1.2)
    assert s is not None and s.grid_size() == 32 and s.r_split() == 1.2
    rng = np.random.default_rng(8)
    pos, m = uniform_particles(rng, 3000, 16.0)
    out = np.zeros((10, 3))
    s.accelerations_for_indices(pos, m, 0.0, 1.0, np.arange(10), out)
    ref = oracle_mod.pm_accelerations(pos, m, 16.0, 32, 1.0, r_split=1.2, idx=np.arange(10))
    assert rms_rel(out, ref) <= 1e-5
    assert pm.HipPmSolver.try_new(100, 1.0) is None  # unsupported mesh -> None, like a failed create


def test_treepm_long_range_matches_oracle(pm, oracle_mod):
    rng = np.random.default_rng(12)
    pos, m = uniform_particles(rng, 20000, 64.0)
    lr = pm.TreePmLongRange(64, 64.0)
    got = lr.long_range_accelerations_for_indices(pos, m, 0.0, 1.0, np.arange(20000))
    ref = oracle_mod.treepm_long_range(pos, m, 64.0, 64, 1.0)
    assert rms_rel(got, ref) <= 1e-5 and max_rel_robust(got, ref) <= 1e-4


def test_empty_and_tiny_inputs(pm):
    plan = pm.Plan(8, 1.0)
    assert plan.accelerations_f64(np.zeros((0, 3)), np.zeros(0), 1.0).shape == (0, 3)
    a = plan.accelerations_f64(np.array([[0.5, 0.5, 0.5]]), np.ones(1), 1.0)
    assert np.all(np.isfinite(a)) and np.abs(a).max() < 1e-3  # a lone particle feels (almost) no self force
    a = plan.accelerations_f64(np.ones((5, 3)) * 0.3, np.zeros(5), 1.0)  # zero masses
    assert np.array_equal(a, np.zeros((5, 3)))
