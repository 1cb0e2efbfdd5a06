"""The reference's own tests for this path, ported onto the FP64 oracle.

Each test names the reference test it restates (paths relative to the
gadget-ng tree) and keeps its inputs and tolerance.
"""
import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st


def direct_accel(pos, mass, eps2, g, i):
    """DirectGravity for one particle (crates/gadget-ng-core/src/gravity.rs:5-12,60-86)."""
    a = np.zeros(3)
    for j in range(pos.shape[0]):
        if j == i:
            continue
        r = pos[j] - pos[i]
        r2 = r @ r + eps2
        a += g * mass[j] * r / r2 ** 1.5
    return a


# ---- crates/gadget-ng-pm/src/cic.rs:461-587 ---------------------------------

def test_assign_single_particle_on_grid_node(oracle_mod):
    rho = oracle_mod.cic_assign(np.zeros((1, 3)), np.ones(1), 4.0, 4)
    assert abs(rho.reshape(-1)[0] - 1.0) < 1e-12
    assert abs(rho.sum() - 1.0) < 1e-12


def test_assign_conserves_total_mass(oracle_mod):
    i = np.arange(27)
    pos = np.stack([(i % 3) / 3.0 + 0.1, ((i // 3) % 3) / 3.0 + 0.1, (i // 9) / 3.0 + 0.1], axis=1)
    rho = oracle_mod.cic_assign(pos, np.full(27, 2.0), 1.0, 8)
    assert abs(rho.sum() - 54.0) < 1e-10


def test_interpolate_constant_field_gives_same_value(oracle_mod):
    nm = 4
    fx = np.ones((nm, nm, nm))
    z = np.zeros((nm, nm, nm))
    acc = oracle_mod.cic_interpolate(fx, z, z, np.array([[0.3, 0.6, 0.7]]), 1.0, nm)
    assert abs(acc[0, 0] - 1.0) < 1e-12 and abs(acc[0, 1]) < 1e-12 and abs(acc[0, 2]) < 1e-12


def test_assign_symmetry_at_center(oracle_mod):
    rho = oracle_mod.cic_assign(np.array([[0.5, 0.5, 0.5]]), np.array([8.0]), 4.0, 4).reshape(-1)
    assert abs(rho[0] - 1.0) < 1e-12
    assert abs(rho[21] - 1.0) < 1e-12  # node (1,1,1)
    assert abs(rho.sum() - 8.0) < 1e-10


def test_interpolation_of_assigned_density_is_positive(oracle_mod):
    nm = 8
    pos = np.array([[0.33, 0.55, 0.77]])
    rho = oracle_mod.cic_assign(pos, np.ones(1), 1.0, nm)
    z = np.zeros_like(rho)
    acc = oracle_mod.cic_interpolate(rho, z, z, pos, 1.0, nm)
    assert acc[0, 0] > 0.0


def test_assign_periodic_wrapping(oracle_mod):
    rho = oracle_mod.cic_assign(np.array([[3.99, 3.99, 3.99]]), np.ones(1), 4.0, 4)
    assert abs(rho.sum() - 1.0) < 1e-12


# ---- crates/gadget-ng-pm/src/fft_poisson.rs:820-854 -------------------------

def test_uniform_density_gives_zero_force(oracle_mod):
    for mode in (oracle_mod.MODE_FAITHFUL, oracle_mod.MODE_LEAN):
        f = oracle_mod.solve_forces(np.ones((8, 8, 8)), 1.0, 8, 1.0, mode=mode)
        assert max(np.abs(c).max() for c in f) < 1e-10


def test_fft3d_roundtrip(oracle_mod):
    nm = 4
    orig = np.arange(nm ** 3, dtype=np.float64).reshape(nm, nm, nm)
    back = oracle_mod.fft3d(oracle_mod.fft3d(orig), inverse=True) / nm ** 3
    assert np.abs(back.real - orig).max() < 1e-10


# ---- crates/gadget-ng-pm/tests/pm_uniform_zero_force.rs:32-80 ----------------

def uniform_lattice(np_side, box):
    step = box / np_side
    i = np.arange(np_side)
    z, y, x = np.meshgrid(i, i, i, indexing="ij")
    pos = np.stack([(x.ravel() + 0.5) * step, (y.ravel() + 0.5) * step, (z.ravel() + 0.5) * step], axis=1)
    return pos, np.ones(pos.shape[0])


def test_uniform_lattice_gives_zero_acceleration(oracle_mod):
    pos, m = uniform_lattice(4, 1.0)
    acc = oracle_mod.pm_accelerations(pos, m, 1.0, 16, 1.0, mode=oracle_mod.MODE_FAITHFUL)
    assert np.linalg.norm(acc, axis=1).max() < 1e-10


def test_partial_indices_subset(oracle_mod):
    pos, m = uniform_lattice(4, 1.0)
    acc = oracle_mod.pm_accelerations(pos, m, 1.0, 8, 1.0, idx=np.arange(4))
    assert acc.shape == (4, 3)
    assert np.linalg.norm(acc, axis=1).max() < 1e-10


# ---- crates/gadget-ng-pm/tests/pm_vs_direct.rs:31-110 -----------------------

def center_mass_probe(box):
    pos = np.array([[0.5, 0.5, 0.5], [0.6, 0.5, 0.5]]) * box
    return pos, np.array([10.0, 0.001]), 1


def test_pm_force_direction_matches_direct(oracle_mod):
    pos, m, probe = center_mass_probe(1.0)
    a_direct = direct_accel(pos, m, 1e-4, 1.0, probe)
    a_pm = oracle_mod.pm_accelerations(pos, m, 1.0, 32, 1.0, idx=[probe])[0]
    assert a_direct[0] < 0 and a_pm[0] < 0
    ratio = abs(a_pm[0]) / abs(a_direct[0])
    assert 0.05 < ratio < 50.0


def test_pm_total_momentum_conserved(oracle_mod):
    pos, m, _ = center_mass_probe(1.0)
    acc = oracle_mod.pm_accelerations(pos, m, 1.0, 16, 1.0)
    s = (acc * m[:, None]).sum(axis=0)
    scale = (np.linalg.norm(acc, axis=1) * m).max()
    assert np.linalg.norm(s) < 1e-8 * max(scale, 1e-15)


# ---- crates/gadget-ng-pm/tests/proptest_invariants.rs:8-22 -------------------

@settings(max_examples=60, deadline=None)
@given(st.lists(st.tuples(st.floats(-0.5, 1.5), st.floats(-0.5, 1.5), st.floats(-0.5, 1.5), st.floats(0.1, 10.0)),
                min_size=1, max_size=99))
def test_cic_conserves_total_mass_property(oracle_mod, pts):
    a = np.array(pts, dtype=np.float64)
    rho = oracle_mod.cic_assign(a[:, :3], a[:, 3], 1.0, 8)
    assert abs(rho.sum() - a[:, 3].sum()) < 1e-10


# ---- crates/gadget-ng-physics/tests/cosmo_pm.rs ------------------------------

def test_cic_mass_conservation_lcg(oracle_mod):
    """cosmo_pm.rs:171-197 (its LCG positions and masses (i+1)*0.1)."""
    state = 17
    mask = (1 << 64) - 1
    pos = np.empty((200, 3))
    for i in range(200):
        for c in range(3):
            state = (state * 6364136223846793005 + 1) & mask
            pos[i, c] = (state >> 33) / float(2 ** 32 - 1)
    m = (np.arange(200) + 1) * 0.1
    rho = oracle_mod.cic_assign(pos, m, 1.0, 16)
    assert abs(rho.sum() - m.sum()) / m.sum() < 1e-12


def test_pm_poisson_single_mode_signs(oracle_mod):
    """cosmo_pm.rs:211-258: sine along x (flat = iz nm^2 + iy nm + ix)."""
    nm = 32
    ix = np.arange(nm)
    dens = np.broadcast_to(np.sin(2 * np.pi * ix / nm)[None, None, :], (nm, nm, nm)).copy()
    fx, _, _ = oracle_mod.solve_forces(dens, 1.0, nm, 1.0, mode=oracle_mod.MODE_FAITHFUL)
    fx = fx.reshape(-1)
    assert fx[0] > 0 and fx[nm // 2] < 0
    assert abs(fx[0] / -fx[nm // 2] - 1.0) < 0.01
    assert abs(fx[nm // 4]) / abs(fx[0]) < 0.01


def test_pm_g_scaling(oracle_mod):
    """cosmo_pm.rs:265-296: F(G/2) = F(G)/2 to 1e-10."""
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 1.0, (27, 3))
    m = np.full(27, 1.0 / 27)
    a1 = oracle_mod.pm_accelerations(pos, m, 1.0, 16, 1.0)
    a2 = oracle_mod.pm_accelerations(pos, m, 1.0, 16, 0.5)
    rel = np.linalg.norm(0.5 * a1 - a2, axis=1) / np.maximum(np.linalg.norm(a1, axis=1), 1e-15)
    assert rel.max() < 1e-10


def test_pm_periodic_force_symmetry(oracle_mod):
    """cosmo_pm.rs:416-449."""
    pos = np.array([[0.1, 0.5, 0.5], [0.9, 0.5, 0.5]])
    acc = oracle_mod.pm_accelerations(pos, np.ones(2), 1.0, 32, 1.0)
    assert acc[0, 0] / (-acc[1, 0] + 1e-30) > 0.5
    for a in acc:
        assert np.hypot(a[1], a[2]) / max(abs(a[0]), 1e-20) < 0.15


# ---- crates/gadget-ng-treepm/src/solver.rs:56-62 ----------------------------

def test_treepm_auto_r_split_is_two_and_a_half_cells(oracle_mod):
    rng = np.random.default_rng(9)
    pos = rng.uniform(0, 10.0, (300, 3))
    m = np.ones(300)
    a_auto = oracle_mod.treepm_long_range(pos, m, 10.0, 16, 1.0, r_split=0.0)
    a_expl = oracle_mod.pm_accelerations(pos, m, 10.0, 16, 1.0, r_split=2.5 * 10.0 / 16)
    assert np.array_equal(a_auto, a_expl)
