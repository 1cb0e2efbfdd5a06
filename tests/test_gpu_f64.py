"""The f64 input path (positions as the reference's Vec3 f64, cell coordinate
formed in FP64 on the device) against the FP64 oracle on the SAME unrounded
positions.

The reference's GravitySolver takes f64 positions (crates/gadget-ng-pm/src/
solver.rs:74-140) and forms c = (x nm / L) rem_euclid nm in f64
(cic.rs:50-52).  Rounding the positions to f32 first costs 8e-4 rms on the
reference's phase-38 legacy 2LPT ICs and 3.2e-2 on the rescaled ones (the
particles sit within 1e-4 of the box of a lattice); the f64 path must stay
within the 1e-5 target on both.  Other checks: the device entry points agree
bit for bit with the host one and with each other, f32-representable inputs
give the f32 path's result bit for bit, and wrapped inputs and index subsets
follow the oracle."""
import numpy as np
import pytest

from helpers import f32_round, max_rel_robust, rms_rel

pytestmark = pytest.mark.gpu


def _uniform64(n, box, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, box, (n, 3)), np.full(n, 1.0 / n)


def test_uniform_unrounded_positions(pm, oracle_mod):
    """Config 1 (32^3 uniform particles, L = 100, 64^3 mesh) in f64."""
    pos, m = _uniform64(32 ** 3, 100.0, 0x6E670001)
    print("case f64 uniform N=32768 nm=64", flush=True)
    plan = pm.Plan(64, 100.0)
    got = plan.accelerations_f64(pos, m, 1.0)
    ref = oracle_mod.pm_accelerations(pos, m, 100.0, 64, 1.0)
    r = rms_rel(got, ref)
    r32 = rms_rel(oracle_mod.pm_accelerations(f32_round(pos), f32_round(m), 100.0, 64, 1.0), ref)
    print(f"f64 path rms {r:.3e}; f32 rounding of the inputs alone {r32:.3e}", flush=True)
    assert r <= 1e-5 and r < r32
    assert max_rel_robust(got, ref) <= 1e-3


@pytest.fixture(scope="module")
def phase38():
    """The reference's phase-38 2LPT initial conditions (32^3, unit box,
    seed 42, a_init = 0.02; legacy and growth-rescaled), built by the oracle."""
    from test_ic_reference_pk import A_INIT, BOX, N, SEED  # noqa: F401
    import ic_oracle as ico
    from gadget_ng_pm_hip import ic
    from gadget_ng_pm_hip import integrate as gi
    p = ic.EisensteinHuParams(0.315, 0.049, 0.674, 2.7255)
    fn = ic.eisenstein_hu_sigma(N, 0.965, 1e-4, p, sigma8=0.8, box_size_mpc_h=100.0)
    sig = ic.sigma_table(N, fn)
    delta = ico.delta_kspace(N, SEED, lambda n2: sig[np.asarray(n2)])
    psi1 = ico.displacement(delta, BOX)
    psi2 = ico.source_to_psi2(ico.source_2lpt(ico.phi_second_derivatives(delta)), BOX)
    out = {}
    for rescale in (False, True):
        v1, v2, d2, sc = ic.factors_2lpt(gi.CosmologyParams(0.315, 0.685, 0.1), A_INIT, 0.1, rescale)
        pos, _ = ico.zeldovich_2lpt(N, BOX, psi1, psi2, v1, v2, d2, sc)
        out[rescale] = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    return N, BOX, out


@pytest.mark.parametrize("rescale", [False, True])
def test_reference_2lpt_ics_unrounded(pm, oracle_mod, phase38, rescale):
    n, box, sets = phase38
    pos = sets[rescale]
    m = np.full(pos.shape[0], 1.0 / pos.shape[0])
    print(f"case f64 phase-38 2LPT ICs rescale={rescale} N={pos.shape[0]} nm={n}", flush=True)
    plan = pm.Plan(n, box)
    got = plan.accelerations_f64(pos, m, 1.0)
    ref = oracle_mod.pm_accelerations(pos, m, box, n, 1.0)
    r = rms_rel(got, ref)
    r32 = rms_rel(oracle_mod.pm_accelerations(f32_round(pos), f32_round(m), box, n, 1.0), ref)
    print(f"f64 path rms {r:.3e}, max {max_rel_robust(got, ref):.3e}; "
          f"f32 rounding of the inputs alone {r32:.3e}", flush=True)
    assert r <= 1e-5
    assert r32 > 10.0 * r


def test_device_entries_agree_with_the_host_entry(pm):
    pos, m = _uniform64(20000, 50.0, 7)
    pos[:100] -= 50.0   # wrapped from below
    pos[100:200] += 50.0  # wrapped from above
    print("case f64 device entries N=20000 nm=32", flush=True)
    plan = pm.Plan(32, 50.0)
    host = plan.accelerations_f64(pos, m, 2.0)
    dpos = pm.DeviceArray.from_host(pos)
    dm = pm.DeviceArray.from_host(m)
    dout = pm.DeviceArray(pos.shape, np.float64)
    plan.solve_device_f64(dpos.ptr, dm.ptr, pos.shape[0], dout.ptr, 2.0)
    pm.device_synchronize()
    assert np.array_equal(dout.numpy(), host)
    # SoA f64 positions with f32 masses and accelerations (the integrator's form)
    xs = [pm.DeviceArray.from_host(np.ascontiguousarray(pos[:, c])) for c in range(3)]
    dm32 = pm.DeviceArray.from_host(m.astype(np.float32))
    outs = [pm.DeviceArray((pos.shape[0],), np.float32) for _ in range(3)]
    plan.solve_device_soa64(*(a.ptr for a in xs), dm32.ptr, pos.shape[0], *(o.ptr for o in outs), 2.0)
    pm.device_synchronize()
    soa = np.stack([o.numpy() for o in outs], axis=1).astype(np.float64)
    # the f64 masses enter the deposit as f32, so both see the same masses
    assert np.array_equal(soa, host)


def test_f32_representable_inputs_give_the_f32_path_bit_for_bit(pm):
    pos, m = _uniform64(30000, 64.0, 11)
    pos, m = f32_round(pos), f32_round(m)
    print("case f64 vs f32 path on f32 values N=30000 nm=64", flush=True)
    plan = pm.Plan(64, 64.0)
    a64 = plan.accelerations_f64(pos, m, 1.0)
    a32 = plan.accelerations_f32(*(pos[:, c].astype(np.float32) for c in range(3)), m.astype(np.float32), 1.0)
    assert np.array_equal(a64, a32.T.astype(np.float64))


def test_wrapped_positions_and_index_subset(pm, oracle_mod):
    rng = np.random.default_rng(3)
    n, box, nm = 8192, 10.0, 16
    pos = rng.uniform(-2.0 * box, 3.0 * box, (n, 3))
    pos[0] = (-1e-17, box, 2.0 * box)  # rem_euclid rounding to L and exact multiples of L
    m = rng.uniform(0.5, 1.5, n)
    sel = rng.choice(n, 777, replace=False).astype(np.int64)
    print("case f64 wrapped positions and index subset N=8192 nm=16", flush=True)
    plan = pm.Plan(nm, box)
    got = plan.accelerations_f64(pos, m, 1.0, idx=sel)
    ref = oracle_mod.pm_accelerations(pos, m, box, nm, 1.0, idx=sel)
    print(f"rms {rms_rel(got, ref):.3e}", flush=True)
    assert rms_rel(got, ref) <= 1e-5


def test_uniform_lattice_gives_zero_force_in_f64(pm):
    """A lattice exactly on the nodes: every cell holds the same mass, the
    mean-subtracted mesh is one constant, and the forces are exactly zero
    (pm_uniform_zero_force.rs:32-53 asks for 1e-10)."""
    ns, box = 16, 16.0
    g = np.arange(ns) * (box / ns)
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    m = np.full(pos.shape[0], 1.0 / pos.shape[0])
    print("case f64 node lattice 16^3 nm=16", flush=True)
    plan = pm.Plan(16, box)
    a = plan.accelerations_f64(pos, m, 1.0)
    assert np.abs(a).max() <= 1e-10


def test_non_finite_coordinates_stay_on_the_mesh(pm, oracle_mod):
    """NaN and infinite coordinates put their particle on node 0 of that
    axis: the solve stays in bounds and equals the oracle's on the same
    particles with those coordinates replaced by 0 (the reference would
    return NaN everywhere)."""
    rng = np.random.default_rng(17)
    n, box, nm = 5000, 20.0, 32
    pos = rng.uniform(0.0, box, (n, 3))
    m = rng.uniform(0.5, 1.5, n)
    bad = np.array([0, 1, 2, 3])
    pos[0, 0] = np.nan
    pos[1, 1] = np.inf
    pos[2, 2] = -np.inf
    pos[3] = (np.nan, np.nan, np.nan)
    print("case f64 non-finite coordinates N=5000 nm=32", flush=True)
    plan = pm.Plan(nm, box)
    got = plan.accelerations_f64(pos, m, 1.0)
    fixed = np.where(np.isfinite(pos), pos, 0.0)
    ref = oracle_mod.pm_accelerations(fixed, m, box, nm, 1.0)
    assert np.all(np.isfinite(got))
    keep = np.setdiff1d(np.arange(n), bad)
    assert rms_rel(got[keep], ref[keep]) <= 1e-5
    a32 = plan.accelerations_f32(*(pos[:, c].astype(np.float32) for c in range(3)), m.astype(np.float32), 1.0)
    assert np.all(np.isfinite(a32))
