"""Every compiled launch variant gives the oracle's forces; the autotuner
picks one of them and the choice carries over to new plans."""
import os

import numpy as np
import pytest

from helpers import rms_rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nm", [8, 16, 24, 64, 96, 128, 256])
def test_every_variant_matches_oracle(pm, oracle_mod, nm):
    print(f"case variants nm={nm}", flush=True)
    g = np.random.default_rng(nm)
    L = 10.0
    n = min(nm ** 3 // 4, 400000)
    pos = g.uniform(0, L, (n, 3)).astype(np.float32).astype(np.float64)
    m = g.uniform(0.5, 1.5, n)
    ref = oracle_mod.pm_accelerations(pos, m, L, nm, r_split=0.3)
    plan = pm.Plan(nm, L, r_split=0.3, device=0)
    for zp, dt, yf, yi in [(False, 256, False, False), (True, 512, True, True), (False, 512, True, False),
                           (True, 256, False, True)]:
        plan.set_tuning(zp, dt, yf, yi)
        got = plan.accelerations_f64(pos, m, 1.0)
        assert rms_rel(got, ref) < 1e-5, (zp, dt, yf, yi)
        assert plan.tuning() == {"z_persistent": zp, "deposit_threads": dt, "y_fwd_persistent": yf,
                                 "y_inv_persistent": yi, "interp_sorted": False, "chunk_fwd": 0, "chunk_inv": 0,
                                 "deposit_split": True, "y_inv_narrow": True, "deposit_pair": False,
                                 "tuned": False}


@pytest.mark.parametrize("nm,chunks", [(32, (4, 16)), (64, (8, 32)), (96, (12, 48)), (128, (16, 64))])
def test_plane_chunks_give_identical_forces(pm, oracle_mod, nm, chunks):
    """The x/y pass pairs run chunk by chunk over z planes touch the same
    values with the same kernels: forces are bit-identical to the whole-mesh
    passes, with and without the potential kept (which turns chunking off for
    the inverse pair)."""
    print(f"case plane chunks nm={nm} chunks={chunks}", flush=True)
    g = np.random.default_rng(nm + 7)
    L = 12.0
    n = min(nm ** 3 // 4, 300000)
    pos = g.uniform(0, L, (n, 3)).astype(np.float32).astype(np.float64)
    m = g.uniform(0.5, 1.5, n)
    plan = pm.Plan(nm, L, r_split=0.4, device=0)
    plan.set_tuning()
    ref = plan.accelerations_f64(pos, m, 1.0)
    assert rms_rel(ref, oracle_mod.pm_accelerations(pos, m, L, nm, r_split=0.4)) < 1e-5
    for c in chunks:
        for cf, ci in ((c, c), (c, 0), (0, c), (chunks[0], chunks[1])):
            plan.set_tuning(chunk_fwd=cf, chunk_inv=ci)
            assert (plan.tuning()["chunk_fwd"], plan.tuning()["chunk_inv"]) == (cf, ci)
            assert np.array_equal(plan.accelerations_f64(pos, m, 1.0), ref), (cf, ci)
        plan.set_tuning(chunk_fwd=c, chunk_inv=c, y_fwd_persistent=True, y_inv_persistent=True)
        assert np.array_equal(plan.accelerations_f64(pos, m, 1.0), ref), (c, "persistent")
    kept = pm.Plan(nm, L, r_split=0.4, device=0, keep_potential=True)
    kept.set_tuning(chunk_fwd=chunks[1], chunk_inv=chunks[0])
    assert np.array_equal(kept.accelerations_f64(pos, m, 1.0), ref)
    for bad in (-1, nm, 2 * nm, 5 if nm % 5 else 7):
        with pytest.raises(Exception):
            plan.set_tuning(chunk_fwd=bad)
        with pytest.raises(Exception):
            plan.set_tuning(chunk_inv=bad)


@pytest.mark.parametrize("nm", [16, 48, 128])
def test_interpolation_orders_agree_bit_for_bit(pm, oracle_mod, nm):
    """The gather in the deposit's sorted order writes the same values as the
    gather in the caller's order; an index subset takes the caller's order."""
    print(f"case interpolation order nm={nm}", flush=True)
    g = np.random.default_rng(nm + 1)
    L = 7.0
    n = min(nm ** 3 // 2, 300000)
    pos = g.uniform(-L, 2 * L, (n, 3)).astype(np.float32).astype(np.float64)  # wrapped by the solver
    m = g.uniform(0.5, 1.5, n)
    plan = pm.Plan(nm, L, device=0)
    plan.set_tuning(interp_sorted=True)
    a = plan.accelerations_f64(pos, m, 1.0)
    plan.set_tuning(interp_sorted=False)
    b = plan.accelerations_f64(pos, m, 1.0)
    assert np.array_equal(a, b)
    idx = g.permutation(n)[: n // 3]
    plan.set_tuning(interp_sorted=True)
    assert np.array_equal(plan.accelerations_f64(pos, m, 1.0, idx=idx), a[idx])
    assert rms_rel(a, oracle_mod.pm_accelerations(pos, m, L, nm)) < 1e-5


def test_autotune_picks_a_variant_and_new_plans_reuse_it(pm, oracle_mod, tmp_path):
    nm, L = 64, 20.0
    plan = pm.Plan(nm, L, device=0)
    t = plan.autotune(reps=2)
    assert t["tuned"]
    again = pm.Plan(nm, L, device=0).tuning()
    assert again == t
    g = np.random.default_rng(3)
    pos = g.uniform(0, L, (5000, 3)).astype(np.float32).astype(np.float64)
    m = np.ones(5000)
    assert rms_rel(plan.accelerations_f64(pos, m, 1.0), oracle_mod.pm_accelerations(pos, m, L, nm)) < 1e-5


def test_batch_autotune_recaptures(pm, oracle_mod):
    from gadget_ng_pm_hip import workloads
    boxes, off, x, y, z, m = workloads.ensemble(n_boxes=4, sides=(8, 16), workers=-1)
    batch = pm.Batch(boxes)
    d = [pm.DeviceArray.from_host(a) for a in (x, y, z, m)]
    out = [pm.DeviceArray((x.shape[0],), np.float32) for _ in range(3)]
    batch.solve_device(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, off, 1.0, out[0].ptr, out[1].ptr, out[2].ptr)
    batch.autotune(reps=1)
    assert all(batch.tuning(i)["tuned"] for i in range(len(boxes)))
    batch.solve_device(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, off, 1.0, out[0].ptr, out[1].ptr, out[2].ptr)
    pm.device_synchronize()
    a = np.stack([o.numpy() for o in out], axis=1).astype(np.float64)
    for b, bx in enumerate(boxes):
        p = bx["particles"]
        ref = oracle_mod.pm_accelerations(p.pos64(), p.mass64(), bx["box_size"], bx["grid_size"], 1.0)
        assert rms_rel(a[off[b]:off[b + 1]], ref) <= 1e-5


def test_split_deposit_of_dense_tiles_is_bit_identical(pm, oracle_mod):
    """A Plummer clump of 150 000 particles inside one deposit tile: with the
    split on, several workgroups share that tile (their exact int64 partial
    tiles are summed in scratch); the density and the forces equal the
    unsplit deposit's bit for bit, twice in a row (the scratch is left zero),
    and match the oracle."""
    from gadget_ng_pm_hip import workloads
    nm, L = 64, 64.0
    rng = np.random.default_rng(7)
    sph, _ = workloads.plummer_spheres(150000, 1, L, 0.3, 11)
    sph = np.asarray(sph).T
    rel = (sph - sph[0] + 0.5 * L) % L - 0.5 * L  # unwrapped about one member
    sph = (rel - rel.mean(axis=0) + 24.0) % L     # one clump, centred inside one 16-cell tile
    back = rng.uniform(0.0, L, (60000, 3))
    pos = np.concatenate([sph, back])
    m = np.full(pos.shape[0], 1.0)
    print(f"case split deposit N={pos.shape[0]} nm={nm}", flush=True)
    plan = pm.Plan(nm, L, keep_density=True)
    plan.set_tuning(deposit_split=False)
    a0 = plan.accelerations_f64(pos, m, 1.0)
    d0 = plan.density()
    plan.set_tuning(deposit_split=True)
    for _ in range(2):
        a1 = plan.accelerations_f64(pos, m, 1.0)
        assert np.array_equal(plan.density(), d0)
        assert np.array_equal(a1, a0)
    ref = oracle_mod.pm_accelerations(pos, m, L, nm, 1.0)
    assert rms_rel(a1, ref) <= 1e-5


def test_inverse_y_tile_widths_are_bit_identical(pm, oracle_mod):
    """At nm = 512 the inverse y pass is compiled for 16-column tiles of 1024
    threads and 8-column tiles of 512: every column's transform is the same,
    so the forces agree bit for bit, chunked or not, and match the oracle."""
    nm, L = 512, 100.0
    g = np.random.default_rng(512)
    n = 200000
    pos = g.uniform(0, L, (n, 3))
    m = g.uniform(0.5, 1.5, n)
    print(f"case inverse y tile widths nm={nm} n={n}", flush=True)
    plan = pm.Plan(nm, L, device=0)
    out = {}
    for narrow in (False, True):
        for ci in (0, 64):
            plan.set_tuning(y_inv_narrow=narrow, chunk_inv=ci)
            assert plan.tuning()["y_inv_narrow"] == narrow
            out[(narrow, ci)] = plan.accelerations_f64(pos, m, 1.0)
    ref = out[(False, 0)]
    for k, v in out.items():
        assert np.array_equal(v, ref), k
    idx = g.choice(n, 20000, replace=False)
    assert rms_rel(ref[idx], oracle_mod.pm_accelerations(pos, m, L, nm, 1.0, idx=idx)) <= 1e-5


@pytest.mark.parametrize("nm", [8, 12, 24, 48, 128, 512])
def test_paired_tile_deposit_is_bit_identical(pm, oracle_mod, nm):
    """Two tiles per deposit workgroup (the second tile's records load while
    the first is written) give the one-tile deposit's density and forces bit
    for bit: uniform particles plus a dense clump, with odd tile counts (1 at
    nm = 8, 12, 24 and 27 at 48: the last workgroup holds one tile) and even
    ones, with and without the split of dense tiles (which takes precedence)."""
    g = np.random.default_rng(nm + 5)
    L = 10.0
    n = min(nm ** 3 // 2, 400000)
    pos = g.uniform(0, L, (n, 3))
    pos[: n // 4] = (g.normal(0.61 * L, 0.01 * L, (n // 4, 3))) % L
    pos[n // 4: n // 4 + 50] = 0.5 * L  # one crowded cell
    m = g.uniform(0.5, 1.5, n)
    print(f"case paired-tile deposit nm={nm} n={n}", flush=True)
    plan = pm.Plan(nm, L, device=0, keep_density=True)
    out = {}
    for threads in (256, 512):
        for split in (False, True):
            for pair in (False, True):
                plan.set_tuning(deposit_threads=threads, deposit_split=split, deposit_pair=pair)
                assert plan.tuning()["deposit_pair"] == pair
                out[(threads, split, pair)] = (plan.accelerations_f64(pos, m, 1.0), plan.density())
    a0, d0 = out[(256, False, False)]
    for k, (a, d) in out.items():
        assert np.array_equal(d, d0), k
        assert np.array_equal(a, a0), k
    idx = g.choice(n, min(n, 20000), replace=False)
    assert rms_rel(a0[idx], oracle_mod.pm_accelerations(pos, m, L, nm, 1.0, idx=idx)) <= 1e-5
