"""The allocation-free entry point (gng_pm_solve_device_noalloc) against the
plan-based gng_pm_solve_device, through the C ABI: two calls in a row on one
GPU, the second smaller (and on another mesh size), each bit-identical to the
plan's result with the same (default) launch choices; sizes beyond the
documented limits are refused before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DEFAULTS = dict(z_persistent=False, deposit_threads=512, y_fwd_persistent=False, y_inv_persistent=False,
                interp_sorted=False, chunk_fwd=0, chunk_inv=0)


def _case(pm, nm, box, n, seed, r_split=0.0, with_idx=False):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5 * box, 1.5 * box, (n, 3)).astype(np.float32)
    m = rng.uniform(0.5, 2.0, n).astype(np.float32)
    d = [pm.DeviceArray.from_host(np.ascontiguousarray(pos[:, c])) for c in range(3)]
    dm = pm.DeviceArray.from_host(m)
    if with_idx:
        sel = rng.choice(n, n // 3, replace=False).astype(np.int64)
        didx, k = pm.DeviceArray.from_host(sel), sel.shape[0]
    else:
        didx, k = None, n
    out_a = [pm.DeviceArray((max(k, 1),), np.float32) for _ in range(3)]
    out_b = [pm.DeviceArray((max(k, 1),), np.float32) for _ in range(3)]
    plan = pm.Plan(nm, box, r_split=r_split)
    plan.set_tuning(**DEFAULTS)
    plan.solve_device(*(a.ptr for a in d), dm.ptr, n, *(o.ptr for o in out_a), 1.5,
                      idx_ptr=didx.ptr if didx else None, k=k)
    pm.solve_device_noalloc(nm, box, *(a.ptr for a in d), dm.ptr, n, *(o.ptr for o in out_b), 1.5, r_split=r_split,
                            idx_ptr=didx.ptr if didx else None, k=k)
    pm.device_synchronize()
    a = np.stack([o.numpy()[:k] for o in out_a])
    b = np.stack([o.numpy()[:k] for o in out_b])
    plan.close()
    return a, b


def test_two_calls_in_a_row_match_the_plan_entry(pm):
    print("case noalloc nm=64 N=200000 then nm=32 N=5000 (idx subset)", flush=True)
    a1, b1 = _case(pm, 64, 80.0, 200000, 1, r_split=1.5)
    assert np.abs(a1).max() > 0.0
    assert np.array_equal(a1, b1)
    a2, b2 = _case(pm, 32, 20.0, 5000, 2, with_idx=True)
    assert np.array_equal(a2, b2)


def test_limits_are_refused(pm):
    print("case noalloc limits", flush=True)
    L = pm.lib()
    cfg = pm.GngConfig()
    cfg.grid_size, cfg.box_size, cfg.device = 128, 10.0, -1
    rc = L.gng_pm_solve_device_noalloc(ctypes.byref(cfg), None, None, None, None, 0, None, 0, 1.0, None, None, None,
                                       None)
    assert rc == pm.GNG_ERR_INVALID and b"at most" in L.gng_last_error_string()
    cfg.grid_size = 32
    rc = L.gng_pm_solve_device_noalloc(ctypes.byref(cfg), None, None, None, None, pm.NOALLOC_MAX_PARTICLES + 1, None,
                                       pm.NOALLOC_MAX_PARTICLES + 1, 1.0, None, None, None, None)
    assert rc == pm.GNG_ERR_INVALID and b"particles" in L.gng_last_error_string()
