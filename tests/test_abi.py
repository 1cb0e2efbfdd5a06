"""C ABI checks that need no GPU: the library loads, exports every symbol
include/gng_pm.h declares, and keeps the reference's error contract
(crates/gadget-ng-hip: hip_pm_create -> NULL on failure, hip_pm_solve with a
NULL handle or n <= 0 is a no-op returning 0; HipPmSolver::try_new -> None).
No compute call is made."""
import ctypes
import os
import re

import numpy as np
import pytest

import gadget_ng_pm_hip as pm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


HEADERS = ("gng_pm.h", "gng_learned.h", "gng_integrate.h", "gng_ic.h")


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b((?:gng|hip)_\w+)\s*\(", text))
    return sorted(names)


def has_device():
    return pm.lib().gng_pm_device_count() > 0


def test_header_declares_the_reference_symbols():
    names = declared_functions()
    for sym in ("hip_pm_create", "hip_pm_destroy", "hip_pm_solve"):
        assert sym in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pm.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_and_mesh_table():
    L = pm.lib()
    assert L.gng_pm_abi_version() == pm.ABI_VERSION == 2
    for nm in (4, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 1024):
        assert L.gng_pm_mesh_supported(nm) == 1, nm
    for nm in (0, 1, 2, 3, 5, 7, 10, 100, 2048, -8):
        assert L.gng_pm_mesh_supported(nm) == 0, nm


def test_unsupported_mesh_is_rejected_before_touching_a_device():
    cfg = pm.GngConfig()
    cfg.grid_size = 100
    cfg.box_size = 1.0
    cfg.device = -1
    h = ctypes.c_void_p()
    rc = pm.lib().gng_pm_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == pm.GNG_ERR_UNSUPPORTED and not h.value
    assert "not compiled in" in pm.last_error()


def test_invalid_box_is_rejected():
    cfg = pm.GngConfig()
    cfg.grid_size = 16
    cfg.box_size = -2.0
    h = ctypes.c_void_p()
    assert pm.lib().gng_pm_create(ctypes.byref(cfg), ctypes.byref(h)) == pm.GNG_ERR_INVALID


def test_null_plan_calls_return_invalid():
    L = pm.lib()
    out = np.zeros(3)
    assert L.gng_pm_accelerations_f64(None, None, None, 0, None, 0, 0.0, 1.0, out.ctypes.data) == pm.GNG_ERR_INVALID
    assert L.gng_pm_density(None, None) == pm.GNG_ERR_INVALID
    L.gng_pm_destroy(None)  # NULL-safe


def test_hip_pm_solve_null_handle_is_noop():
    """pm_gravity.hip:370: `if (!h || n <= 0) return 0;`"""
    L = pm.lib()
    a = np.zeros(4, dtype=np.float32)
    p = a.ctypes.data
    assert L.hip_pm_solve(None, p, p, p, p, p, p, p, 4, 0.0, 1.0, 0.0) == 0
    assert L.hip_pm_solve(None, p, p, p, p, p, p, p, 0, 0.0, 1.0, 0.0) == 0
    L.hip_pm_destroy(None)


@pytest.mark.skipif("has_device()")
def test_without_device_create_fails_cleanly():
    L = pm.lib()
    assert L.gng_pm_device_count() == 0
    assert not L.hip_pm_create(16, ctypes.c_float(1.0))
    cfg = pm.GngConfig()
    cfg.grid_size = 16
    cfg.box_size = 1.0
    cfg.device = -1
    h = ctypes.c_void_p()
    assert L.gng_pm_create(ctypes.byref(cfg), ctypes.byref(h)) == pm.GNG_ERR_NO_DEVICE


@pytest.mark.skipif("has_device()")
def test_without_device_solver_is_unavailable():
    av = pm.HipPmSolver.availability()
    assert av.compiled and not av.runtime_available
    assert not pm.HipPmSolver.is_available()
    assert pm.HipPmSolver.try_new(16, 1.0) is None
    assert pm.HipPmSolver.try_new_with_r_split(16, 1.0, 0.2) is None
    with pytest.raises(pm.HipUnavailable):
        pm.HipPmSolver.try_new_checked(16, 1.0)
    with pytest.raises(pm.HipExecutionError):
        pm.PmSolver(16, 1.0).accelerations_for_indices(np.zeros((1, 3)), np.ones(1), 0.0, 1.0, [0])


def test_learned_symbols_declared_and_parameter_count():
    names = declared_functions()
    for sym in ("gng_sr_accelerations", "gng_net_forward", "gng_net_backward", "gng_net_train_step",
                "gng_pm_keep_potential", "gng_dev_axpy"):
        assert sym in names
    # w1[27][2][16] b1[16] w2[27][16][16] b2[16] w3[27][16][3] b3[3]
    assert pm.lib().gng_net_param_count() == 27 * 8 * 16 + 16 + 27 * 16 * 16 + 16 + 27 * 16 * 3 + 3


@pytest.mark.skipif("has_device()")
def test_learned_objects_need_a_device():
    L = pm.lib()
    h = ctypes.c_void_p()
    assert L.gng_net_create(16, 1, 0, ctypes.byref(h)) == pm.GNG_ERR_NO_DEVICE and not h.value
    assert L.gng_sr_create(0, ctypes.byref(h)) == pm.GNG_ERR_NO_DEVICE and not h.value
    assert L.gng_net_create(24, 1, 0, ctypes.byref(h)) == pm.GNG_ERR_INVALID
    L.gng_net_destroy(None)
    L.gng_sr_destroy(None)


def test_config_struct_layout_matches_header():
    # int32, int32, double, double, double, int32, int32[7] -> 64 bytes with padding
    assert ctypes.sizeof(pm.GngConfig) == 64
    assert pm.GngConfig.box_size.offset == 8 and pm.GngConfig.device.offset == 32


def test_treepm_effective_r_split():
    lr = pm.TreePmLongRange(64, 128.0)
    assert lr.effective_r_split() == pytest.approx(2.5 * 128.0 / 64)
    assert pm.TreePmLongRange(64, 128.0, 0.7).effective_r_split() == 0.7


def test_hilbert_key_first_order_matches_reference_doc():
    """crates/gadget-ng-parallel/src/sfc.rs:113-116 lists the p = 1 visiting order."""
    order = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (1, 0, 0)]
    assert [pm.hilbert_key3(*c, 1) for c in order] == list(range(8))


@pytest.mark.parametrize("bits", [2, 3, 4, 5])
def test_hilbert_curve_is_a_face_adjacent_bijection(bits):
    n = 1 << bits
    keys = {}
    for z in range(n):
        for y in range(n):
            for x in range(n):
                keys[pm.hilbert_key3(x, y, z, bits)] = (x, y, z)
    assert sorted(keys) == list(range(n ** 3))
    for k in range(n ** 3 - 1):
        a, b = keys[k], keys[k + 1]
        assert sum(abs(u - v) for u, v in zip(a, b)) == 1, (k, a, b)


def _reference_coords_to_hilbert(ix, iy, iz, p):
    """Restatement of the reference's coords_to_hilbert
    (crates/gadget-ng-parallel/src/sfc.rs:118-183): axes visited in reverse in
    the AxesToTranspose loop (:134), then Gray encode and the XOR correction,
    then the transposed bits packed from level k = 0 (least significant)
    upward (:177-181)."""
    x = [ix, iy, iz]
    m = 1 << (p - 1)
    q = m
    while q > 1:
        pm_ = q - 1
        for i in (2, 1, 0):
            if x[i] & q:
                x[0] ^= pm_
            else:
                t = (x[0] ^ x[i]) & pm_
                x[0] ^= t
                x[i] ^= t
        q >>= 1
    for i in (1, 2):
        x[i] ^= x[i - 1]
    t, q = 0, m
    while q > 1:
        if x[2] & q:
            t ^= q - 1
        q >>= 1
    x = [v ^ t for v in x]
    h = 0
    for k in range(p):
        for v in x:
            h = (h << 1) | ((v >> k) & 1)
    return h


def test_reference_hilbert_key_diverges_beyond_first_order():
    """The reference curve equals this key at p = 1 and differs beyond it;
    its own order at p >= 2 is not face-adjacent (a reference defect, recorded
    so that nobody takes the two keys as interchangeable)."""
    order = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (1, 0, 0)]
    assert [_reference_coords_to_hilbert(*c, 1) for c in order] == list(range(8))
    for p, differ in ((2, 59), (3, 493)):
        n = 1 << p
        cells = [(x, y, z) for z in range(n) for y in range(n) for x in range(n)]
        ref = {c: _reference_coords_to_hilbert(*c, p) for c in cells}
        ours = {c: pm.hilbert_key3(*c, p) for c in cells}
        assert sorted(ref.values()) == list(range(n ** 3))  # still a bijection
        assert sum(ref[c] != ours[c] for c in cells) == differ
        inv = {v: c for c, v in ref.items()}
        jumps = sum(1 for k in range(n ** 3 - 1)
                    if sum(abs(a - b) for a, b in zip(inv[k], inv[k + 1])) != 1)
        assert jumps > 0
