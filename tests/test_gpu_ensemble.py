"""Ensembles of independent boxes solved together through one captured HIP graph."""
import numpy as np
import pytest

from helpers import rms_rel

pytestmark = pytest.mark.gpu


def _run(pm, batch, x, y, z, m, off):
    n = x.shape[0]
    d = [pm.DeviceArray.from_host(a) for a in (x, y, z, m)]
    out = [pm.DeviceArray((n,), np.float32) for _ in range(3)]
    batch.solve_device(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, off, 1.0, out[0].ptr, out[1].ptr, out[2].ptr)
    pm.device_synchronize()
    a1 = np.stack([o.numpy() for o in out], axis=1).astype(np.float64)
    batch.solve_device(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, off, 1.0, out[0].ptr, out[1].ptr, out[2].ptr)
    pm.device_synchronize()
    a2 = np.stack([o.numpy() for o in out], axis=1).astype(np.float64)
    return a1, a2


def test_small_ensemble_each_box_matches_oracle(pm, oracle_mod):
    from gadget_ng_pm_hip import workloads
    boxes, off, x, y, z, m = workloads.ensemble(n_boxes=10, sides=(8, 12, 16, 24, 32), workers=-1)
    print(f"case ensemble small boxes={len(boxes)} N={x.shape[0]}", flush=True)
    batch = pm.Batch(boxes)
    a1, a2 = _run(pm, batch, x, y, z, m, off)
    assert np.array_equal(a1, a2), "graph replays must agree bit for bit"
    for b, bx in enumerate(boxes):
        p = bx["particles"]
        ref = oracle_mod.pm_accelerations(p.pos64(), p.mass64(), bx["box_size"], bx["grid_size"], 1.0)
        got = a1[off[b]:off[b + 1]]
        assert rms_rel(got, ref) <= 1e-5, (b, bx["grid_size"])


def test_ensemble_recaptures_when_offsets_change(pm, oracle_mod):
    from gadget_ng_pm_hip import workloads
    boxes, off, x, y, z, m = workloads.ensemble(n_boxes=3, sides=(8, 16), workers=-1)
    batch = pm.Batch(boxes)
    a1, _ = _run(pm, batch, x, y, z, m, off)
    # drop the last particle of box 0: offsets change, the graph is captured again
    keep = np.ones(x.shape[0], bool)
    keep[off[1] - 1] = False
    off2 = off.copy()
    off2[1:] -= 1
    a2, _ = _run(pm, batch, x[keep], y[keep], z[keep], m[keep], off2)
    p0 = boxes[0]["particles"]
    ref = oracle_mod.pm_accelerations(p0.pos64()[:-1], p0.mass64()[:-1], boxes[0]["box_size"],
                                      boxes[0]["grid_size"], 1.0)
    assert rms_rel(a2[off2[0]:off2[1]], ref) <= 1e-5
    # boxes 1.. unchanged
    assert np.array_equal(a2[off2[1]:], a1[off[1]:])


def test_full_ensemble_64_boxes(pm, oracle_mod):
    from gadget_ng_pm_hip import workloads
    boxes, off, x, y, z, m = workloads.ensemble(workers=-1)
    print(f"case ensemble full boxes={len(boxes)} N={x.shape[0]}", flush=True)
    batch = pm.Batch(boxes)
    a1, a2 = _run(pm, batch, x, y, z, m, off)
    assert np.array_equal(a1, a2)
    worst = 0.0
    for b, bx in enumerate(boxes):
        p = bx["particles"]
        ref = oracle_mod.pm_accelerations(p.pos64(), p.mass64(), bx["box_size"], bx["grid_size"], 1.0)
        worst = max(worst, rms_rel(a1[off[b]:off[b + 1]], ref))
    print(f"worst per-box rms err {worst:.3e}", flush=True)
    assert worst <= 1e-5


def test_box_tables_equal_per_box_pipelines(pm):
    """One launch per stage per mesh side (box tables, the default) against
    one pipeline per box: the same kernels on the same data, bit for bit,
    also with boxes of one side holding different particle counts."""
    from gadget_ng_pm_hip import workloads
    boxes, off, x, y, z, m = workloads.ensemble(n_boxes=12, sides=(8, 12, 16, 32), workers=-1)
    # uneven counts within a mesh side: drop a few particles of every third box
    keep = np.ones(x.shape[0], bool)
    off2 = off.copy()
    for b in range(0, len(boxes), 3):
        drop = min(5 * (b + 1), int(off[b + 1] - off[b]) - 1)
        keep[off[b + 1] - drop:off[b + 1]] = False
        off2[b + 1:] -= drop
    print(f"case ensemble box tables vs per-box boxes={len(boxes)} N={int(keep.sum())}", flush=True)
    batch = pm.Batch(boxes)
    a_tab, _ = _run(pm, batch, x[keep], y[keep], z[keep], m[keep], off2)
    batch.set_box_tables(False)
    a_box, _ = _run(pm, batch, x[keep], y[keep], z[keep], m[keep], off2)
    assert np.abs(a_tab).max() > 0.0
    assert np.array_equal(a_tab, a_box)


def test_launch_split_tuner_keeps_a_valid_split_and_the_result(pm):
    from gadget_ng_pm_hip import workloads
    boxes, off, x, y, z, m = workloads.ensemble(n_boxes=9, sides=(8, 16), workers=-1)
    n = x.shape[0]
    print(f"case ensemble launch-split tuner boxes={len(boxes)} N={n}", flush=True)
    batch = pm.Batch(boxes)
    ref, _ = _run(pm, batch, x, y, z, m, off)
    d = [pm.DeviceArray.from_host(a) for a in (x, y, z, m)]
    out = [pm.DeviceArray((n,), np.float32) for _ in range(3)]
    k = batch.tune_launches(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, off, 1.0, out[0].ptr, out[1].ptr, out[2].ptr,
                            reps=2)
    assert k >= 0 and batch.box_tables() == k
    got, _ = _run(pm, batch, x, y, z, m, off)
    assert np.array_equal(got, ref)
