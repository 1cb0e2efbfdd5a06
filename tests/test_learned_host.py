"""Host-side bookkeeping of the learned-correction evaluation (no GPU)."""
import numpy as np

from gadget_ng_pm_hip.learned import force_errors


def test_force_error_populations_partition_the_particles():
    g = np.random.default_rng(4)
    n = 1000
    t = g.normal(size=(n, 3))
    f = t + 0.1 * g.normal(size=(n, 3))
    field = g.random(n) < 0.7
    near = g.random(n) < 0.2
    e = force_errors(f, t, 600, field, near)
    assert e["count"] == n
    assert e["lattice_count"] + e["sphere_count"] == n
    assert e["field_count"] + e["dense_count"] == n
    assert e["field_near_count"] + e["field_far_count"] == e["field_count"]
    # squared errors and squared forces add over a partition
    d2 = np.sum((f - t) ** 2, axis=1)
    t2 = np.sum(t ** 2, axis=1)
    for tag, sel in (("field_near_", field & near), ("field_far_", field & ~near), ("dense_", ~field)):
        assert np.isclose(e[tag + "rms_rel"] ** 2, d2[sel].sum() / t2[sel].sum())
    assert np.isclose(e["rms_rel"] ** 2, d2.sum() / t2.sum())
