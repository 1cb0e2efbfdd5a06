"""The deposit's particle order (gng_pm_deposit_order) against a CPU
restatement of its bucket key: the Hilbert rank of the particle's deposit
tile times 8 plus the particle's class inside the tile.  The tile ranks come
from gng_hilbert_key3 evaluated on the host (the curve itself is pinned in
tests/test_abi.py), the cell from the FP64 cell coordinate with its 43-bit
fraction (tests/test_deposit_scheme.py::cell_fix)."""
import numpy as np
import pytest

from test_deposit_scheme import cell_fix

pytestmark = pytest.mark.gpu


def deposit_tile(nm):
    return 16 if nm % 16 == 0 else 12 if nm % 12 == 0 else 8 if nm % 8 == 0 else 4


def expected_buckets(pm, pos, box, nm):
    T = deposit_tile(nm)
    nt = nm // T
    bits = 0
    while (1 << bits) < nt:
        bits += 1
    keys = np.array([pm.lib().gng_hilbert_key3(tx, ty, tz, bits)
                     for tz in range(nt) for ty in range(nt) for tx in range(nt)], dtype=np.uint64)
    rank = np.empty(nt ** 3, np.int64)
    rank[np.argsort(keys, kind="stable")] = np.arange(nt ** 3)
    c = [cell_fix(pos[:, a], box, nm)[0] for a in range(3)]
    tile = ((c[2] // T) * nt + c[1] // T) * nt + c[0] // T
    cls = (c[0] % T == T - 1).astype(np.int64) | ((c[1] % T == T - 1).astype(np.int64) << 1) | \
        ((c[2] % T == T - 1).astype(np.int64) << 2)
    return rank[tile] * 8 + cls


@pytest.mark.parametrize("nm,box", [(64, 64.0), (48, 30.0), (96, 200.0)])
def test_deposit_order_is_bucket_sorted(pm, nm, box):
    rng = np.random.default_rng(nm)
    n = 120_000
    pos = rng.uniform(0.0, box, (n, 3))
    pos[: n // 5] = rng.normal(0.37 * box, 0.02 * box, (n // 5, 3))   # a clump over a few tiles
    pos[n // 5: n // 5 + 2000] += box * rng.integers(-2, 3, (2000, 3))  # wrapped inputs
    m = rng.uniform(0.5, 1.5, n)
    print(f"case deposit order nm={nm} n={n}", flush=True)
    plan = pm.Plan(nm, box)
    order = plan.deposit_order(pos, m).astype(np.int64)
    assert np.array_equal(np.sort(order), np.arange(n))
    key = expected_buckets(pm, pos, box, nm)
    assert np.all(np.diff(key[order]) >= 0)
    # the deposit's forces are unchanged by the query
    ref = plan.accelerations_f64(pos, m, 1.0)
    plan.deposit_order(pos[::-1].copy(), m[::-1].copy())
    assert np.array_equal(plan.accelerations_f64(pos, m, 1.0), ref)
