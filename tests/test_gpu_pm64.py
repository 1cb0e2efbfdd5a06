"""The FP64 GPU form of the reference algorithm against the FP64 oracle."""
import numpy as np
import pytest

from helpers import rms_rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nm", [4, 8, 16, 32, 64])
def test_fp64_form_matches_oracle(pm, oracle_mod, nm):
    print(f"case fp64 form nm={nm}", flush=True)
    g = np.random.default_rng(nm)
    L = 7.0
    n = 4 * nm ** 3 // 2 + 5
    pos = g.uniform(-0.5, L + 0.5, (n, 3))  # also exercises the periodic wrap
    m = g.uniform(0.5, 1.5, n)
    got = pm.Plan64(nm, L, device=0).accelerations(pos, m, 1.3)
    ref = oracle_mod.pm_accelerations(pos, m, L, nm, g=1.3, mode=oracle_mod.MODE_FAITHFUL)
    assert rms_rel(got, ref) < 1e-11


def test_fp64_form_filters_and_subset(pm, oracle_mod):
    g = np.random.default_rng(9)
    nm, L, n = 32, 5.0, 20000
    pos = g.uniform(0, L, (n, 3))
    m = g.uniform(0.5, 1.5, n)
    idx = g.choice(n, 300, replace=False)
    got = pm.Plan64(nm, L, r_split=0.4, plummer_eps=0.1, deconvolve=True, device=0).accelerations(pos, m, 1.0, idx)
    ref = oracle_mod.pm_accelerations(pos, m, L, nm, idx=idx, r_split=0.4, plummer_eps=0.1, deconvolve=True,
                                      mode=oracle_mod.MODE_FAITHFUL)
    assert rms_rel(got, ref) < 1e-11


def test_fp64_form_agrees_with_fp32_product(pm):
    g = np.random.default_rng(10)
    nm, L, n = 64, 10.0, 100000
    pos = g.uniform(0, L, (n, 3)).astype(np.float32).astype(np.float64)
    m = np.ones(n)
    a64 = pm.Plan64(nm, L, device=0).accelerations(pos, m, 1.0)
    a32 = pm.Plan(nm, L, device=0).accelerations_f64(pos, m, 1.0)
    assert rms_rel(a32, a64) < 1e-5
