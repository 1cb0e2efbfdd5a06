"""HIP product path vs the FP64 CPU oracle on the same f32-rounded inputs.

Tolerance: the FP32 pipeline's rms relative force error must be <= 1e-5
(the target of the build) and the robust max relative error <= 1e-4.
"""
import numpy as np
import pytest

from helpers import f32_round, max_rel_robust, rms_rel, uniform_particles

pytestmark = pytest.mark.gpu

RMS_TOL = 1e-5
MAX_TOL = 1e-4


@pytest.mark.parametrize("nm", [4, 8, 12, 16, 24, 32, 48, 64, 96, 128])
def test_random_uniform_matches_oracle(pm, oracle_mod, nm):
    rng = np.random.default_rng(1000 + nm)
    box = 100.0
    n = max(64, nm ** 3 // 4)
    pos, m = uniform_particles(rng, n, box)
    print(f"case nm={nm} n={n}", flush=True)
    got = pm.Plan(nm, box).accelerations_f64(pos, m, 1.0)
    ref = oracle_mod.pm_accelerations(pos, m, box, nm, 1.0)
    assert rms_rel(got, ref) <= RMS_TOL
    assert max_rel_robust(got, ref) <= MAX_TOL


@pytest.mark.parametrize("r_split,eps,deconv", [(2.0, 0.0, False), (0.0, 1.5, False), (0.0, 0.0, True),
                                                (3.1, 0.7, True)])
def test_filters_match_oracle(pm, oracle_mod, r_split, eps, deconv):
    rng = np.random.default_rng(7)
    nm, box, n = 32, 64.0, 6000
    pos, m = uniform_particles(rng, n, box)
    m = f32_round(rng.uniform(0.5, 2.0, n))
    print(f"case r_split={r_split} eps={eps} deconv={deconv}", flush=True)
    got = pm.Plan(nm, box, r_split=r_split, plummer_eps=eps, deconvolve=deconv).accelerations_f64(pos, m, 2.5)
    ref = oracle_mod.pm_accelerations(pos, m, box, nm, 2.5, r_split=r_split, plummer_eps=eps, deconvolve=deconv)
    assert rms_rel(got, ref) <= RMS_TOL
    assert max_rel_robust(got, ref) <= MAX_TOL


def test_wrapping_and_subset(pm, oracle_mod):
    rng = np.random.default_rng(11)
    nm, box, n = 16, 10.0, 3000
    pos = f32_round(rng.uniform(-1.5 * box, 2.5 * box, (n, 3)))
    m = f32_round(rng.uniform(0.1, 3.0, n))
    idx = rng.choice(n, 257, replace=False)
    got = pm.Plan(nm, box).accelerations_f64(pos, m, 1.0, idx=idx)
    ref = oracle_mod.pm_accelerations(pos, m, box, nm, 1.0, idx=idx)
    assert got.shape == (257, 3)
    assert rms_rel(got, ref) <= RMS_TOL


def test_density_bit_identical_and_close_to_oracle(pm, oracle_mod):
    rng = np.random.default_rng(5)
    nm, box, n = 32, 50.0, 20000
    pos, m = uniform_particles(rng, n, box)
    plan = pm.Plan(nm, box, keep_density=True)
    plan.accelerations_f64(pos, m, 1.0)
    q1 = plan.density()
    perm = rng.permutation(n)
    plan.accelerations_f64(pos[perm], m[perm], 1.0)
    q2 = plan.density()
    assert np.array_equal(q1, q2), "the density must not depend on particle order"
    dens = q2
    ref = oracle_mod.cic_assign(pos, m, box, nm)
    assert abs(dens.sum() - m.sum()) <= 1e-6 * m.sum()
    assert np.abs(dens - ref).max() <= 1e-6 * np.abs(ref).max()
