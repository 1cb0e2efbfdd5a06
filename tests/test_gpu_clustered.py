"""Clustered inputs (TreePM long range at a clustered shape): a Zel'dovich lattice
plus Plummer spheres, with the Gaussian split r_s = 1.25 cells in the
reference convention exp(-k^2 r_s^2 / 2) (crates/gadget-ng-pm/src/fft_poisson.rs:528-531).
Heavy clusters stress the deposit (thousands of particles per cell)."""
import numpy as np
import pytest

from helpers import max_rel_robust, rms_rel

pytestmark = pytest.mark.gpu


def test_small_clustered_long_range_matches_oracle(pm, oracle_mod):
    from gadget_ng_pm_hip import workloads
    p = workloads.clustered_box(n_total=64 ** 3, lattice_side=56, n_spheres=40, box=64.0, nm=128, workers=-1)
    rs = 1.25 * p.box / 128
    print(f"case clustered small N={p.n} nm=128", flush=True)
    plan = pm.Plan(128, p.box, r_split=rs, keep_density=True)
    got = plan.accelerations_f32(p.x, p.y, p.z, p.m, 1.0).T.astype(np.float64)
    ref, dens = oracle_mod.pm_accelerations(p.pos64(), p.mass64(), p.box, 128, 1.0, r_split=rs, want_density=True)
    assert rms_rel(got, ref) <= 1e-5
    assert max_rel_robust(got, ref) <= 1e-4
    d = plan.density()
    assert np.abs(d - dens).max() <= 1e-6 * dens.max()
    assert abs(d.sum() - p.n) <= 1e-6 * p.n


def test_full_clustered_512cubed_on_1024cubed(pm, oracle_mod):
    from gadget_ng_pm_hip import workloads
    p = workloads.clustered_box(workers=-1)
    rs = 1.25 * p.box / 1024
    print(f"case clustered full N={p.n} nm=1024", flush=True)
    plan = pm.Plan(1024, p.box, r_split=rs)
    got = plan.accelerations_f32(p.x, p.y, p.z, p.m, 1.0).T.astype(np.float64)
    again = plan.accelerations_f32(p.x, p.y, p.z, p.m, 1.0).T.astype(np.float64)
    assert np.array_equal(got, again)
    plan.close()
    ref = oracle_mod.pm_accelerations(p.pos64(), p.mass64(), p.box, 1024, 1.0, r_split=rs)
    assert rms_rel(got, ref) <= 1e-5
    assert max_rel_robust(got, ref) <= 1e-3
