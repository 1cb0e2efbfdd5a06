"""Power spectrum on the GPU (gng_pm_power_spectrum) against the FP64 oracle
restatement of crates/gadget-ng-analysis/src/power_spectrum.rs.

Shell wavenumbers and mode counts are integers and must agree exactly; the
power agrees within 5e-6 relative per shell (FP32 transform, f32 mass per
cell) -- except for shells whose power is at the FP32 noise floor of the
transform, which are held to 1e-6 of the largest shell instead."""
import math

import numpy as np
import pytest

import analysis_oracle as ao
from helpers import f32_round

pytestmark = pytest.mark.gpu


def _check(got, ref, tol=5e-6):
    k, pk, nm = got
    rk, rpk, rnm = ref
    assert np.array_equal(nm, rnm)
    np.testing.assert_allclose(k, rk, rtol=1e-15)
    floor = 1e-6 * rpk.max()
    err = np.abs(pk - rpk) / np.maximum(rpk, floor)
    print(f"max shell error {err.max():.2e} over {len(k)} shells", flush=True)
    assert err.max() < tol


@pytest.mark.parametrize("mesh,n,box", [(12, 3000, 6.0), (32, 20000, 32.0), (64, 200000, 100.0),
                                        (128, 400000, 128.0)])
def test_uniform_random_particles(pm, mesh, n, box):
    from gadget_ng_pm_hip import analysis
    print(f"case pk uniform mesh={mesh} N={n}", flush=True)
    rng = np.random.default_rng(mesh)
    pos = f32_round(rng.uniform(0, box, (n, 3)))
    m = f32_round(rng.uniform(0.5, 1.5, n))
    plan = pm.Plan(mesh, box)
    _check(analysis.power_spectrum_arrays(plan, pos, m), ao.power_spectrum(pos, m, box, mesh))


def test_zeldovich_lattice(pm):
    from gadget_ng_pm_hip import analysis, workloads
    print("case pk zeldovich 64^3 on 128", flush=True)
    p = workloads.zeldovich_lattice(n_side=64, nm=128, box=64.0, spectral_index=-2.0, rms_disp_cells=0.5, seed=3)
    plan = pm.Plan(128, 64.0)
    _check(analysis.power_spectrum_arrays(plan, p.pos64(), p.mass64()),
           ao.power_spectrum(p.pos64(), p.mass64(), 64.0, 128))


def test_clustered_particles_on_the_device_path(pm):
    from gadget_ng_pm_hip import analysis
    print("case pk clustered device pointers mesh=64", flush=True)
    rng = np.random.default_rng(9)
    box = 50.0
    cen = rng.uniform(0, box, (20, 3))
    pos = np.concatenate([rng.uniform(0, box, (30000, 3)),
                          cen[rng.integers(0, 20, 30000)] + rng.normal(scale=0.5, size=(30000, 3))])
    pos = f32_round(np.mod(pos, box))
    pos[pos >= box] = 0.0
    m = np.ones(len(pos))
    plan = pm.Plan(64, box)
    d = [pm.DeviceArray.from_host(np.ascontiguousarray(pos[:, a], dtype=np.float32)) for a in range(3)]
    dm = pm.DeviceArray.from_host(m.astype(np.float32))
    got = analysis.power_spectrum_device(plan, d[0].ptr, d[1].ptr, d[2].ptr, dm.ptr, len(pos))
    _check(got, ao.power_spectrum(pos, m, box, 64))
    # the same plan still solves forces afterwards
    a = plan.accelerations_f64(pos[:100], m[:100], 1.0)
    assert np.isfinite(a).all()


def test_uniform_lattice_and_single_particle(pm):
    from gadget_ng_pm_hip import analysis
    print("case pk lattice / one particle", flush=True)
    n, box = 8, 2.0
    cell = box / n
    g = (np.arange(n) + 0.5) * cell
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    bins = analysis.power_spectrum(pos, np.ones(len(pos)), box, n)
    assert len(bins) == n // 2 and all(b.pk < 1e-20 for b in bins)
    # one particle on a node: every shell holds V^2 <1/W^2>
    one = np.array([[3 * cell, 1 * cell, 6 * cell]])
    got = analysis.power_spectrum_arrays(pm.Plan(n, box), one, np.array([1.5]))
    _check(got, ao.power_spectrum(one, np.array([1.5]), box, n), tol=1e-6)


def test_repeatable_and_errors(pm):
    from gadget_ng_pm_hip import analysis
    rng = np.random.default_rng(4)
    pos = f32_round(rng.uniform(0, 10.0, (5000, 3)))
    m = np.ones(5000)
    plan = pm.Plan(32, 10.0)
    a = analysis.power_spectrum_arrays(plan, pos, m)
    b = analysis.power_spectrum_arrays(plan, pos, m)
    assert np.array_equal(a[2], b[2]) and np.allclose(a[1], b[1], rtol=1e-13, atol=0)
    with pytest.raises(pm.HipExecutionError):
        analysis.power_spectrum_arrays(plan, np.zeros((0, 3)), np.zeros(0))
