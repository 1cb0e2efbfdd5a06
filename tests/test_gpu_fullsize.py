"""Full-size headline workload: 256^3 Zel'dovich particles on a 512^3 mesh.

The FP64 oracle runs the whole case in seconds on the host, so parity is
checked directly (rms <= 1e-5), plus size-independent properties: bitwise
reproducibility of density and forces, momentum conservation, and equivariance
under a translation by whole cells."""
import numpy as np
import pytest

from helpers import max_rel_robust, rms_rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def headline():
    from gadget_ng_pm_hip import workloads
    return workloads.zeldovich_lattice(256, 512, 256.0, workers=-1, keep_f64=True)


def test_headline_parity_with_fp64(pm, oracle_mod, headline):
    p = headline
    print(f"case headline N={p.n} nm=512", flush=True)
    plan = pm.Plan(512, p.box)
    got = plan.accelerations_f32(p.x, p.y, p.z, p.m, 1.0).T.astype(np.float64)
    ref = oracle_mod.pm_accelerations(p.pos64(), p.mass64(), p.box, 512, 1.0)
    assert rms_rel(got, ref) <= 1e-5
    assert max_rel_robust(got, ref) <= 1e-3
    # momentum: sum m a vanishes to FP32 rounding of the force field
    s = np.abs(got.sum(axis=0))
    assert s.max() <= 1e-5 * np.abs(got).sum(axis=0).max()
    # bitwise reproducible
    again = plan.accelerations_f32(p.x, p.y, p.z, p.m, 1.0).T.astype(np.float64)
    assert np.array_equal(got, again)


def test_headline_unrounded_f64_parity(pm, oracle_mod, headline):
    """The same workload on its unrounded f64 positions through the f64 entry
    point: rounding to f32 alone costs about 2e-5 rms here."""
    p = headline
    print(f"case headline f64 N={p.n} nm=512", flush=True)
    plan = pm.Plan(512, p.box)
    m = p.mass64()
    got = plan.accelerations_f64(p.p64, m, 1.0)
    ref = oracle_mod.pm_accelerations(p.p64, m, p.box, 512, 1.0)
    r = rms_rel(got, ref)
    print(f"f64 path rms {r:.3e}, max {max_rel_robust(got, ref):.3e}", flush=True)
    assert r <= 1e-5
    assert max_rel_robust(got, ref) <= 1e-3
    again = plan.accelerations_f64(p.p64, m, 1.0)
    assert np.array_equal(got, again)


def test_headline_translation_by_whole_cells(pm, headline):
    """Positions quantised to multiples of 2^-15 (exact in float32 below 512) are
    shifted by whole cells (nm / L = 2 per unit): the fixed-point density must
    roll bit for bit and the forces follow to FP32 rounding."""
    p = headline
    q = [np.round(c.astype(np.float64) * 2.0 ** 15) / 2.0 ** 15 % p.box for c in (p.x, p.y, p.z)]
    x0 = [c.astype(np.float32) for c in q]
    plan = pm.Plan(512, p.box, keep_density=True)
    a0 = plan.accelerations_f32(x0[0], x0[1], x0[2], p.m, 1.0)
    d0 = plan.density()
    sh = (1.0, 2.0, 3.0)  # 2, 4, 6 cells
    x1 = [np.mod(c + s, p.box).astype(np.float32) for c, s in zip(q, sh)]
    a1 = plan.accelerations_f32(x1[0], x1[1], x1[2], p.m, 1.0)
    d1 = plan.density()
    assert np.array_equal(np.roll(d0, (6, 4, 2), axis=(0, 1, 2)), d1)
    assert rms_rel(a1.T, a0.T) <= 2e-6
