"""Seeded workload generators and the algorithmic-byte model (CPU only)."""
import numpy as np

from gadget_ng_pm_hip import workloads


def test_algorithmic_bytes_match_the_stated_model():
    assert workloads.algorithmic_bytes(256 ** 3, 512) == 9_277_800_448
    assert workloads.algorithmic_bytes(512 ** 3, 1024) == 74_155_294_720


def test_uniform_box_is_seeded_and_in_range():
    a = workloads.uniform_box(1000, 100.0)
    b = workloads.uniform_box(1000, 100.0)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.z, b.z)
    for c in (a.x, a.y, a.z):
        assert c.dtype == np.float32 and c.min() >= 0 and c.max() < 100.0
    assert np.allclose(a.m.sum(), 1.0, rtol=1e-5)


def test_zeldovich_lattice_shape_and_displacement():
    p = workloads.zeldovich_lattice(n_side=16, nm=32, box=16.0, workers=1)
    q = np.arange(16.0)
    d = [((c.reshape(16, 16, 16) - qq + 8.0) % 16.0) - 8.0
         for c, qq in ((p.x, q[None, None, :]), (p.y, q[None, :, None]), (p.z, q[:, None, None]))]
    rms = np.sqrt(sum((di ** 2 for di in d)).mean()) / (16.0 / 32)
    assert abs(rms - 0.5) < 1e-3  # 3-D rms displacement of 0.5 cell (float32 rounding aside)
    p2 = workloads.zeldovich_lattice(n_side=16, nm=32, box=16.0, workers=1)
    assert np.array_equal(p.x, p2.x)
    assert p.x.min() >= 0 and p.x.max() < 16.0


def test_ap08_curve_matches_reference_bench_input():
    """crates/gadget-ng-cuda/benches/cuda_vs_simd.rs:17-31: (sin t, cos t, sin t/2)*10, t = 0.37 i."""
    p = workloads.ap08_curve(32)
    i = 7
    t = 0.37 * i
    assert p.n == 1024
    assert np.isclose(p.x[i], np.sin(t) * 10, atol=1e-5) and np.isclose(p.z[i], np.sin(t / 2) * 10, atol=1e-5)
