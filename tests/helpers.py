"""Shared metrics and input builders for the tests."""
import numpy as np


def rms_rel(a, ref):
    """sqrt(sum|a - ref|^2 / sum|ref|^2) over all particles."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = float((ref ** 2).sum())
    return float(np.sqrt(((a - ref) ** 2).sum() / den)) if den > 0 else float(np.sqrt(((a - ref) ** 2).sum()))


def max_rel_robust(a, ref):
    """max |a - ref| / max(|ref|, 0.05 mean|ref|): the robust floor of the
    reference's GPU smoke metric (crates/gadget-ng-cuda/tests/cuda_pm_smoke.rs:156-162)."""
    a = np.asarray(a, dtype=np.float64).reshape(-1, 3)
    ref = np.asarray(ref, dtype=np.float64).reshape(-1, 3)
    nr = np.linalg.norm(ref, axis=1)
    floor = 0.05 * max(nr.mean(), 1e-300)
    den = np.maximum(np.maximum(nr, floor), 1e-12)
    return float((np.linalg.norm(a - ref, axis=1) / den).max())


def f32_round(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def uniform_particles(rng, n, box, lo=0.0, hi=None, mass=None):
    hi = box if hi is None else hi
    pos = f32_round(rng.uniform(lo, hi, (n, 3)))
    m = f32_round(np.full(n, 1.0 / n) if mass is None else mass)
    return pos, m
