"""JSONL snapshots (gadget_ng_pm_hip.snapshot) against files the reference's
own writer produced (tests/golden/snapshot/, copied from the data the
reference ships; see make_snapshot_fixture.py).  No GPU needed.

Those files come from an earlier writer without the SPH fields, so a line we
write must equal the reference line with the three SPH fields appended at
their defaults; every f64 text (7 per line, 200 lines) must be identical,
which pins the number formatting; the pretty-printed meta.json and
provenance.json must be identical byte for byte.
"""
import json
import math
import os

import numpy as np
import pytest

from gadget_ng_pm_hip import snapshot as snap

FIX = os.path.join(os.path.dirname(__file__), "golden", "snapshot")
SPH_DEFAULTS = ',"internal_energy":0.0,"smoothing_length":0.0,"ptype":"DarkMatter"}'


def _fixture_dir(tmp_path, meta_name="p16_meta.json"):
    d = tmp_path / "ref"
    d.mkdir()
    with open(os.path.join(FIX, "p16_particles_head.jsonl"), "rb") as f:
        (d / "particles.jsonl").write_bytes(f.read())
    with open(os.path.join(FIX, meta_name), "rb") as f:
        (d / "meta.json").write_bytes(f.read())
    return d


def _provenance(meta):
    p = meta["provenance"]
    return snap.Provenance(p["crate_version"], p["git_commit"], p["build_profile"], p["enabled_features"],
                           p["command_line"], p["config_hash"])


def test_fixture_checksums():
    import hashlib
    with open(os.path.join(FIX, "fixtures.json")) as f:
        sums = json.load(f)
    for name, entry in sums.items():
        with open(os.path.join(FIX, name), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == entry["sha256"], name


def test_reads_reference_snapshot_with_serde_defaults(tmp_path):
    d = _fixture_dir(tmp_path)
    s = snap.read_jsonl(str(d))
    assert s.pos.shape == (200, 3) and s.vel.shape == (200, 3)
    assert np.array_equal(s.global_id, np.arange(200))
    assert s.time == 0.125 and s.redshift == 0.0 and s.box_size == 20.0
    assert np.all(s.internal_energy == 0.0) and np.all(s.smoothing_length == 0.0) and np.all(s.ptype == 0)
    assert s.mass[0] == 0.0005 and s.pos[0, 0] == 0.20255550697565908


def test_rewrite_reproduces_the_reference_bytes(tmp_path):
    d = _fixture_dir(tmp_path)
    s = snap.read_jsonl(str(d))
    with open(os.path.join(FIX, "p16_meta.json")) as f:
        meta = json.load(f)
    out = tmp_path / "out"
    snap.write_jsonl(str(out), s.pos, s.vel, s.mass, s.global_id, time=s.time, redshift=s.redshift,
                     box_size=s.box_size, provenance=_provenance(meta))
    ref_lines = (d / "particles.jsonl").read_text().splitlines()
    got_lines = (out / "particles.jsonl").read_text().splitlines()
    assert len(got_lines) == len(ref_lines) == 200
    for r, g in zip(ref_lines, got_lines):
        assert g == r[:-1] + SPH_DEFAULTS
    # the fixture keeps the first 200 of the snapshot's 2000 particles
    text = (out / "meta.json").read_text().replace('"particle_count": 200,', '"particle_count": 2000,')
    with open(os.path.join(FIX, "p16_meta.json")) as f:
        assert text == f.read()
    with open(os.path.join(FIX, "p16_provenance.json"), "rb") as f:
        assert (out / "provenance.json").read_bytes() == f.read()


def test_meta_with_empty_feature_list_is_byte_identical(tmp_path):
    with open(os.path.join(FIX, "p38_meta.json")) as f:
        meta = json.load(f)
    out = tmp_path / "o"
    snap.write_jsonl(str(out), np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0), time=meta["time"],
                     redshift=meta["redshift"], box_size=meta["box_size"], provenance=_provenance(meta))
    got = json.loads((out / "meta.json").read_text())
    assert got["particle_count"] == 0
    # the rest of the file is the reference's byte for byte once the count agrees
    text = (out / "meta.json").read_text().replace('"particle_count": 0', '"particle_count": 32768')
    with open(os.path.join(FIX, "p38_meta.json")) as f:
        assert text == f.read()


@pytest.mark.parametrize("x,s", [(0.0, "0.0"), (-0.0, "-0.0"), (1.0, "1.0"), (-1.5, "-1.5"), (100.0, "100.0"),
                                 (1e15, "1000000000000000.0"), (1e16, "1e16"), (1.5e16, "1.5e16"),
                                 (123456789012345680.0, "1.2345678901234568e17"), (0.001, "0.001"),
                                 (1.234e-5, "0.00001234"), (1e-5, "0.00001"), (1e-6, "1e-6"),
                                 (1.5e-7, "1.5e-7"), (5e-324, "5e-324"),
                                 (1.7976931348623157e308, "1.7976931348623157e308"), (0.1, "0.1"),
                                 (float("nan"), "null"), (float("inf"), "null")])
def test_format_follows_ryu_rules(x, s):
    assert snap.format_f64(x) == s


def test_format_round_trips():
    g = np.random.default_rng(0)
    bits = g.integers(0, 2 ** 63 - 1, 20000, dtype=np.int64).view(np.float64)
    vals = [v for v in bits if math.isfinite(v)] + list(g.normal(size=2000) * 10.0 ** g.integers(-30, 30, 2000))
    for v in vals:
        assert float(snap.format_f64(v)) == v


def test_round_trip_of_a_generated_set(tmp_path):
    g = np.random.default_rng(1)
    n = 500
    pos, vel = g.uniform(0, 50, (n, 3)), g.normal(size=(n, 3))
    mass = np.full(n, 1.0 / n)
    t = g.integers(0, 3, n)
    u = np.where(t == 1, g.uniform(0, 2, n), 0.0)
    units = {"length_in_kpc": 1000.0, "mass_in_msun": 1e10, "velocity_in_km_s": 1.0, "time_in_gyr": 977.79,
             "g_internal": 43007.1}
    snap.write_jsonl(str(tmp_path), pos, vel, mass, np.arange(10, 10 + n), time=0.5, redshift=1.0, box_size=50.0,
                     units=units, internal_energy=u, ptype=t)
    s = snap.read_jsonl(str(tmp_path))
    assert np.array_equal(s.pos, pos) and np.array_equal(s.vel, vel) and np.array_equal(s.mass, mass)
    assert np.array_equal(s.global_id, np.arange(10, 10 + n)) and np.array_equal(s.ptype, t)
    assert np.array_equal(s.internal_energy, u) and s.units == units
    assert (s.time, s.redshift, s.box_size) == (0.5, 1.0, 50.0)
    with pytest.raises(ValueError):
        snap.write_jsonl(str(tmp_path / "bad"), pos, vel, mass[:-1])
