"""JSONL snapshots in the reference's format, to exchange particle sets with
the reference CLI (crates/gadget-ng-io/src/snapshot.rs).

A snapshot is a directory with
* ``meta.json``       SnapshotMeta (:71-81): schema_version 1, provenance,
                      particle_count, time, redshift, box_size, units when set;
* ``provenance.json`` Provenance (crates/gadget-ng-io/src/provenance.rs:5-12);
* ``particles.jsonl`` one ParticleRecord per line (:16-35): global_id, mass,
                      px, py, pz, vx, vy, vz, internal_energy,
                      smoothing_length, ptype.

Both JSON files are written as serde_json's pretty printer does (two-space
indent, ``"key": value``), the particle lines as its compact printer does, and
every f64 as serde_json prints it: the shortest digits that round-trip
(Python's repr gives the same digits as the ryu algorithm serde_json uses)
laid out by ryu's rules -- integral values keep ".0", plain decimals for
decimal exponents -5 < e <= 16, otherwise d.ddde<exp> with no "+" and no
zero padding.  Reading accepts any valid JSON numbers, as serde does.

The particle arrays come from the host (for example ``Leapfrog.state()`` or
``ic.zeldovich_ics``); nothing here touches the GPU.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

PTYPES = ("DarkMatter", "Gas", "Star")


def format_f64(x: float) -> str:
    """serde_json's text for a finite f64 (ryu's shortest representation).
    Non-finite values are written as null, as serde_json does."""
    x = float(x)
    if not math.isfinite(x):
        return "null"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign = "-" if x < 0 else ""
    r = repr(abs(x))
    # split repr into its digit string and decimal exponent: value = digits * 10^k
    if "e" in r:
        mant, ex = r.split("e")
        ex = int(ex)
    else:
        mant, ex = r, 0
    if "." in mant:
        ip, fp = mant.split(".")
    else:
        ip, fp = mant, ""
    if fp == "0":
        fp = ""
    digits = (ip + fp).lstrip("0")
    k = ex - len(fp)
    # strip trailing zeros into the exponent
    stripped = digits.rstrip("0")
    k += len(digits) - len(stripped)
    digits = stripped
    n = len(digits)
    kk = n + k  # position of the decimal point relative to the digit string
    if 0 <= k and kk <= 16:
        out = digits + "0" * k + ".0"
    elif 0 < kk <= 16:
        out = digits[:kk] + "." + digits[kk:]
    elif -5 < kk <= 0:
        out = "0." + "0" * (-kk) + digits
    elif n == 1:
        out = digits + "e" + str(kk - 1)
    else:
        out = digits[0] + "." + digits[1:] + "e" + str(kk - 1)
    return sign + out


def _dump(v, indent: int, level: int) -> str:
    pad = " " * (indent * (level + 1))
    end = " " * (indent * level)
    if isinstance(v, dict):
        if not v:
            return "{}"
        items = [f"{pad}{json.dumps(k, ensure_ascii=False)}: {_dump(x, indent, level + 1)}" for k, x in v.items()]
        return "{\n" + ",\n".join(items) + "\n" + end + "}"
    if isinstance(v, (list, tuple)):
        if not v:
            return "[]"
        return "[\n" + ",\n".join(pad + _dump(x, indent, level + 1) for x in v) + "\n" + end + "]"
    if v is None:
        return "null"
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    if isinstance(v, (float, np.floating)):
        return format_f64(float(v))
    return json.dumps(v, ensure_ascii=False)


@dataclass
class Provenance:
    """crates/gadget-ng-io/src/provenance.rs:5-12."""
    crate_version: str = "gadget-ng-amd"
    git_commit: Optional[str] = None
    build_profile: str = "release"
    enabled_features: List[str] = field(default_factory=list)
    command_line: List[str] = field(default_factory=list)
    config_hash: str = ""

    def as_dict(self) -> Dict:
        return {"crate_version": self.crate_version, "git_commit": self.git_commit,
                "build_profile": self.build_profile, "enabled_features": list(self.enabled_features),
                "command_line": list(self.command_line), "config_hash": self.config_hash}


@dataclass
class Snapshot:
    """Particles as arrays (rows in file order) plus the header scalars."""
    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray
    global_id: np.ndarray
    internal_energy: np.ndarray
    smoothing_length: np.ndarray
    ptype: np.ndarray  # indices into PTYPES
    time: float = 0.0
    redshift: float = 0.0
    box_size: float = 0.0
    units: Optional[Dict] = None


def write_jsonl(out_dir: str, pos, vel, mass, global_id=None, *, time: float = 0.0, redshift: float = 0.0,
                box_size: float = 0.0, provenance: Optional[Provenance] = None, units: Optional[Dict] = None,
                internal_energy=None, smoothing_length=None, ptype=None) -> None:
    """JsonlWriter::write (snapshot.rs:126-146) for n particles: pos, vel
    (n, 3) f64, mass (n,); global_id defaults to 0 .. n-1; the SPH fields to
    0 and the type to DarkMatter.  units: the five SnapshotUnits fields."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    vel = np.asarray(vel, dtype=np.float64).reshape(-1, 3)
    mass = np.asarray(mass, dtype=np.float64).reshape(-1)
    n = pos.shape[0]
    if vel.shape[0] != n or mass.shape[0] != n:
        raise ValueError("pos, vel and mass need the same particle count")
    gid = np.arange(n, dtype=np.int64) if global_id is None else np.asarray(global_id, dtype=np.int64)
    if gid.shape[0] != n or (n and gid.min() < 0):
        raise ValueError("global_id needs one non-negative id per particle")
    u = np.zeros(n) if internal_energy is None else np.asarray(internal_energy, dtype=np.float64)
    h = np.zeros(n) if smoothing_length is None else np.asarray(smoothing_length, dtype=np.float64)
    t = np.zeros(n, dtype=np.int64) if ptype is None else np.asarray(ptype, dtype=np.int64)
    if t.size and (t.min() < 0 or t.max() >= len(PTYPES)):
        raise ValueError("ptype indices must be 0 (DarkMatter), 1 (Gas) or 2 (Star)")
    prov = (provenance or Provenance()).as_dict()
    meta = {"schema_version": 1, "provenance": prov, "particle_count": n, "time": float(time),
            "redshift": float(redshift), "box_size": float(box_size)}
    if units is not None:
        keys = ("length_in_kpc", "mass_in_msun", "velocity_in_km_s", "time_in_gyr", "g_internal")
        meta["units"] = {k: float(units[k]) for k in keys}
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, "meta.json"), "w", encoding="utf-8") as f:
        f.write(_dump(meta, 2, 0))
    with open(os.path.join(out_dir, "provenance.json"), "w", encoding="utf-8") as f:
        f.write(_dump(prov, 2, 0))
    ff = format_f64
    with open(os.path.join(out_dir, "particles.jsonl"), "w", encoding="utf-8") as f:
        for i in range(n):
            p, v = pos[i], vel[i]
            f.write(f'{{"global_id":{int(gid[i])},"mass":{ff(mass[i])},"px":{ff(p[0])},"py":{ff(p[1])},'
                    f'"pz":{ff(p[2])},"vx":{ff(v[0])},"vy":{ff(v[1])},"vz":{ff(v[2])},'
                    f'"internal_energy":{ff(u[i])},"smoothing_length":{ff(h[i])},"ptype":"{PTYPES[t[i]]}"}}\n')


def read_jsonl(snap_dir: str) -> Snapshot:
    """JsonlReader::read (snapshot.rs:151-171): the SPH fields and the type
    default to 0 and DarkMatter when a line lacks them (serde defaults)."""
    with open(os.path.join(snap_dir, "meta.json"), encoding="utf-8") as f:
        meta = json.load(f)
    rows = []
    with open(os.path.join(snap_dir, "particles.jsonl"), encoding="utf-8") as f:
        for line in f:
            if line.strip():
                rows.append(json.loads(line))
    n = len(rows)
    pos = np.array([[r["px"], r["py"], r["pz"]] for r in rows], dtype=np.float64).reshape(n, 3)
    vel = np.array([[r["vx"], r["vy"], r["vz"]] for r in rows], dtype=np.float64).reshape(n, 3)
    return Snapshot(pos=pos, vel=vel, mass=np.array([r["mass"] for r in rows], dtype=np.float64),
                    global_id=np.array([r["global_id"] for r in rows], dtype=np.int64),
                    internal_energy=np.array([r.get("internal_energy", 0.0) for r in rows], dtype=np.float64),
                    smoothing_length=np.array([r.get("smoothing_length", 0.0) for r in rows], dtype=np.float64),
                    ptype=np.array([PTYPES.index(r.get("ptype", "DarkMatter")) for r in rows], dtype=np.int64),
                    time=float(meta["time"]), redshift=float(meta["redshift"]),
                    box_size=float(meta["box_size"]), units=meta.get("units"))
