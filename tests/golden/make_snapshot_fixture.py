"""Writes tests/golden/snapshot/: the first 200 lines of a particles.jsonl and
two meta.json files written by the reference's own JSONL snapshot writer
(serde_json), taken from the data the reference tree ships under
experiments/nbody/.  Data only: the files are copied as bytes, nothing of the
reference is run.  Run from the repository root where /root/reference exists."""
import hashlib
import json
import os
import shutil

REF = "/root/reference/experiments/nbody"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "snapshot")
SOURCES = {
    "p16_particles_head.jsonl": ("phase16_tiled_simd/results/p16/validation_N2000_P2/frames/snap_000005/"
                                 "particles.jsonl", 200),
    "p16_meta.json": ("phase16_tiled_simd/results/p16/validation_N2000_P2/frames/snap_000005/meta.json", None),
    "p16_provenance.json": ("phase16_tiled_simd/results/p16/validation_N2000_P2/frames/snap_000005/"
                            "provenance.json", None),
    "p38_meta.json": ("phase38_class_validation/output/rescaled/cli_snapshot/meta.json", None),
}


def main():
    os.makedirs(OUT, exist_ok=True)
    sums = {}
    for name, (src, lines) in SOURCES.items():
        path = os.path.join(REF, src)
        dst = os.path.join(OUT, name)
        if lines is None:
            shutil.copyfile(path, dst)
        else:
            with open(path, "rb") as f:
                head = b"".join(f.readline() for _ in range(lines))
            with open(dst, "wb") as f:
                f.write(head)
        with open(dst, "rb") as f:
            sums[name] = {"source": "experiments/nbody/" + src, "lines": lines,
                          "sha256": hashlib.sha256(f.read()).hexdigest()}
    with open(os.path.join(OUT, "fixtures.json"), "w") as f:
        json.dump(sums, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
