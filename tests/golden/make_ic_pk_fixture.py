"""Copies the reference's measured P(k) of two 2LPT initial-condition
snapshots (32^3 particles, seed 42, Eisenstein-Hu with sigma8 = 0.8, a_init =
0.02; normalisation modes "legacy" and "z0_sigma8") into ic_pk/.  Both files
were written by the reference CLI's `analyse --pk-mesh 32` and ship with the
reference under experiments/nbody/phase38_class_validation/output/; the
configuration is experiments/nbody/phase38_class_validation/configs/
lcdm_N32_2lpt_pm_{legacy,rescaled}.toml.  Run from the repository root where
/root/reference exists."""
import hashlib
import json
import os
import shutil

REF = "/root/reference/experiments/nbody/phase38_class_validation/output"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ic_pk")
SOURCES = {"legacy_power_spectrum.jsonl": "legacy/cli_analysis/power_spectrum.jsonl",
           "rescaled_power_spectrum.jsonl": "rescaled/cli_analysis/power_spectrum.jsonl"}


def main():
    os.makedirs(OUT, exist_ok=True)
    sums = {}
    for name, src in SOURCES.items():
        dst = os.path.join(OUT, name)
        shutil.copyfile(os.path.join(REF, src), dst)
        with open(dst, "rb") as f:
            sums[name] = {"source": "experiments/nbody/phase38_class_validation/output/" + src,
                          "sha256": hashlib.sha256(f.read()).hexdigest()}
    with open(os.path.join(OUT, "fixtures.json"), "w") as f:
        json.dump(sums, f, indent=2)
        f.write("\n")


if __name__ == "__main__":
    main()
