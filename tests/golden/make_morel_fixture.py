"""Writes tests/golden/morel_coefficients.json: the Morel (1988) spectral
table that OceanBioME's MultiBand light uses (src/Light/morel_coefficients.jl:
MOREL_λ, MOREL_kʷ, MOREL_e, MOREL_χ), read as data from the reference
source tree.  Run in the build container, where the reference is present:

    python tests/golden/make_morel_fixture.py /root/reference

The fixture holds the numbers only; tests/test_oracle_pisces_pins.py checks
the package's table against it."""
import json
import os
import re
import sys

NAMES = {"MOREL_λ": "lambda", "MOREL_kʷ": "kw", "MOREL_e": "e", "MOREL_χ": "chi"}


def parse(text):
    out = {}
    for m in re.finditer(r"const\s+(MOREL_\S+)\s*=\s*\[([^\]]*)\]", text):
        name, body = m.group(1), m.group(2)
        if name in NAMES:
            out[NAMES[name]] = [float(x) for x in re.split(r"[,\s]+", body.strip()) if x]
    return out


def main(ref_root):
    src = os.path.join(ref_root, "src", "Light", "morel_coefficients.jl")
    table = parse(open(src, encoding="utf-8").read())
    assert set(table) == set(NAMES.values()), sorted(table)
    n = len(table["lambda"])
    assert all(len(v) == n for v in table.values()), {k: len(v) for k, v in table.items()}
    table["source"] = "src/Light/morel_coefficients.jl (OceanBioME.jl v0.17.6)"
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "morel_coefficients.json"), "w") as f:
        json.dump(table, f, indent=1)
    print(f"wrote {n} wavelengths")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
