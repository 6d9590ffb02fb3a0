"""Writes tests/golden/sh3_native.json from the reference's sh3 sample input.

The sample (sample/toml_input/sh3_AICG2+.toml, MIT licence) holds the native
Calpha coordinates of SH3 and the AICG2+ parameters derived from them; its
v0 values are the native bond lengths, 1-3 distances, contact distances and
dihedral angles.  Only numbers are extracted: positions (Angstrom), masses,
and the index/parameter lists of the bonded terms.

    python tests/golden/make_sh3_fixture.py /root/reference/sample/toml_input/sh3_AICG2+.toml
"""
import json
import os
import sys

import tomli


def main(path):
    with open(path, "rb") as f:
        d = tomli.load(f)
    parts = d["systems"][0]["particles"]
    out = {
        "source": "sample/toml_input/sh3_AICG2+.toml",
        "positions_A": [p["pos"] for p in parts],
        "masses": [p["m"] for p in parts],
        "temperature": d["systems"][0]["attributes"]["temperature"],
        "local": [],
        "global": [],
    }
    for ff in d["forcefields"][0]["local"]:
        out["local"].append({k: ff[k] for k in ("interaction", "potential", "topology") if k in ff}
                            | {"parameters": ff["parameters"]})
    for ff in d["forcefields"][0]["global"]:
        g = {k: v for k, v in ff.items() if k != "parameters"}
        g["parameters"] = ff["parameters"]
        out["global"].append(g)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sh3_native.json")
    with open(dst, "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1])
