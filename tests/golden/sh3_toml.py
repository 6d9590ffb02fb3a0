"""Writes an sh3 TOML input (same schema as the reference's inputs) from the
numbers in tests/golden/sh3_native.json, for the C++ front-end tests."""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def _fmt(v):
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, str):
        return '"' + v + '"'
    if isinstance(v, list):
        return "[" + ", ".join(_fmt(x) for x in v) + "]"
    if isinstance(v, dict):
        return "{" + ", ".join(f'"{k}" = {_fmt(x)}' if not k.isascii() else f"{k} = {_fmt(x)}" for k, x in v.items()) + "}"
    return repr(v)


def write(path, total_step=1000, save_step=100, with_velocity=False):
    d = json.load(open(os.path.join(HERE, "sh3_native.json")))
    lines = ["[files]", 'output.prefix = "sh3"', 'output.path = "./"', 'output.format = "pdb"', "",
             "[simulator]", 'boundary_type = "Unlimited"', "delta_t = 0.01", f"total_step = {total_step}",
             f"save_step = {save_step}", "seed = 12345", "", "[[systems]]",
             f"attributes.temperature = {d['temperature']}", "particles = ["]
    for i, (m, p) in enumerate(zip(d["masses"], d["positions_A"])):
        extra = ', vel = [0.01, -0.02, 0.03]' if with_velocity else ""
        lines.append(f'{{m = {m}, pos = {_fmt(p)}{extra}, name = "CA", group = "PRO"}},')
    lines += ["]", "", "[[forcefields]]"]
    for ff in d["local"]:
        lines.append("[[forcefields.local]]")
        for k in ("interaction", "potential", "topology"):
            if k in ff:
                lines.append(f"{k} = {_fmt(ff[k])}")
        lines.append("parameters = [")
        lines += [_fmt(p) + "," for p in ff["parameters"]]
        lines.append("]")
    for g in d["global"]:
        lines.append("[[forcefields.global]]")
        for k, v in g.items():
            if k == "parameters":
                continue
            if isinstance(v, dict):
                for k2, v2 in v.items():
                    if isinstance(v2, dict):
                        for k3, v3 in v2.items():
                            lines.append(f"{k}.{k2}.{k3} = {_fmt(v3)}")
                    else:
                        lines.append(f"{k}.{k2} = {_fmt(v2)}")
            else:
                lines.append(f"{k} = {_fmt(v)}")
        lines.append("parameters = [")
        lines += [_fmt(p) + "," for p in g["parameters"]]
        lines.append("]")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path
