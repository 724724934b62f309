"""Writes tests/golden/serde_layout.json: the serialised field order of the
reference's structs that make up its checkpoint and products files, read
as text from the reference's Rust sources.

For each struct, the fields in declaration order with their type and serde
attributes; the tests compare the keys this repository writes (in
Configuration, Tasks, App, State, BlockState, Products, BlockProducts) with
these lists.  Run from the repository root with the reference checked out
at /root/reference:  python tests/golden/make_serde_layout.py
"""
import json
import os
import re

REF = "/root/reference"
SOURCES = {
    "RelativisticHydro": "src/physics/relativistic_hydro.rs",
    "NewtonianHydro": "src/physics/newtonian_hydro.rs",
    "Mesh": "src/mesh.rs",
    "Control": "src/app.rs",
    "Configuration": "src/app.rs",
    "App": "src/app.rs",
    "RecurringTask": "src/tasks.rs",
    "Tasks": "src/tasks.rs",
    "State": "src/state.rs",
    "BlockState": "src/state.rs",
    "Products": "src/products.rs",
    "BlockProducts": "src/products.rs",
    "KineticBomb": "src/models/kinetic_bomb.rs",
    "JetInCloud": "src/models/jet_in_cloud.rs",
    "JetInStar": "src/models/jet_in_star.rs",
    "HaloKilonova": "src/models/halo_kilonova.rs",
    "WindShock": "src/models/wind_shock.rs",
}
ENUMS = {"AnyHydro": "src/app.rs", "AnyModel": "src/app.rs", "AnyState": "src/app.rs"}


def struct_fields(text, name):
    m = re.search(r"pub struct " + name + r"(<[^>]*>)?\s*\{(.*?)\n\}", text, re.S)
    if not m:
        raise SystemExit(f"struct {name} not found")
    fields, attrs = [], []
    for line in m.group(2).splitlines():
        line = line.strip()
        if line.startswith("#[serde("):
            attrs.append(line[len("#[serde("):-2])
            continue
        f = re.match(r"pub (\w+)\s*:\s*(.+?),?$", line)
        if f:
            fields.append({"name": f.group(1), "type": f.group(2), "serde": attrs})
            attrs = []
        elif line and not line.startswith("//") and not line.startswith("///") and not line.startswith("#["):
            attrs = []
    return fields


def enum_info(text, name):
    m = re.search(r"((?:#\[[^\n]*\]\s*)*)pub enum " + name + r"\s*\{(.*?)\n\}", text, re.S)
    if not m:
        raise SystemExit(f"enum {name} not found")
    rename = "snake_case" if 'rename_all = "snake_case"' in m.group(1) else None
    variants = re.findall(r"^\s*(\w+)\(", m.group(2), re.M)
    return {"rename_all": rename, "variants": variants}


def main():
    out = {"_source": "Field order and serde attributes read from the reference's sources by "
                      "tests/golden/make_serde_layout.py", "structs": {}, "enums": {}}
    for name, path in SOURCES.items():
        with open(os.path.join(REF, path)) as f:
            out["structs"][name] = {"file": path, "fields": struct_fields(f.read(), name)}
    for name, path in ENUMS.items():
        with open(os.path.join(REF, path)) as f:
            out["enums"][name] = dict(enum_info(f.read(), name), file=path)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "serde_layout.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
