"""Recipe: regenerate ``residue_templates.json`` from origami's side-chain tables.

The chain builder must place every atom at exactly the coordinates origami's
NeRF builder produces, so the per-residue side-chain internal coordinates
(bond partner, angle/dihedral partners, bond length, bond angle, dihedral
rule) are data that has to match origami's table entry for entry.  This
script reads that table as text from ``crates/chem/src/topology_data.rs``
(and the named bond/angle constants from ``crates/chem/src/topology.rs``),
evaluates the constant expressions with the same IEEE-754 double operations
(``deg * PI / 180``), and writes a JSON data file.  It is run once, by hand,
where the reference tree is mounted; nothing at run time reads the reference.

Usage:  python origami_amd/data/extract_residue_tables.py [/root/reference]
"""
import json
import math
import os
import re
import sys

PI = math.pi


def d(deg):
    # const fn d(deg) = deg * PI / 180.0  (same evaluation order)
    return deg * PI / 180.0


def load_constants(topology_rs):
    consts = {}
    for mod in ("bond", "angle"):
        m = re.search(r"pub mod %s \{(.*?)\n\}" % mod, topology_rs, re.S)
        for name, expr in re.findall(r"pub const ([A-Z_0-9]+): f64 = ([^;]+);", m.group(1)):
            consts["%s::%s" % (mod, name)] = expr.strip()
    return consts


def eval_expr(expr, consts, local):
    expr = expr.strip()
    if expr in local:
        return local[expr]
    if expr in consts:
        return eval_expr(consts[expr], consts, local)
    m = re.fullmatch(r"d\((-?[0-9_.]+)\)", expr)
    if m:
        return d(float(m.group(1).replace("_", "")))
    m = re.fullmatch(r"(-?[0-9_.]+) \* PI / 180\.0", expr)
    if m:
        return float(m.group(1).replace("_", "")) * PI / 180.0
    m = re.fullmatch(r"-?[0-9_.]+", expr)
    if m:
        return float(expr.replace("_", ""))
    raise ValueError("unhandled expression %r" % expr)


def split_top_level(body):
    """Split a comma-separated list at depth 0 of (), {}, []."""
    out, depth, cur = [], 0, []
    for ch in body:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            s = "".join(cur).strip()
            if s:
                out.append(s)
            cur = []
        else:
            cur.append(ch)
    s = "".join(cur).strip()
    if s:
        out.append(s)
    return out


def strip_comments(text):
    return re.sub(r"//[^\n]*", "", text)


def parse_dihedral(expr, consts, local):
    expr = expr.strip()
    m = re.fullmatch(r"DihedralValue::Fixed\((.*)\)", expr)
    if m:
        return {"kind": "fixed", "value": eval_expr(m.group(1), consts, local)}
    m = re.fullmatch(r"DihedralValue::Chi\((\d+)\)", expr)
    if m:
        return {"kind": "chi", "chi": int(m.group(1))}
    m = re.fullmatch(r"DihedralValue::ChiPlus\((\d+),\s*(.*)\)", expr)
    if m:
        return {"kind": "chi_plus", "chi": int(m.group(1)),
                "offset": eval_expr(m.group(2), consts, local)}
    raise ValueError("unhandled dihedral %r" % expr)


def main(ref_root):
    src = os.path.join(ref_root, "crates/chem/src")
    topo = open(os.path.join(src, "topology.rs")).read()
    data = strip_comments(open(os.path.join(src, "topology_data.rs")).read())
    consts = load_constants(topo)
    local = {"TRANS": PI, "GAUCHE_PLUS": PI / 3.0, "GAUCHE_MINUS": -PI / 3.0,
             "CB_DIHEDRAL": d(-122.55), "CB_ANGLE": d(110.5)}

    def sc_atom(fields):
        kv = {}
        for f in split_top_level(fields):
            k, v = f.split(":", 1)
            kv[k.strip()] = v.strip()
        return {
            "name": kv["name"].strip('"'),
            "element": kv["element"].split("::")[1],
            "bond_to": kv["bond_to"].strip('"'),
            "angle_at": kv["angle_at"].strip('"'),
            "dihedral_to": kv["dihedral_to"].strip('"'),
            "bond_length": eval_expr(kv["bond_length_a"], consts, local),
            "bond_angle": eval_expr(kv["bond_angle_rad"], consts, local),
            "dihedral": parse_dihedral(kv["dihedral"], consts, local),
        }

    def helper(name, args):
        a = [x.strip() for x in split_top_level(args)]
        nm, c, b, aa = (s.strip('"') for s in a[:4])
        base = {"name": nm, "element": "H", "bond_to": c, "angle_at": b,
                "dihedral_to": aa, "bond_length": eval_expr("bond::C_H_SP3", consts, local),
                "bond_angle": eval_expr("angle::TETRAHEDRAL", consts, local)}
        if name in ("methyl_h", "beta_h"):
            base["dihedral"] = {"kind": "fixed", "value": eval_expr(a[4], consts, local)}
        elif name == "ch2_chi":
            off = d(120.0) if int(a[5]) >= 0 else d(-120.0)
            base["dihedral"] = {"kind": "chi_plus", "chi": int(a[4]), "offset": off}
        else:
            raise ValueError(name)
        return base

    residues = {}
    for m in re.finditer(r"const ([A-Z]{3}): ResidueTopology = ResidueTopology \{(.*?)\n\};", data, re.S):
        name, body = m.group(1), m.group(2)
        sm = re.search(r"sidechain: &\[(.*?)\n    \],", body, re.S)
        entries = split_top_level(sm.group(1)) if sm else []
        atoms = []
        for e in entries:
            e = e.strip()
            if e == "cb_atom()":
                atoms.append({"name": "CB", "element": "C", "bond_to": "CA", "angle_at": "N",
                              "dihedral_to": "C", "bond_length": eval_expr("bond::C_C", consts, local),
                              "bond_angle": local["CB_ANGLE"],
                              "dihedral": {"kind": "fixed", "value": local["CB_DIHEDRAL"]}})
            elif e.startswith("SidechainAtom {"):
                atoms.append(sc_atom(e[len("SidechainAtom {"):-1]))
            else:
                hm = re.fullmatch(r"(\w+)\((.*)\)", e, re.S)
                atoms.append(helper(hm.group(1), hm.group(2)))
        cm = re.search(r"default_chi_rad: &\[(.*?)\]", body, re.S)
        chis = [eval_expr(x, consts, local) for x in split_top_level(cm.group(1))]
        amide = re.search(r"has_amide_h: (true|false)", body).group(1) == "true"
        gly = re.search(r"is_glycine: (true|false)", body).group(1) == "true"
        residues[name.capitalize()] = {"sidechain": atoms, "default_chi": chis,
                                       "has_amide_h": amide, "is_glycine": gly}
    if len(residues) != 20:
        raise SystemExit("expected 20 residue tables, got %d" % len(residues))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "residue_templates.json")
    with open(out, "w") as f:
        json.dump({"source": "origami crates/chem/src/topology_data.rs (data table)",
                   "residues": residues}, f, indent=1)
    print("wrote", out, sum(len(r["sidechain"]) for r in residues.values()), "side-chain atoms")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
