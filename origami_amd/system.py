"""Host-side system construction: CHARMM36m parameters, atom typing, the
extended-chain NeRF builder, the bonded topology graph and the resolved
per-atom / per-term arrays the force engine consumes.

This is input preparation, not the hot path.  It restates origami's
``crates/chem`` and ``crates/geom`` input side so that a peptide built here
has the same atoms, coordinates, parameters, exclusions and bonded terms as
the one origami would build, and then flattens all of it into plain arrays
(``ResolvedSystem``) -- the content of origami's ``ForceScratch``
(crates/energy/src/scratch.rs:196-251) plus the resolved bonded-term and
SHAKE lists.  Those arrays are what ``og_system_create`` takes over the C ABI.

Reference citations (file:line in origami):
  * amino-acid order ``AminoAcid::ALL``           crates/chem/src/amino_acid.rs:31-52
  * atom types (declaration order = Ord)          crates/chem/src/atom_type.rs:17-106
  * ``classify``                                  crates/chem/src/atom_type.rs:164-348
  * prm / rtf parsing, name translation, lookups  crates/chem/src/forcefield.rs:85-596
  * element masses / vdW radii                    crates/chem/src/element.rs:27-48
  * NeRF ``place_atom``                           crates/geom/src/nerf.rs:15-33
  * chain builder                                 crates/geom/src/builder.rs:19-214
  * topology graph                                crates/geom/src/topology_graph.rs:81-389
  * exclusion mask bits                           crates/energy/src/scratch.rs:31-32,233-251
  * GB intrinsic radii / HCT scales               crates/energy/src/gb.rs:58-83
  * SASA radii / surface tension                  crates/energy/src/powersasa/mod.rs:34-52
  * SHAKE X-H constraint list                     crates/dynamics/src/shake.rs:47-76
  * xoshiro256++ / SplitMix64                     crates/dynamics/src/rng.rs:18-98
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

# --------------------------------------------------------------------------
# Constants (units: Å, kJ/mol, K, Da)
# --------------------------------------------------------------------------
KCAL_TO_KJ = 4.184                      # crates/energy/src/units.rs:5
COULOMB_KCAL = 332.0637                 # crates/energy/src/nonbonded.rs:34
EPS_WATER = 78.5                        # crates/energy/src/gb.rs:23
OBC_OFFSET = 0.09                       # crates/energy/src/gb.rs:32
PROBE_RADIUS = 1.4                      # crates/energy/src/powersasa/mod.rs:22
BOLTZMANN_KJ = 8.314462618e-3           # crates/dynamics/src/langevin.rs:46
ACCEL_FACTOR = 1.0e-4                   # crates/dynamics/src/langevin.rs:49

AMINO_ACIDS = ["Ala", "Arg", "Asn", "Asp", "Cys", "Gln", "Glu", "Gly", "His", "Ile",
               "Leu", "Lys", "Met", "Phe", "Pro", "Ser", "Thr", "Trp", "Tyr", "Val"]
ONE_LETTER = dict(zip(AMINO_ACIDS, "ARNDCQEGHILKMFPSTWYV"))
FROM_ONE_LETTER = {v: k for k, v in ONE_LETTER.items()}

# AtomType in declaration order: the integer index is the Ord used by the
# canonical-key functions of the parameter lookups.
ATOM_TYPES = ["C", "CA", "CAI", "CC", "CT1", "CT2", "CT2A", "CT3", "CP1", "CP2", "CP3",
              "CPH1", "CPH2", "CPT", "CY",
              "N", "NH1", "NH2", "NH3", "NC2", "NR1", "NR2", "NY",
              "O", "OC", "OH1",
              "S",
              "H", "HA", "HA1", "HA2", "HA3", "HB1", "HB2", "HC", "HP", "HR1", "HR3", "HS"]
TYPE_INDEX = {n: i for i, n in enumerate(ATOM_TYPES)}

ELEMENTS = ["H", "C", "N", "O", "P", "S"]
ELEMENT_INDEX = {e: i for i, e in enumerate(ELEMENTS)}
MASS = {"H": 1.008, "C": 12.011, "N": 14.007, "O": 15.999, "P": 30.974, "S": 32.06}
GB_RHO = {"H": 1.20, "C": 1.70, "N": 1.55, "O": 1.50, "P": 1.85, "S": 1.80}
GB_SCALE = {"H": 0.85, "C": 0.72, "N": 0.79, "O": 0.85, "P": 0.86, "S": 0.96}
SASA_VDW = {"H": 1.20, "C": 1.70, "N": 1.55, "O": 1.52, "P": 1.80, "S": 1.80}
SASA_GAMMA_KCAL = {"C": 0.005, "S": 0.005}


def type_element(t: int) -> str:
    name = ATOM_TYPES[t]
    if name == "S":
        return "S"
    return name[0]


# --------------------------------------------------------------------------
# Atom typing (crates/chem/src/atom_type.rs:164-348)
# --------------------------------------------------------------------------
def _sc(pairs):
    out = {}
    for names, t in pairs:
        for n in names.split():
            out[n] = t
    return out


_SIDECHAIN_TYPES = {
    "Ala": _sc([("CB", "CT3"), ("HB1 HB2 HB3", "HA3")]),
    "Val": _sc([("CB", "CT1"), ("HB", "HA1"), ("CG1 CG2", "CT3"),
                ("HG11 HG12 HG13 HG21 HG22 HG23", "HA3")]),
    "Leu": _sc([("CB", "CT2"), ("HB2 HB3", "HA2"), ("CG", "CT1"), ("HG", "HA1"),
                ("CD1 CD2", "CT3"), ("HD11 HD12 HD13 HD21 HD22 HD23", "HA3")]),
    "Ile": _sc([("CB", "CT1"), ("HB", "HA1"), ("CG2", "CT3"), ("HG21 HG22 HG23", "HA3"),
                ("CG1", "CT2"), ("HG12 HG13", "HA2"), ("CD1", "CT3"), ("HD11 HD12 HD13", "HA3")]),
    "Met": _sc([("CB CG", "CT2"), ("HB2 HB3 HG2 HG3", "HA2"), ("SD", "S"), ("CE", "CT3"),
                ("HE1 HE2 HE3", "HA3")]),
    "Pro": _sc([("CB CG", "CP2"), ("CD", "CP3"), ("HB2 HB3 HG2 HG3 HD2 HD3", "HA2")]),
    "Ser": _sc([("CB", "CT2"), ("HB2 HB3", "HA2"), ("OG", "OH1"), ("HG", "H")]),
    "Thr": _sc([("CB", "CT1"), ("HB", "HA1"), ("OG1", "OH1"), ("HG1", "H"), ("CG2", "CT3"),
                ("HG21 HG22 HG23", "HA3")]),
    "Cys": _sc([("CB", "CT2"), ("HB2 HB3", "HA2"), ("SG", "S"), ("HG", "HS")]),
    "Asn": _sc([("CB", "CT2"), ("HB2 HB3", "HA2"), ("CG", "CC"), ("OD1", "O"), ("ND2", "NH2"),
                ("HD21 HD22", "H")]),
    "Gln": _sc([("CB", "CT2"), ("HB2 HB3", "HA2"), ("CG", "CT2"), ("HG2 HG3", "HA2"),
                ("CD", "CC"), ("OE1", "O"), ("NE2", "NH2"), ("HE21 HE22", "H")]),
    "Asp": _sc([("CB", "CT2A"), ("HB2 HB3", "HA2"), ("CG", "CC"), ("OD1 OD2", "OC")]),
    "Glu": _sc([("CB", "CT2A"), ("HB2 HB3", "HA2"), ("CG", "CT2"), ("HG2 HG3", "HA2"),
                ("CD", "CC"), ("OE1 OE2", "OC")]),
    "Lys": _sc([("CB CG CD CE", "CT2"), ("HB2 HB3 HG2 HG3 HD2 HD3 HE2 HE3", "HA2"),
                ("NZ", "NH3"), ("HZ1 HZ2 HZ3", "HC")]),
    "Arg": _sc([("CB CG CD", "CT2"), ("HB2 HB3 HG2 HG3 HD2 HD3", "HA2"), ("NE", "NC2"),
                ("HE", "HC"), ("CZ", "C"), ("NH1 NH2", "NC2"), ("HH11 HH12 HH21 HH22", "HC")]),
    "His": _sc([("CB", "CT2"), ("HB2 HB3", "HA2"), ("CG", "CPH1"), ("ND1", "NR1"), ("HD1", "H"),
                ("CE1", "CPH2"), ("HE1", "HR1"), ("NE2", "NR2"), ("CD2", "CPH1"), ("HD2", "HR3")]),
    "Phe": _sc([("CB", "CT2"), ("HB2 HB3", "HA2"), ("CG CD1 CD2 CE1 CE2 CZ", "CA"),
                ("HD1 HD2 HE1 HE2 HZ", "HP")]),
    "Tyr": _sc([("CB", "CT2"), ("HB2 HB3", "HA2"), ("CG CD1 CD2 CE1 CE2 CZ", "CA"),
                ("HD1 HD2 HE1 HE2", "HP"), ("OH", "OH1"), ("HH", "H")]),
    "Trp": _sc([("CB", "CT2"), ("HB2 HB3", "HA2"), ("CG", "CY"), ("CD1", "CA"), ("HD1", "HP"),
                ("NE1", "NY"), ("HE1", "H"), ("CE2", "CPT"), ("CD2", "CPT"), ("CE3 CZ2", "CAI"),
                ("CZ3 CH2", "CA"), ("HE3 HZ2 HZ3 HH2", "HP")]),
    "Gly": {},
}


def classify(aa: str, name: str) -> Optional[int]:
    """Atom type index for (residue, PDB atom name), or None if not modelled."""
    if aa == "Pro":
        bb = {"N": "N", "CA": "CP1", "C": "C", "O": "O", "HA": "HB1"}
    elif aa == "Gly":
        bb = {"N": "NH1", "CA": "CT2", "C": "C", "O": "O", "H": "H", "HA2": "HB2", "HA3": "HB2"}
    else:
        bb = {"N": "NH1", "CA": "CT1", "C": "C", "O": "O", "H": "H", "HA": "HB1"}
    if name in bb:
        return TYPE_INDEX[bb[name]]
    # Pro/Gly backbone fall-throughs: origami's match falls to the generic
    # backbone arms for names the special arms do not cover.
    generic = {"N": "NH1", "CA": "CT1", "C": "C", "O": "O", "H": "H", "HA": "HB1"}
    if aa in ("Pro", "Gly") and name in generic:
        return TYPE_INDEX[generic[name]]
    t = _SIDECHAIN_TYPES[aa].get(name)
    return None if t is None else TYPE_INDEX[t]


# --------------------------------------------------------------------------
# Residue templates (data table restated from origami's topology_data.rs)
# --------------------------------------------------------------------------
_TEMPLATES = None


def residue_templates():
    global _TEMPLATES
    if _TEMPLATES is None:
        with open(os.path.join(_DATA, "residue_templates.json")) as f:
            _TEMPLATES = json.load(f)["residues"]
    return _TEMPLATES


# --------------------------------------------------------------------------
# CHARMM36m force field (crates/chem/src/forcefield.rs)
# --------------------------------------------------------------------------
def _canon_pair(a, b):
    return (a, b) if a <= b else (b, a)


def _canon_triple(a, b, c):
    return (a, b, c) if a <= c else (c, b, a)


def _canon_quad(a, b, c, d):
    return (a, b, c, d) if (b, a) <= (c, d) else (d, c, b, a)


_WILD = -1
_UNKNOWN = -2


def _parse_atom(tok):
    if tok in ("X", "x"):
        return _WILD
    return TYPE_INDEX.get(tok, _UNKNOWN)


def _float(tok):
    try:
        return float(tok)
    except ValueError:
        return None


def _charmm_to_pdb(aa, charmm):
    if charmm == "HN":
        return "H"
    table = {
        ("Gly", "HA1"): "HA2", ("Gly", "HA2"): "HA3",
        ("Ile", "CD"): "CD1", ("Ile", "HD1"): "HD11", ("Ile", "HD2"): "HD12",
        ("Ile", "HD3"): "HD13", ("Ile", "HG11"): "HG12", ("Ile", "HG12"): "HG13",
        ("Ser", "HG1"): "HG", ("Cys", "HG1"): "HG",
    }
    if (aa, charmm) in table:
        return table[(aa, charmm)]
    if aa in ("Leu", "Met", "Pro", "Ser", "Cys", "Asn", "Gln", "Asp", "Glu", "Lys", "Arg",
              "His", "Phe", "Tyr", "Trp"):
        if charmm == "HB1":
            return "HB2"
        if charmm == "HB2":
            return "HB3"
    if aa in ("Met", "Pro", "Gln", "Glu", "Lys", "Arg"):
        if charmm == "HG1":
            return "HG2"
        if charmm == "HG2":
            return "HG3"
    if aa in ("Pro", "Lys", "Arg"):
        if charmm == "HD1":
            return "HD2"
        if charmm == "HD2":
            return "HD3"
    if aa == "Lys":
        if charmm == "HE1":
            return "HE2"
        if charmm == "HE2":
            return "HE3"
    if charmm in ("N", "CA", "C", "O", "HA"):
        return charmm
    for sc in residue_templates()[aa]["sidechain"]:
        if sc["name"] == charmm:
            return charmm
    return ""


_RTF_RESI = {"ALA": "Ala", "ARG": "Arg", "ASN": "Asn", "ASP": "Asp", "CYS": "Cys",
             "GLN": "Gln", "GLU": "Glu", "GLY": "Gly", "HSD": "His", "ILE": "Ile",
             "LEU": "Leu", "LYS": "Lys", "MET": "Met", "PHE": "Phe", "PRO": "Pro",
             "SER": "Ser", "THR": "Thr", "TRP": "Trp", "TYR": "Tyr", "VAL": "Val"}


class ForceField:
    """Parsed CHARMM36m tables with origami's canonical-key lookups."""

    def __init__(self, prm_text: str, rtf_text: str):
        self.bonds: Dict[tuple, tuple] = {}
        self.angles: Dict[tuple, tuple] = {}
        self.dihedrals: Dict[tuple, list] = {}
        self.wildcard_dihedrals: Dict[tuple, list] = {}
        self.impropers: Dict[tuple, tuple] = {}
        self.wildcard_impropers: Dict[tuple, tuple] = {}
        self.nonbonded: Dict[int, tuple] = {}
        self.charges: Dict[tuple, float] = {}
        self._parse_prm(prm_text)
        self._parse_rtf(rtf_text)
        # Disulfide S-S bond patch (forcefield.rs:206-212).
        self.bonds.setdefault(_canon_pair(TYPE_INDEX["S"], TYPE_INDEX["S"]), (173.0, 2.029))

    # -- prm ---------------------------------------------------------------
    def _parse_prm(self, text):
        headers = {"BONDS": "b", "ANGLES": "a", "DIHEDRALS": "d", "IMPROPER": "i",
                   "IMPROPERS": "i", "CMAP": "skip", "NONBONDED": "n", "HBOND": "skip",
                   "END": "end"}
        section = None
        for raw in text.splitlines():
            line = raw.split("!", 1)[0].strip()
            if not line:
                continue
            first = line.split()[0]
            if first in headers:
                section = headers[first]
                if section == "end":
                    break
                continue
            if section is None or section == "skip":
                continue
            tok = line.split()
            if section == "b" and len(tok) >= 4:
                a, b = _parse_atom(tok[0]), _parse_atom(tok[1])
                k, r0 = _float(tok[2]), _float(tok[3])
                if a >= 0 and b >= 0 and k is not None and r0 is not None:
                    self.bonds.setdefault(_canon_pair(a, b), (k, r0))
            elif section == "a" and len(tok) >= 5:
                a, b, c = (_parse_atom(t) for t in tok[:3])
                k, th = _float(tok[3]), _float(tok[4])
                if a >= 0 and b >= 0 and c >= 0 and k is not None and th is not None:
                    self.angles.setdefault(_canon_triple(a, b, c), (k, th))
            elif section == "d" and len(tok) >= 7:
                a, b, c, d = (_parse_atom(t) for t in tok[:4])
                k = _float(tok[4])
                try:
                    n = int(tok[5])
                except ValueError:
                    continue
                delta = _float(tok[6])
                if k is None or delta is None or n < 0:
                    continue
                term = (k, n, delta)
                if a >= 0 and b >= 0 and c >= 0 and d >= 0:
                    self.dihedrals.setdefault(_canon_quad(a, b, c, d), []).append(term)
                elif a == _WILD and b >= 0 and c >= 0 and d == _WILD:
                    self.wildcard_dihedrals.setdefault(_canon_pair(b, c), []).append(term)
            elif section == "i" and len(tok) >= 7:
                a, b, c, d = (_parse_atom(t) for t in tok[:4])
                k, psi0 = _float(tok[4]), _float(tok[6])
                if k is None or psi0 is None:
                    continue
                if a >= 0 and b == _WILD and c == _WILD and d >= 0:
                    self.wildcard_impropers.setdefault((a, d), (k, psi0))
                elif a >= 0 and b >= 0 and c >= 0 and d >= 0:
                    off = sorted([b, c, d])
                    self.impropers.setdefault((off[0], a, off[1], off[2]), (k, psi0))
            elif section == "n":
                if "cutnb" in line or "ctofnb" in line or len(tok) < 4:
                    continue
                t = _parse_atom(tok[0])
                if t < 0:
                    continue
                eps, rmin = _float(tok[2]), _float(tok[3])
                if eps is None or rmin is None:
                    continue
                e14 = r14 = None
                if len(tok) >= 7:
                    a5, a6 = _float(tok[5]), _float(tok[6])
                    if a5 is not None and a6 is not None:
                        e14, r14 = -a5, a6
                self.nonbonded.setdefault(t, (-eps, rmin, e14, r14))

    # -- rtf ---------------------------------------------------------------
    def _parse_rtf(self, text):
        current = None
        for raw in text.splitlines():
            line = raw.split("!", 1)[0].strip()
            if not line:
                continue
            tok = line.split()
            head = tok[0]
            if head == "RESI":
                current = _RTF_RESI.get(tok[1] if len(tok) > 1 else "")
                continue
            if head == "PRES":
                current = None
                continue
            if head != "ATOM" or current is None:
                continue
            name = tok[1] if len(tok) > 1 else ""
            q = _float(tok[3]) if len(tok) > 3 else None
            if q is None:
                continue
            self.charges[(current, _charmm_to_pdb(current, name))] = q

    # -- lookups -----------------------------------------------------------
    def bond(self, a, b):
        return self.bonds.get(_canon_pair(a, b))

    def angle(self, a, b, c):
        return self.angles.get(_canon_triple(a, b, c))

    def dihedral(self, a, b, c, d):
        t = self.dihedrals.get(_canon_quad(a, b, c, d))
        if t is not None:
            return t
        return self.wildcard_dihedrals.get(_canon_pair(b, c))

    def improper(self, a, b, c, d):
        # forcefield.rs:111-147: try each of b, a, c, d as the central type;
        # the off atoms are the others *by type* (padded with central).
        for central in (b, a, c, d):
            others = [t for t in (a, b, c, d) if t != central]
            while len(others) < 3:
                others.append(central)
            off = sorted(others[:3])
            p = self.impropers.get((off[0], central, off[1], off[2]))
            if p is not None:
                return p
            for spec in off:
                p = self.wildcard_impropers.get((central, spec))
                if p is not None:
                    return p
        return None

    def nonbonded_params(self, t):
        return self.nonbonded.get(t)

    def partial_charge(self, aa, name):
        return self.charges.get((aa, name))


_FF = None


def standard_ff() -> ForceField:
    global _FF
    if _FF is None:
        d = os.path.join(_DATA, "charmm36")
        with open(os.path.join(d, "par_all36m_prot.prm")) as f:
            prm = f.read()
        with open(os.path.join(d, "top_all36_prot.rtf")) as f:
            rtf = f.read()
        _FF = ForceField(prm, rtf)
    return _FF


# --------------------------------------------------------------------------
# Geometry helpers and NeRF builder
# --------------------------------------------------------------------------
def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _norm(a):
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def _normalize(a):
    n = _norm(a)
    return (a[0] / n, a[1] / n, a[2] / n)


def place_atom(a, b, c, length, angle, dihedral):
    """NeRF placement of D from A, B, C (crates/geom/src/nerf.rs:15-33)."""
    bc = _normalize(_sub(c, b))
    ba = _sub(a, b)
    n = _normalize(_cross(bc, ba))
    y = _cross(n, bc)
    st, ct = math.sin(angle), math.cos(angle)
    sd, cd = math.sin(dihedral), math.cos(dihedral)
    lx = length * ct
    ly = length * st * cd
    lz = length * st * sd
    return (c[0] + (-bc[0] * lx + y[0] * ly + n[0] * lz),
            c[1] + (-bc[1] * lx + y[1] * ly + n[1] * lz),
            c[2] + (-bc[2] * lx + y[2] * ly + n[2] * lz))


PI = math.pi
DEFAULT_PHI = -120.0 * PI / 180.0
DEFAULT_PSI = 140.0 * PI / 180.0
DEFAULT_OMEGA = PI
CB_DIHEDRAL = -122.55 * PI / 180.0
PEPTIDE_C_N, N_CA, CA_C, C_O, N_H_AMIDE, CA_HA = 1.329, 1.458, 1.525, 1.231, 1.010, 1.090
C_N_CA_ANGLE = 121.7 * PI / 180.0
CA_C_O_ANGLE = 120.8 * PI / 180.0
HN_BOND_ANGLE = 119.0 * PI / 180.0
HA_BOND_ANGLE = 109.0 * PI / 180.0
N_CA_C_ANGLE = 111.2 * PI / 180.0
CA_C_N_ANGLE = 116.2 * PI / 180.0


@dataclass
class Residue:
    aa: str
    names: List[str] = field(default_factory=list)
    elements: List[str] = field(default_factory=list)
    pos: List[tuple] = field(default_factory=list)
    chain: str = "A"

    def position(self, name):
        for n, p in zip(self.names, self.pos):
            if n == name:
                return p
        return None

    def push(self, name, el, p):
        self.names.append(name)
        self.elements.append(el)
        self.pos.append(p)


def append_residue(residues: List[Residue], aa: str, phi=DEFAULT_PHI, psi=DEFAULT_PSI,
                   omega=DEFAULT_OMEGA) -> Residue:
    """Place residue aa after the last one of residues, at the given backbone
    dihedrals, from that residue's current coordinates (builder.rs:78-214,
    append_residue: the co-translational driver extends a chain that has
    already moved), and append it."""
    tmpl = residue_templates()
    idx = len(residues)
    topo = tmpl[aa]
    r = Residue(aa)
    if idx == 0:
        n = (0.0, 0.0, 0.0)
        ca = (N_CA, 0.0, 0.0)
        dx = -CA_C * math.cos(N_CA_C_ANGLE)
        dy = CA_C * math.sin(N_CA_C_ANGLE)
        c = (ca[0] + dx, ca[1] + dy, 0.0)
    else:
        prev = residues[idx - 1]
        pn, pca, pc = prev.position("N"), prev.position("CA"), prev.position("C")
        n = place_atom(pn, pca, pc, PEPTIDE_C_N, CA_C_N_ANGLE, psi)
        ca = place_atom(pca, pc, n, N_CA, C_N_CA_ANGLE, omega)
        c = place_atom(pc, n, ca, CA_C, N_CA_C_ANGLE, phi)
    r.push("N", "N", n)
    r.push("CA", "C", ca)
    r.push("C", "C", c)
    o = place_atom(n, ca, c, C_O, CA_C_O_ANGLE, psi + PI)
    r.push("O", "O", o)
    if topo["has_amide_h"] and idx > 0:
        prev = residues[idx - 1]
        h = place_atom(prev.position("O"), prev.position("C"), n, N_H_AMIDE,
                       HN_BOND_ANGLE, PI)
        r.push("H", "H", h)
    elif topo["has_amide_h"] and idx == 0:
        h = place_atom(c, ca, n, N_H_AMIDE, HN_BOND_ANGLE, PI)
        r.push("H", "H", h)
    if topo["is_glycine"]:
        r.push("HA2", "H", place_atom(c, n, ca, CA_HA, HA_BOND_ANGLE, -CB_DIHEDRAL))
        r.push("HA3", "H", place_atom(c, n, ca, CA_HA, HA_BOND_ANGLE, CB_DIHEDRAL))
    else:
        r.push("HA", "H", place_atom(c, n, ca, CA_HA, HA_BOND_ANGLE, -CB_DIHEDRAL))
    chi = topo["default_chi"]
    for sc in topo["sidechain"]:
        def look(nm):
            p = r.position(nm)
            if p is None and idx > 0:
                p = residues[idx - 1].position(nm)
            if p is None:
                raise ValueError("residue %d references unplaced atom %s" % (idx, nm))
            return p
        pa, pb, pc_ = look(sc["dihedral_to"]), look(sc["angle_at"]), look(sc["bond_to"])
        dv = sc["dihedral"]
        if dv["kind"] == "fixed":
            dih = dv["value"]
        elif dv["kind"] == "chi":
            dih = chi[dv["chi"] - 1]
        else:
            dih = chi[dv["chi"] - 1] + dv["offset"]
        r.push(sc["name"], sc["element"],
               place_atom(pa, pb, pc_, sc["bond_length"], sc["bond_angle"], dih))
    residues.append(r)
    return r


def build_chain(sequence: Sequence[str], phi=DEFAULT_PHI, psi=DEFAULT_PSI,
                omega=DEFAULT_OMEGA) -> List[Residue]:
    """Extended all-atom chain (crates/geom/src/builder.rs:50-214)."""
    if len(sequence) == 0:
        raise ValueError("empty sequence")
    residues: List[Residue] = []
    for aa in sequence:
        append_residue(residues, aa, phi, psi, omega)
    return residues


# --------------------------------------------------------------------------
# Topology graph (crates/geom/src/topology_graph.rs:134-389)
# --------------------------------------------------------------------------
_RING_CLOSURES = {"Pro": [("N", "CD")], "Phe": [("CE2", "CZ")], "Tyr": [("CE2", "CZ")],
                  "His": [("CE1", "NE2")], "Trp": [("CE2", "CD2"), ("CH2", "CZ3")]}

_IMPROPER_CENTRES = {"Asn": ("CB", "CG", "OD1", "ND2"), "Gln": ("CG", "CD", "OE1", "NE2"),
                     "Asp": ("CB", "CG", "OD1", "OD2"), "Glu": ("CG", "CD", "OE1", "OE2"),
                     "Arg": ("NE", "CZ", "NH1", "NH2")}


@dataclass
class Topology:
    bonds: List[Tuple[int, int]]
    angles: List[Tuple[int, int, int]]
    dihedrals: List[Tuple[int, int, int, int]]
    impropers: List[Tuple[int, int, int, int]]
    bonded_to: List[List[int]]


def build_topology(residues: List[Residue]) -> Topology:
    index: Dict[Tuple[int, str], int] = {}
    total = 0
    for ri, r in enumerate(residues):
        for nm in r.names:
            index[(ri, nm)] = total
            total += 1

    def lk(ri, nm):
        return index.get((ri, nm))

    bonds = set()

    def add(a, b):
        if a is not None and b is not None and a != b:
            bonds.add((a, b) if a < b else (b, a))

    tmpl = residue_templates()
    for ri, r in enumerate(residues):
        n, ca, c, o = lk(ri, "N"), lk(ri, "CA"), lk(ri, "C"), lk(ri, "O")
        if n is not None and ca is not None:
            add(n, ca)
        if ca is not None and c is not None:
            add(ca, c)
        if c is not None and o is not None:
            add(c, o)
        add(n, lk(ri, "H")) if n is not None else None
        if r.aa == "Gly":
            for ha in ("HA2", "HA3"):
                if ca is not None:
                    add(ca, lk(ri, ha))
        elif ca is not None:
            add(ca, lk(ri, "HA"))
        if ri > 0 and residues[ri - 1].chain == r.chain:
            pc = lk(ri - 1, "C")
            if pc is not None and n is not None:
                add(pc, n)
        for sc in tmpl[r.aa]["sidechain"]:
            add(lk(ri, sc["bond_to"]), lk(ri, sc["name"]))
        for na, nb in _RING_CLOSURES.get(r.aa, []):
            add(lk(ri, na), lk(ri, nb))
    # Disulfides: SG-SG within 2.5 Å.
    sgs = []
    for ri, r in enumerate(residues):
        if r.aa == "Cys":
            i, p = lk(ri, "SG"), r.position("SG")
            if i is not None and p is not None:
                sgs.append((i, p))
    for a in range(len(sgs)):
        for b in range(a + 1, len(sgs)):
            dd = _sub(sgs[a][1], sgs[b][1])
            if dd[0] ** 2 + dd[1] ** 2 + dd[2] ** 2 <= 2.5 * 2.5:
                add(sgs[a][0], sgs[b][0])
    bonds = sorted(bonds)
    bonded_to: List[List[int]] = [[] for _ in range(total)]
    for a, b in bonds:
        bonded_to[a].append(b)
        bonded_to[b].append(a)
    for nb in bonded_to:
        nb.sort()
    angles = []
    for b, nb in enumerate(bonded_to):
        for i in range(len(nb)):
            for j in range(i + 1, len(nb)):
                angles.append((nb[i], b, nb[j]))
    seen = set()
    dihedrals = []
    for b, c in bonds:
        for a in bonded_to[b]:
            if a == c:
                continue
            for d in bonded_to[c]:
                if d == b or d == a:
                    continue
                key = (a, b, c, d) if b < c else (d, c, b, a)
                if key not in seen:
                    seen.add(key)
                    dihedrals.append(key)
    impropers = []
    nres = len(residues)
    for ri, r in enumerate(residues):
        if ri + 1 < nres:
            q = (lk(ri, "CA"), lk(ri, "C"), lk(ri, "O"), lk(ri + 1, "N"))
            if None not in q:
                impropers.append((q[1], q[0], q[2], q[3]))
        if r.aa in _IMPROPER_CENTRES:
            s_, cen, o1, o2 = (lk(ri, x) for x in _IMPROPER_CENTRES[r.aa])
            if None not in (s_, cen, o1, o2):
                impropers.append((cen, s_, o1, o2))
    return Topology(bonds, angles, dihedrals, impropers, bonded_to)


def exclusion_pairs(topo: Topology):
    """(i, j, kind) for i < j: kind 1 = 1-2/1-3 excluded, 2 = 1-4
    (scratch.rs:233-251 with topology_graph.rs:81-115 semantics)."""
    bt = [set(x) for x in topo.bonded_to]
    out = {}
    for a in range(len(bt)):
        one_two = bt[a]
        one_three = set()
        for n1 in one_two:
            one_three |= bt[n1]
        one_three.discard(a)
        one_four = set()
        for n1 in one_two:
            for n2 in bt[n1]:
                one_four |= bt[n2]
        for b in one_two | one_three:
            if b > a:
                out[(a, b)] = 1
        for b in one_four:
            if b > a and (a, b) not in out and b not in one_two and b not in one_three:
                out[(a, b)] = 2
    return sorted((a, b, k) for (a, b), k in out.items())


# --------------------------------------------------------------------------
# RNG (crates/dynamics/src/rng.rs:18-98)
# --------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & _M64


class Xoshiro256pp:
    def __init__(self, seed: int):
        s = seed & _M64
        st = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & _M64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
            st.append(z ^ (z >> 31))
        self.s = st
        self.cached = None

    def next_u64(self):
        s = self.s
        result = (_rotl((s[0] + s[3]) & _M64, 23) + s[0]) & _M64
        t = (s[1] << 17) & _M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def next_f64(self):
        return (self.next_u64() >> 11) * (1.0 / float(1 << 53))

    def gaussian(self):
        if self.cached is not None:
            g, self.cached = self.cached, None
            return g
        u1 = self.next_f64()
        while u1 <= 0.0:
            u1 = self.next_f64()
        u2 = self.next_f64()
        r = math.sqrt(-2.0 * math.log(u1))
        th = 2.0 * math.pi * u2
        self.cached = r * math.sin(th)
        return r * math.cos(th)


def random_sequence(rng: Xoshiro256pp, length: int) -> List[str]:
    """Uniform draw over ``AminoAcid::ALL`` order (amino_acid.rs:31-52)."""
    return [AMINO_ACIDS[int(rng.next_f64() * 20.0)] for _ in range(length)]


# --------------------------------------------------------------------------
# Resolved system (the C-ABI input)
# --------------------------------------------------------------------------
@dataclass
class ResolvedSystem:
    n_atoms: int
    sequence: List[str]
    names: List[str]
    residue_index: np.ndarray      # int32 [N]
    element: np.ndarray            # int32 [N] index into ELEMENTS
    atom_type: np.ndarray          # int32 [N]
    pos: np.ndarray                # f64 [N,3]
    charge: np.ndarray             # f64 [N] e
    rmin_half: np.ndarray          # f64 [N] Å
    epsilon: np.ndarray            # f64 [N] kcal/mol
    rmin_half_14: np.ndarray
    epsilon_14: np.ndarray
    gb_rho: np.ndarray             # f64 [N] intrinsic radius ρ
    gb_scale: np.ndarray           # f64 [N] HCT scale
    sasa_radius: np.ndarray        # f64 [N] vdW + probe
    sasa_gamma: np.ndarray         # f64 [N] kJ/mol/Å² (unscaled)
    mass: np.ndarray               # f64 [N] Da
    excl_pairs: np.ndarray         # int32 [P,2] i<j, 1-2 or 1-3
    pair14: np.ndarray             # int32 [Q,2] i<j, 1-4
    bonds: np.ndarray              # int32 [B,2]
    bond_params: np.ndarray        # f64 [B,2] (K kcal/mol/Å², r0 Å)
    angles: np.ndarray             # int32 [A,3]
    angle_params: np.ndarray       # f64 [A,2] (K kcal/mol/rad², θ0 rad)
    dihedrals: np.ndarray          # int32 [D,4]  one row per (dihedral, term)
    dihedral_params: np.ndarray    # f64 [D,3] (k kcal/mol, n, δ rad)
    dihedral_first: np.ndarray     # int32 [D] 1 if the row is a dihedral's first term
    impropers: np.ndarray          # int32 [I,4]
    improper_params: np.ndarray    # f64 [I,2] (K kcal/mol/rad², ψ0 rad)
    constraints: np.ndarray        # int32 [C,2] X-H
    constraint_len: np.ndarray     # f64 [C] r0
    missing_params: int = 0

    def positions_copy(self):
        return self.pos.copy()


def deg_to_rad(x):
    return x * math.pi / 180.0


def resolve(residues: List[Residue], ff: Optional[ForceField] = None) -> ResolvedSystem:
    """Flatten a built structure into the arrays the engine consumes."""
    ff = ff or standard_ff()
    topo = build_topology(residues)
    names, elems, types, ridx, pos = [], [], [], [], []
    charge, rmin, eps, rmin14, eps14 = [], [], [], [], []
    for ri, r in enumerate(residues):
        for nm, el, p in zip(r.names, r.elements, r.pos):
            t = classify(r.aa, nm)
            if t is None:
                raise ValueError("unclassified atom %s %s" % (r.aa, nm))
            names.append(nm)
            elems.append(ELEMENT_INDEX[el])
            types.append(t)
            ridx.append(ri)
            pos.append(p)
            q = ff.partial_charge(r.aa, nm)
            charge.append(0.0 if q is None else q)
            nb = ff.nonbonded_params(t)
            if nb is None:
                rmin.append(0.0), eps.append(0.0), rmin14.append(0.0), eps14.append(0.0)
            else:
                e, rm, e14, r14 = nb
                eps.append(e)
                rmin.append(rm)
                eps14.append(e if e14 is None else e14)
                rmin14.append(rm if r14 is None else r14)
    n = len(names)
    el_names = [ELEMENTS[e] for e in elems]
    missing = 0
    bonds, bparams = [], []
    for a, b in topo.bonds:
        p = ff.bond(types[a], types[b])
        if p is None:
            missing += 1
            continue
        bonds.append((a, b))
        bparams.append(p)
    angles, aparams = [], []
    for a, b, c in topo.angles:
        p = ff.angle(types[a], types[b], types[c])
        if p is None:
            missing += 1
            continue
        angles.append((a, b, c))
        aparams.append((p[0], deg_to_rad(p[1])))
    dih, dparams, dfirst = [], [], []
    for a, b, c, d in topo.dihedrals:
        terms = ff.dihedral(types[a], types[b], types[c], types[d])
        if terms is None:
            missing += 1
            continue
        for ti, (k, nn, delta) in enumerate(terms):
            dih.append((a, b, c, d))
            dparams.append((k, float(nn), deg_to_rad(delta)))
            dfirst.append(1 if ti == 0 else 0)
    imps, iparams = [], []
    for a, b, c, d in topo.impropers:
        p = ff.improper(types[a], types[b], types[c], types[d])
        if p is None:
            missing += 1
            continue
        imps.append((a, b, c, d))
        iparams.append((p[0], deg_to_rad(p[1])))
    cons, clen = [], []
    for a, b in topo.bonds:
        if el_names[a] == "H" or el_names[b] == "H":
            p = ff.bond(types[a], types[b])
            if p is None:
                continue
            cons.append((a, b))
            clen.append(p[1])
    ex = exclusion_pairs(topo)
    excl = [(a, b) for a, b, k in ex if k == 1]
    p14 = [(a, b) for a, b, k in ex if k == 2]

    def i32(x, w):
        return np.asarray(x, dtype=np.int32).reshape(-1, w)

    def f64(x, w=None):
        a = np.asarray(x, dtype=np.float64)
        return a.reshape(-1, w) if w else a

    return ResolvedSystem(
        n_atoms=n, sequence=[r.aa for r in residues], names=names,
        residue_index=np.asarray(ridx, np.int32), element=np.asarray(elems, np.int32),
        atom_type=np.asarray(types, np.int32), pos=f64(pos, 3), charge=f64(charge),
        rmin_half=f64(rmin), epsilon=f64(eps), rmin_half_14=f64(rmin14), epsilon_14=f64(eps14),
        gb_rho=f64([GB_RHO[e] for e in el_names]), gb_scale=f64([GB_SCALE[e] for e in el_names]),
        sasa_radius=f64([SASA_VDW[e] + PROBE_RADIUS for e in el_names]),
        sasa_gamma=f64([SASA_GAMMA_KCAL.get(e, 0.0) * KCAL_TO_KJ for e in el_names]),
        mass=f64([MASS[e] for e in el_names]),
        excl_pairs=i32(excl, 2), pair14=i32(p14, 2),
        bonds=i32(bonds, 2), bond_params=f64(bparams, 2),
        angles=i32(angles, 3), angle_params=f64(aparams, 2),
        dihedrals=i32(dih, 4), dihedral_params=f64(dparams, 3),
        dihedral_first=np.asarray(dfirst, np.int32),
        impropers=i32(imps, 4), improper_params=f64(iparams, 2),
        constraints=i32(cons, 2), constraint_len=f64(clen), missing_params=missing)


def build_system(sequence: Sequence[str]) -> ResolvedSystem:
    """Extended chain for a sequence of three-letter or one-letter codes."""
    seq = [FROM_ONE_LETTER.get(s, s) for s in sequence]
    return resolve(build_chain(seq))


def seeded_chain(seed: int, n_residues: int) -> ResolvedSystem:
    rng = Xoshiro256pp(seed)
    return resolve(build_chain(random_sequence(rng, n_residues)))
