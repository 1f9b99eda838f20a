"""PDB ATOM-record input and output, and Cα RMSD.

Host-side formats on either side of the engine:

  read_pdb / read_pdb_trajectory   io/src/pdb_read.rs:38-297 (ATOM records of
                                   the 20 standard residues, alternate
                                   location ' ' or 'A' only, PDB v2 names such
                                   as "1HD1" read as "HD11", terminal-patch
                                   atoms H1/H2/H3/OXT/HXT skipped, a repeated
                                   atom name within a residue kept once)
  write_pdb / write_pdb_trajectory io/src/pdb.rs:19-178 (column-exact v3.3
                                   records, one MODEL block per frame, TER at
                                   chain changes and at the end)
  rmsd_ca / rmsd_points            geom/src/rmsd.rs:30-92 (Kabsch)

Structures are lists of ``system.Residue``; ``system_from_residues`` resolves
them into the engine's input (topology and CHARMM36 parameters) the way the
seeded chains are resolved.
"""
from __future__ import annotations

import io
from typing import Iterable, List, Optional, Sequence, TextIO, Union

import numpy as np

from . import system as S

TERMINAL_PATCH = {"H1", "H2", "H3", "OXT", "HXT"}
BACKBONE = ("N", "CA", "C", "O", "H", "HA", "HA2", "HA3")
ELEMENTS = {"H", "C", "N", "O", "S"}


class PdbReadError(ValueError):
    pass


def _three_letter():
    return {aa.upper(): aa for aa in S.AMINO_ACIDS}


def _sidechain_names():
    tmpl = S.residue_templates()
    return {aa: {sc["name"] for sc in tmpl[aa]["sidechain"]} for aa in S.AMINO_ACIDS}


def normalise_atom_name(name: str) -> str:
    """PDB v2 "1HD1" -> v3.3 "HD11" (pdb_read.rs:353-367)."""
    if len(name) >= 4 and name[0].isdigit():
        return name[1:] + name[0]
    return name


def _canonical(aa: str, name: str, sc_names) -> Optional[str]:
    if name in BACKBONE:
        return name
    if name in sc_names[aa]:
        return name
    return None


def _element(rec_el: str, name: str) -> Optional[str]:
    e = rec_el.strip().upper()
    if e in ELEMENTS:
        return e
    c = name[:1]
    return c if c in ELEMENTS else None


def _parse_atom(line: str, lineno: int):
    s = line.ljust(80)
    try:
        res_seq = int(s[22:26])
    except ValueError:
        raise PdbReadError("line %d: malformed ATOM record: residue seq number" % lineno)
    xyz = []
    for a, b, what in ((30, 38, "x"), (38, 46, "y"), (46, 54, "z")):
        try:
            xyz.append(float(s[a:b]))
        except ValueError:
            raise PdbReadError("line %d: malformed ATOM record: %s coordinate" % (lineno, what))
    return dict(name=s[12:16].strip(), alt=s[16], res=s[17:20].strip(), chain=s[21],
                seq=res_seq, xyz=tuple(xyz), element=s[76:78])


def _lines(src: Union[str, TextIO]):
    if hasattr(src, "read"):
        return src.read().splitlines()
    if "\n" in src:
        return src.splitlines()
    with open(src) as f:
        return f.read().splitlines()


def _assemble(atoms) -> List[S.Residue]:
    residues: List[S.Residue] = []
    key = None
    for (chain, seq, aa), name, el, xyz in atoms:
        if (chain, seq) != key:
            residues.append(S.Residue(aa=aa, chain=chain))
            key = (chain, seq)
        r = residues[-1]
        if name not in r.names:
            r.push(name, el, xyz)
    return residues


def read_pdb_trajectory(src: Union[str, TextIO]) -> List[List[S.Residue]]:
    """Every MODEL of a PDB (a file without MODEL records is one frame)."""
    three, sc = _three_letter(), _sidechain_names()
    frames, cur = [], []
    for lineno, line in enumerate(_lines(src), 1):
        if line.startswith("ENDMDL"):
            if cur:
                frames.append(_assemble(cur))
                cur = []
            continue
        if line.startswith("END"):
            if cur:
                frames.append(_assemble(cur))
                cur = []
            break
        if line.startswith("MODEL"):
            if cur:
                frames.append(_assemble(cur))
                cur = []
            continue
        if not line.startswith("ATOM"):
            continue
        atom = _parse_atom(line, lineno)
        if atom["alt"] not in (" ", "A"):
            continue
        aa = three.get(atom["res"].upper()) if len(atom["res"]) == 3 else None
        if aa is None:
            if atom["res"].upper() in ("A", "U", "G", "C", "RA", "RU", "RG", "RC"):
                continue            # ribonucleotides are skipped (pdb_read.rs:83-91)
            raise PdbReadError("line %d: residue name %r is not one of the 20 standard amino acids"
                               % (lineno, atom["res"]))
        name = normalise_atom_name(atom["name"])
        canon = _canonical(aa, name, sc)
        if canon is None:
            if name in TERMINAL_PATCH:
                continue
            raise PdbReadError("line %d: atom %r is not part of residue %s" % (lineno, atom["name"], aa))
        el = _element(atom["element"], canon)
        if el is None:
            raise PdbReadError("line %d: can't determine element for atom %r" % (lineno, canon))
        cur.append(((atom["chain"], atom["seq"], aa), canon, el, atom["xyz"]))
    if cur:
        frames.append(_assemble(cur))
    if not frames:
        raise PdbReadError("PDB contains no ATOM records")
    return frames


def read_pdb(src: Union[str, TextIO]) -> List[S.Residue]:
    """The first model (NMR ensembles), every chain (pdb_read.rs:158-297)."""
    lines = []
    seen_atom = False
    for line in _lines(src):
        if line.startswith("END") or (line.startswith("MODEL") and seen_atom):
            if seen_atom:
                break
            continue
        if line.startswith("ATOM"):
            seen_atom = True
        lines.append(line)
    return read_pdb_trajectory("\n".join(lines) + "\n")[0]


def _atom_name_field(name: str) -> str:
    """Columns 13-16 (pdb.rs:162-178): 4-character names fill the field,
    shorter ones start in column 14."""
    return name if len(name) == 4 else " " + name.ljust(3)


def _write_atoms(out: List[str], residues: Sequence[S.Residue], pos: Optional[np.ndarray]):
    serial = 1
    k = 0
    prev_chain = None
    for ri, r in enumerate(residues):
        if prev_chain is not None and prev_chain != r.chain:
            out.append("TER   %5d      %3s %s%4d" % (serial, "", prev_chain, ri))
        name3 = r.aa.upper()
        for nm, el, p in zip(r.names, r.elements, r.pos):
            x, y, z = (pos[k] if pos is not None else p)
            k += 1
            out.append("ATOM  %5d %-4s %3s %s%4d    %8.3f%8.3f%8.3f  1.00  0.00          %2s"
                       % (serial, _atom_name_field(nm), name3, r.chain, ri + 1, x, y, z, el))
            serial += 1
        prev_chain = r.chain
    if residues:
        last = residues[-1]
        out.append("TER   %5d      %3s %s%4d" % (serial, last.aa.upper(), last.chain, len(residues)))


def _header(out: List[str], title: str):
    out.append("REMARK   1 Generated by origami (experimental physics-based folder)")
    if title:
        out.append("REMARK   2 Title: %s" % title)


def write_pdb(residues: Sequence[S.Residue], title: str = "", pos: Optional[np.ndarray] = None) -> str:
    """Single-model PDB text (no MODEL records); pos [N,3] overrides the
    residues' coordinates."""
    out: List[str] = []
    _header(out, title)
    _write_atoms(out, residues, pos)
    out.append("END")
    return "\n".join(out) + "\n"


def write_pdb_trajectory(residues: Sequence[S.Residue], frames: Iterable[np.ndarray],
                         title: str = "") -> str:
    """One MODEL block per frame (frames: [N,3] coordinate arrays in the
    residues' atom order)."""
    out: List[str] = []
    _header(out, title)
    for i, f in enumerate(frames):
        out.append("MODEL     %4d" % (i + 1))
        _write_atoms(out, residues, np.asarray(f, dtype=np.float64).reshape(-1, 3))
        out.append("ENDMDL")
    out.append("END")
    return "\n".join(out) + "\n"


def rmsd_points(p: np.ndarray, q: np.ndarray) -> float:
    """Optimal-rotation RMSD (Kabsch): centre both sets, H = P^T Q,
    R = V diag(1, 1, sign det(V U^T)) U^T."""
    p = np.asarray(p, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    if p.shape != q.shape or len(p) == 0:
        raise ValueError("rmsd_points: mismatched or empty inputs")
    pc = p - p.mean(0)
    qc = q - q.mean(0)
    u, _, vt = np.linalg.svd(pc.T @ qc)
    v = vt.T
    d = np.sign(np.linalg.det(v @ u.T)) or 1.0
    rot = v @ np.diag([1.0, 1.0, d]) @ u.T
    diff = pc @ rot.T - qc
    return float(np.sqrt((diff ** 2).sum() / len(p)))


def ca_positions(residues: Sequence[S.Residue], pos: Optional[np.ndarray] = None) -> Optional[np.ndarray]:
    out = []
    k = 0
    for r in residues:
        found = None
        for nm, p in zip(r.names, r.pos):
            if nm == "CA":
                found = pos[k] if pos is not None else p
            k += 1
        if found is None:
            return None
        out.append(found)
    return np.asarray(out, dtype=np.float64)


def rmsd_ca(a: Sequence[S.Residue], b: Sequence[S.Residue], pos_a=None, pos_b=None) -> Optional[float]:
    """Cα RMSD of two structures with the same residues (None when the
    residue counts differ or a residue lacks its Cα)."""
    if len(a) != len(b) or len(a) == 0:
        return None
    p, q = ca_positions(a, pos_a), ca_positions(b, pos_b)
    if p is None or q is None:
        return None
    return rmsd_points(p, q)


def system_from_residues(residues: Sequence[S.Residue]) -> S.ResolvedSystem:
    """Topology and CHARMM36 parameters for a structure read from a PDB."""
    return S.resolve(list(residues))
