"""PDB input/output and Cα RMSD (io/src/pdb.rs, io/src/pdb_read.rs,
geom/src/rmsd.rs), and the reference's own systems as inputs: the public PDB
entries its io tests hold (copied into tests/golden/), pinned by the atom
counts and energy gap its README publishes."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from origami_amd import pdb as P
from origami_amd import system as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    return os.path.join(GOLDEN, name + ".pdb")


def test_writes_valid_alanine_pdb():
    # pdb.rs:182-195
    text = P.write_pdb(S.build_chain(["Ala"]), "single Ala test")
    assert text.startswith("REMARK") and "ALA" in text and text.endswith("END\n")
    assert sum(1 for l in text.splitlines() if l.startswith("ATOM")) == 10


def test_atom_columns_are_correctly_aligned():
    # pdb.rs:197-212
    line = next(l for l in P.write_pdb(S.build_chain(["Gly"])).splitlines() if l.startswith("ATOM"))
    assert line[0:6] == "ATOM  " and line[12:16] == " N  " and line[17:20] == "GLY"
    assert line[21:22] == "A"
    assert line[76:78] == " N"
    # a 4-character name fills columns 13-16
    res = S.build_chain(["Ile"])
    four = next(l for l in P.write_pdb(res).splitlines() if l[12:16].strip() == "HD11")
    assert four[12:16] == "HD11"


def test_trajectory_writes_one_model_per_frame():
    # pdb.rs:214-227
    res = S.build_chain(["Ala", "Gly"])
    pos = np.array([p for r in res for p in r.pos])
    text = P.write_pdb_trajectory(res, [pos, pos, pos], "traj test")
    lines = text.splitlines()
    assert sum(l.startswith("MODEL") for l in lines) == 3
    assert sum(l.startswith("ENDMDL") for l in lines) == 3
    assert text.endswith("END\n")


def test_round_trip_all_residues_and_trajectory():
    # pdb_read.rs:428-458, :494-518
    res = S.build_chain(list(S.AMINO_ACIDS))
    back = P.read_pdb(P.write_pdb(res, "all-20"))
    assert len(back) == 20
    assert [r.aa for r in back] == [r.aa for r in res]
    for a, b in zip(res, back):
        assert a.names == b.names and a.elements == b.elements
        assert np.abs(np.array(a.pos) - np.array(b.pos)).max() < 1e-3
    pos = np.array([p for r in res for p in r.pos])
    p2 = pos.copy()
    p2[0, 0] += 0.5
    frames = P.read_pdb_trajectory(P.write_pdb_trajectory(res, [pos, p2, pos]))
    assert len(frames) == 3
    assert abs((frames[1][0].pos[0][0] - frames[0][0].pos[0][0]) - 0.5) < 1e-3


def test_normalise_legacy_atom_name_and_errors():
    # pdb_read.rs:460-480
    assert P.normalise_atom_name("1HD1") == "HD11"
    assert P.normalise_atom_name("HD11") == "HD11"
    assert P.normalise_atom_name("CA") == "CA"
    with pytest.raises(P.PdbReadError, match="standard amino acids"):
        P.read_pdb("ATOM      1  N   XXX A   1       0.000   0.000   0.000  1.00  0.00           N\n")
    with pytest.raises(P.PdbReadError, match="no ATOM"):
        P.read_pdb("REMARK only\n")


def test_handles_alt_loc():
    # pdb_read.rs:520-536
    pdb = ("ATOM      1  N  AALA A   1       0.000   0.000   0.000  0.50  0.00           N\n"
           "ATOM      2  N  BALA A   1       1.000   0.000   0.000  0.50  0.00           N\n"
           "ATOM      3  CA  ALA A   1       1.458   0.000   0.000  1.00  0.00           C\n"
           "TER\n")
    r = P.read_pdb(pdb)[0]
    assert r.names == ["N", "CA"] and r.pos[0] == (0.0, 0.0, 0.0)


def test_reads_the_reference_systems():
    # README.md:406-410 lists 134, 300 and 520 atoms; pdb_read.rs:482-492
    # reads 1L2Y as NLYIQWLKDGGPSSGRPPPS
    for name, n_atoms in (("1UAO_chignolin", 134), ("1L2Y_model1", 300), ("2F4K_villin_hp35", 520)):
        res = P.read_pdb(fixture(name))
        s = P.system_from_residues(res)
        assert s.n_atoms == n_atoms, name
        assert s.missing_params == 0, name
    seq = "".join(S.ONE_LETTER[r.aa] for r in P.read_pdb(fixture("1L2Y_model1")))
    assert seq == "NLYIQWLKDGGPSSGRPPPS"
    frames = P.read_pdb_trajectory(fixture("1UAO_chignolin"))
    assert len(frames) == 18                            # the NMR ensemble's models
    assert all(sum(len(r.names) for r in f) == 134 for f in frames)


def test_native_trp_cage_energy_gap():
    # README.md:33-34: native 1L2Y scores 49 170 kJ/mol below the extended
    # chain with origami's minimiser energy (energy_eval.rs:20-34: bonded +
    # 10 Å LJ / reaction-field Coulomb + all-pairs GB)
    nat = P.system_from_residues(P.read_pdb(fixture("1L2Y_model1")))
    ext = S.build_system([S.FROM_ONE_LETTER[c] for c in "NLYIQWLKDGGPSSGRPPPS"])
    e_nat = O.Oracle(nat).total_energy(nat.pos, cut_nb=10.0)["total"]
    e_ext = O.Oracle(ext).total_energy(ext.pos, cut_nb=10.0)["total"]
    assert abs((e_ext - e_nat) - 49170.0) < 5e-3 * 49170.0


def rotation(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def test_rmsd_kabsch():
    # rmsd.rs:94-160: zero for a rigid motion, the plain RMS for a pure offset
    # pattern that no rotation improves, and scipy's optimal alignment
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(3)
    q = rng.normal(size=(40, 3)) * 5.0
    p = q @ rotation([1, 2, 3], 1.1).T + np.array([4.0, -2.0, 9.0])
    assert P.rmsd_points(p, q) < 1e-10
    # a reflection is not a rotation: the sign correction keeps it > 0
    assert P.rmsd_points(q * np.array([1.0, 1.0, -1.0]), q) > 1.0
    p2 = q + rng.normal(size=q.shape) * 0.7
    _, rssd = Rotation.align_vectors(q - q.mean(0), p2 - p2.mean(0))
    assert abs(P.rmsd_points(p2, q) - rssd / math.sqrt(len(q))) < 1e-9
    res = P.read_pdb(fixture("1L2Y_model1"))
    assert P.rmsd_ca(res, res) < 1e-12
    assert P.rmsd_ca(res, res[:-1]) is None
