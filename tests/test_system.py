"""Input construction pinned by origami's own known answers
(builder / topology / force-field tests)."""
import math

import numpy as np
import pytest

from origami_amd import system as S


def test_atom_counts():
    # builder.rs:383-389 (Ala = 10 atoms); cotranslate_acceptance.rs:47-49 (AGAGA = 44)
    assert len(S.build_chain(["Ala"])[0].names) == 10
    assert sum(len(r.names) for r in S.build_chain(["Ala", "Gly", "Ala", "Gly", "Ala"])) == 44
    g = S.build_chain(["Gly"])[0]
    assert "HA2" in g.names and "HA3" in g.names and "CB" not in g.names
    pro = S.build_chain(["Ala", "Pro"])[1]
    assert "H" not in pro.names and "CD" in pro.names


def test_bond_counts():
    # topology_graph.rs:397-414
    assert len(S.build_topology(S.build_chain(["Ala"])).bonds) == 9
    assert len(S.build_topology(S.build_chain(["Gly"])).bonds) == 6


def test_backbone_geometry():
    # builder.rs:404-424 / io/tests/build_all_residues.rs:22-37
    res = S.build_chain(["Ala", "Ala", "Ala"])
    for i, r in enumerate(res):
        n, ca, c, o = (np.array(r.position(x)) for x in ("N", "CA", "C", "O"))
        assert abs(np.linalg.norm(ca - n) - 1.458) < 1e-9
        assert abs(np.linalg.norm(c - ca) - 1.525) < 1e-9
        assert abs(np.linalg.norm(o - c) - 1.231) < 1e-9
        if i > 0:
            pc = np.array(res[i - 1].position("C"))
            assert abs(np.linalg.norm(n - pc) - 1.329) < 1e-9


def test_no_clash_extended_ala3():
    # builder.rs:440-462
    s = S.build_system(["Ala", "Ala", "Ala"])
    d = np.linalg.norm(s.pos[:, None] - s.pos[None], axis=-1)
    np.fill_diagonal(d, 9.0)
    assert d.min() > 0.7


def test_nerf_reproduces_internal_coordinates():
    # nerf.rs:125-149
    a, b, c = (0.0, 0.0, 0.0), (1.5, 0.0, 0.0), (2.5, 1.2, 0.0)
    d = np.array(S.place_atom(a, b, c, 1.4, math.radians(110.0), math.radians(60.0)))
    assert abs(np.linalg.norm(d - c) - 1.4) < 1e-9
    cb = np.array(b) - c
    cd = d - c
    ang = math.acos(np.dot(cb, cd) / np.linalg.norm(cb) / np.linalg.norm(cd))
    assert abs(ang - math.radians(110.0)) < 1e-9


def test_force_field_known_answers():
    # forcefield.rs:622-714 checks these within tolerances; the values are the
    # exact entries of par_all36m_prot.prm / top_all36_prot.rtf
    ff = S.standard_ff()
    T = S.TYPE_INDEX
    assert ff.bond(T["CT1"], T["CT3"]) == (222.5, 1.538)          # prm: CT3 CT1 222.500 1.5380
    assert ff.angle(T["NH1"], T["CT1"], T["C"]) == (50.0, 107.0)   # prm: NH1 CT1 C 50.000 107.0000
    assert ff.nonbonded_params(T["C"])[:2] == (0.11, 2.0)         # prm: C 0.0 -0.110000 2.000000
    assert ff.nonbonded_params(T["H"])[:2] == (0.046, 0.2245)     # prm: H 0.0 -0.046000 0.224500
    # CT1 carries 1-4 parameters: -0.0320 2.000 0.0 -0.01 1.9
    assert ff.nonbonded_params(T["CT1"]) == (0.032, 2.0, 0.01, 1.9)
    # wildcard X CT2 CT3 X 0.1600 3 0.00
    assert ff.dihedral(T["HA3"], T["CT3"], T["CT2"], T["HA2"]) == [(0.16, 3, 0.0)]
    assert ff.partial_charge("Ala", "N") == -0.47
    assert ff.partial_charge("Ala", "CA") == 0.07
    assert ff.partial_charge("Ala", "H") == 0.31
    assert ff.partial_charge("Leu", "HB2") == 0.09
    assert ff.partial_charge("Leu", "HB3") == 0.09
    assert ff.partial_charge("Ile", "CD1") == -0.27                # rtf ILE CD CT3 -0.27
    assert ff.partial_charge("Ile", "HD11") == 0.09
    assert ff.partial_charge("His", "HD1") == 0.32                 # HSD
    assert ff.partial_charge("His", "ND1") == -0.36
    # O X X C 120.0000 0 0.0000 (peptide carbonyl improper)
    assert ff.improper(T["CT1"], T["C"], T["O"], T["NH1"]) == (120.0, 0.0)


def test_classification_known_answers():
    # atom_type.rs:354-400
    T = S.TYPE_INDEX
    assert S.classify("Ala", "N") == T["NH1"]
    assert S.classify("Pro", "N") == T["N"]
    assert S.classify("Gly", "CA") == T["CT2"]
    assert S.classify("Pro", "CA") == T["CP1"]
    assert S.classify("Gly", "HA2") == T["HB2"]
    assert S.classify("Asp", "CB") == T["CT2A"]
    assert S.classify("Phe", "CZ") == T["CA"]
    assert S.classify("Phe", "HZ") == T["HP"]
    assert S.classify("Gly", "CB") is None


def test_every_residue_builds_and_resolves():
    for aa in S.AMINO_ACIDS:
        s = S.build_system(["Ala", aa, "Ala"])
        assert s.missing_params == 0, aa
        assert np.isfinite(s.pos).all()


def test_exclusions_one_three_one_four():
    # topology_graph.rs:455-476 (N-C is 1-3, N-O is 1-4 within Ala)
    s = S.build_system(["Ala"])
    idx = {n: i for i, n in enumerate(s.names)}
    excl = {tuple(p) for p in s.excl_pairs.tolist()}
    p14 = {tuple(p) for p in s.pair14.tolist()}
    key = lambda a, b: (min(idx[a], idx[b]), max(idx[a], idx[b]))
    assert key("N", "CA") in excl and key("N", "C") in excl
    assert key("N", "O") in p14


def test_rng_streams():
    # rng.rs:27-98.  SplitMix64 from seed 0 starts 0xE220A8397B1DCDAF.
    r = S.Xoshiro256pp(0)
    assert r.s[0] == 0xE220A8397B1DCDAF
    a, b = S.Xoshiro256pp(123), S.Xoshiro256pp(123)
    assert [a.next_u64() for _ in range(100)] == [b.next_u64() for _ in range(100)]
    r = S.Xoshiro256pp(42)
    u = np.array([r.next_f64() for _ in range(100000)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.002
    r = S.Xoshiro256pp(7)
    g = np.array([r.gaussian() for _ in range(200000)])
    assert abs(g.mean()) < 0.02 and abs(g.var() - 1.0) < 0.02


def test_seeded_configs_sizes():
    assert S.seeded_chain(1, 20).n_atoms == 340
    assert S.seeded_chain(2, 80).n_atoms == 1313
