"""Shrake-Rupley SASA energy on the GPU (og_energies with OG_SASA_SR, the
term of origami's REMD swap energies, remd.rs:529-539, sasa.rs:113-178)
against the oracle's restatement: the same dots and the same buried test,
so the same dot counts and energies."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from origami_amd import pdb as P
from origami_amd import system as S

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def E():
    from origami_amd import engine
    return engine


def jittered(s, seed, scale=0.02):
    return s.pos + np.random.default_rng(seed).normal(scale=scale, size=s.pos.shape)


def test_sr_energy_matches_oracle(E, c1_system, c2_system):
    fx = P.system_from_residues(P.read_pdb(os.path.join(GOLDEN, "2F4K_villin_hp35.pdb")))
    systems = [c1_system, c2_system, fx]
    pos = [jittered(c1_system, 3), jittered(c2_system, 4), fx.pos]
    eng = E.Engine(systems)
    got = eng.total_energy(np.concatenate(pos), mode_a=True, include_sasa=True,
                           sasa_method="shrake_rupley")
    for k, s in enumerate(systems):
        ref, area = O.sr_energy(pos[k], s.sasa_radius, s.sasa_gamma)
        print("system %d: Shrake-Rupley energy gpu %.12f oracle %.12f kJ/mol (area %.2f A^2)"
              % (k, got[k]["sasa"], ref, area.sum()))
        assert abs(got[k]["sasa"] - ref) <= 1e-9 * abs(ref) + 1e-6
    # the other terms are those of the evaluation without SASA
    plain = eng.total_energy(np.concatenate(pos), mode_a=True)
    for k in range(3):
        for t in ("bond", "lj", "coulomb", "gb_pair"):
            assert got[k][t] == plain[k][t]
    # a force evaluation does not take the dot-density term (it has no gradient)
    f = np.zeros((eng.n_atoms, 3))
    with pytest.raises(E.EngineError, match="energy term only"):
        E._check(E.lib().og_forces(eng._h, None, f.ctypes.data_as(E._dp), E.SASA_SR, 10.0, 1.0, None))


def test_remd_swap_energies_carry_shrake_rupley(E, c3_system):
    """remd::potential_energy with include_sasa: the swap energies are the
    10 A energies plus the Shrake-Rupley term of each replica."""
    from origami_amd.minimize import minimize_device
    e0 = E.Engine([c3_system])
    minimize_device(e0, max_steps=200, mode_a=True)
    s = S.ResolvedSystem(**{**c3_system.__dict__, "pos": e0.positions()})
    temps = E.geometric_ladder(280.0, 395.0, 4)
    eng, summ = E.run_remd(s, temps, total_time_fs=200.0, swap_interval_fs=50.0, dt_fs=2.0,
                           friction_ps_inv=2.0, seed=9, constrain_h_bonds=True, include_sasa=True,
                           sasa_gamma_scale=0.25)
    assert summ.swap_attempts == [2, 2, 2] or sum(summ.swap_attempts) > 0
    xs = eng.split(eng.positions())
    for k in range(4):
        o = O.Oracle(s)
        ref = o.total_energy(xs[k], cut_nb=10.0, cut_descreen=20.0)["total"]
        ref += O.sr_energy(xs[k], s.sasa_radius, s.sasa_gamma)[0]
        assert abs(summ.final_potential_energy[k] - ref) <= 1e-9 * abs(ref) + 1e-6, k
