"""The five benchmark workloads run on the GPU and checked against the FP64 oracle.

  C1  20-residue seeded chain, 1000 BAOAB steps at dt = 1 fs, 310 K, 2/ps:
      every step's forces and energies against the oracle at the GPU's own
      coordinates, and single integrator steps against the oracle's step.
  C2  80-residue seeded chain, L-BFGS, then 2000 steps at dt = 2 fs with
      SHAKE and PowerSasa forces (gamma scale 0.25): force RMS against the
      oracle every 100 steps.
  C3  REMD ladder of 64 replicas of a seeded 10-residue chain, 280-395 K,
      swaps every 1 ps at dt = 2 fs with SHAKE.
  C4  the uneven batch of 64 chains (10-300 residues) with SASA: every
      chain's forces and energies against the oracle.
  C5  learned Born radii on the C4 batch: force error against the oracle
      and the fallback verdict per chain.

Dynamics start from L-BFGS-relaxed structures: origami's extended-chain
builder leaves ~1e4 kJ/mol/Å forces (a 5.5 Å Trp ring bond, a 1.95 Å Pro
N-CD bond), on which BAOAB at dt = 1 fs runs away within 100 steps on the
oracle itself.
"""
import concurrent.futures as cf
import copy
import os

import numpy as np
import pytest

from oracle import oracle as O
from origami_amd import system as S

pytestmark = pytest.mark.gpu

FORCE_RMS_TOL = 1e-5


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def energy_close(got, ref, rel=1e-9, floor=1e-6):
    """FP64 energy terms: relative 1e-9 with an absolute floor of 1e-6 kJ/mol
    (fixed-point units are 2.3e-10 kJ/mol per contribution)."""
    return abs(got - ref) <= rel * abs(ref) + floor


def threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


@pytest.fixture(scope="module")
def E():
    from origami_amd import engine
    return engine


def relaxed(E, s, steps=300):
    from origami_amd.minimize import minimize
    eng = E.Engine([s])
    res = minimize(eng, max_steps=steps, mode_a=False)[0]
    out = copy.copy(s)
    out.pos = eng.positions()
    return out, res


def test_c1_thousand_steps_per_step_parity(E, c1_system):
    s, _ = relaxed(E, c1_system)
    orc = O.Oracle(s)
    eng = E.Engine([s])
    eng.set_velocities(E.maxwell_boltzmann(s, 310.0, S.Xoshiro256pp(1)))
    seed = 20240601
    eng.dynamics_setup([310.0], dt_fs=1.0, friction_ps_inv=2.0, seed=seed)
    worst_f = 0.0
    worst_e = {}
    for step in range(1000):
        x0, v0, f0 = eng.positions(), eng.velocities(), eng.forces()
        eng.run(1)
        x, f = eng.positions(), eng.forces()
        # the step's own forces at the step's own coordinates
        fo = orc.total_force(x)
        err = rel_rms(f, fo)
        worst_f = max(worst_f, err)
        assert err < FORCE_RMS_TOL, (step, err)
        got = eng.total_energy()[0]
        ref = orc.total_energy(x)
        for k in ("bond", "angle", "dihedral", "improper", "lj", "coulomb", "gb_self",
                  "gb_pair"):
            assert energy_close(got[k], ref[k]), (step, k, got[k], ref[k])
            worst_e[k] = max(worst_e.get(k, 0.0), abs(got[k] - ref[k]) / max(abs(ref[k]), 1e-300))
        if step % 250 == 0 or step == 999:
            # the integrator step itself from the GPU's state, with its noise
            xi = O.philox_normals(seed, 0, s.n_atoms, step)
            px, pv, pf, fails = orc.baoab_step(x0, v0, f0, xi, 1.0, 310.0, 2.0, False)
            assert fails == 0
            assert np.abs(x - px).max() < 1e-9, step
            assert rel_rms(f, pf) < FORCE_RMS_TOL
    print("C1 1000 steps: worst force rel RMS %.3e, worst energy rel err %s" % (worst_f, worst_e))
    assert np.isfinite(eng.positions()).all()
    assert not eng.status()["diverged"].any()


def test_c2_lbfgs_then_sasa_dynamics(E, c2_system):
    from origami_amd.minimize import minimize
    eng0 = E.Engine([c2_system])
    res = minimize(eng0, mode_a=False)[0]          # minimize.rs:27-38 defaults, all pairs
    print("C2 L-BFGS", res)
    assert res.final_energy < res.initial_energy
    s = copy.copy(c2_system)
    s.pos = eng0.positions()
    orc = O.Oracle(s)
    eng = E.Engine([s])
    eng.set_velocities(E.maxwell_boltzmann(s, 310.0, S.Xoshiro256pp(2)))
    eng.dynamics_setup([310.0], dt_fs=2.0, friction_ps_inv=2.0, seed=2, constrain_h_bonds=True,
                       include_sasa=True, sasa_gamma_scale=0.25)
    rms = []
    for chunk in range(20):
        eng.run(100)
        x, f = eng.positions(), eng.forces()
        fo = orc.total_force(x) + orc.sasa_forces(x, 0.25)[0]
        rms.append(rel_rms(f, fo))
    print("C2 2000 steps, force rel RMS every 100 steps: max %.3e mean %.3e"
          % (max(rms), float(np.mean(rms))))
    assert max(rms) < FORCE_RMS_TOL
    st = eng.status()
    assert not st["diverged"].any() and st["shake_failures"][0] == 0
    x = eng.positions()
    d = np.linalg.norm(x[s.constraints[:, 0]] - x[s.constraints[:, 1]], axis=1)
    assert np.abs(d ** 2 - s.constraint_len ** 2).max() < 1e-6


def test_c3_remd_64_replicas(E, c3_system):
    s, _ = relaxed(E, c3_system)
    temps = E.geometric_ladder(280.0, 395.0, 64)
    eng, summ = E.run_remd(s, temps, total_time_fs=4000.0, swap_interval_fs=1000.0, dt_fs=2.0,
                           friction_ps_inv=2.0, seed=3, constrain_h_bonds=True)
    acc = summ.acceptance_ratios()
    print("C3 64 replicas: attempts %d, accepts %d, acceptance min %.2f mean %.2f"
          % (sum(summ.swap_attempts), sum(summ.swap_accepts), min(acc), float(np.mean(acc))))
    # 4 rounds of alternating parity: every adjacent pair is attempted twice
    assert summ.swap_attempts == [2] * 63
    assert sum(summ.swap_accepts) > 0
    assert not any(summ.diverged)
    assert np.isfinite(summ.final_potential_energy).all()
    # every replica's forces against the oracle at its final coordinates
    orc = O.Oracle(s)
    xs, fs = eng.split(eng.positions()), eng.split(eng.forces())
    # run_remd's forces use origami's integrator cutoffs (remd.rs:209-214)
    errs = [rel_rms(fs[k], orc.total_force(xs[k], **O.MODE_A)) for k in range(64)]
    print("C3 replica force rel RMS max %.3e" % max(errs))
    assert max(errs) < FORCE_RMS_TOL
    assert eng.status()["shake_failures"].sum() == 0


@pytest.fixture(scope="module")
def c4_batch():
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    systems, temps = bench.build_config4()
    rng = np.random.default_rng(44)
    pos = [s.pos + rng.normal(scale=0.02, size=s.pos.shape) for s in systems]
    return systems, temps, pos


def oracle_forces_sasa(systems, pos, gamma_scale):
    def one(k):
        orc = O.Oracle(systems[k])
        return orc.total_force(pos[k]) + orc.sasa_forces(pos[k], gamma_scale)[0]
    with cf.ThreadPoolExecutor(max_workers=threads()) as ex:
        return list(ex.map(one, range(len(systems))))


def test_c4_batch_every_chain_against_oracle(E, c4_batch):
    systems, temps, pos = c4_batch
    eng = E.Engine(systems)
    assert eng.n_systems == 64 and eng.n_atoms > 150000
    f, en = eng.total_force(np.concatenate(pos), include_sasa=True, sasa_gamma_scale=0.25,
                            energies=True)
    ref = oracle_forces_sasa(systems, pos, 0.25)
    errs = [rel_rms(fk, rk) for fk, rk in zip(eng.split(f), ref)]
    worst = int(np.argmax(errs))
    print("C4 per-chain force rel RMS: max %.3e (chain %d, %d atoms), median %.3e"
          % (errs[worst], worst, systems[worst].n_atoms, float(np.median(errs))))
    assert max(errs) < FORCE_RMS_TOL

    def energies(k):
        orc = O.Oracle(systems[k])
        e = orc.total_energy(pos[k])
        e["sasa"] = float((systems[k].sasa_gamma * orc.sasa_area(pos[k])[1]).sum())
        return e
    with cf.ThreadPoolExecutor(max_workers=threads()) as ex:
        eref = list(ex.map(energies, range(len(systems))))
    e_sasa = eng.total_energy(np.concatenate(pos), include_sasa=True)
    for k in range(len(systems)):
        for t in ("bond", "angle", "dihedral", "improper", "lj", "coulomb", "gb_self", "gb_pair"):
            assert energy_close(en[k][t], eref[k][t]), (k, t, en[k][t], eref[k][t])
        assert energy_close(e_sasa[k]["sasa"], eref[k]["sasa"]), (k, e_sasa[k]["sasa"], eref[k]["sasa"])


def test_c5_surrogate_on_c4_batch(E, c4_batch):
    systems, temps, pos = c4_batch
    eng = E.Engine(systems)
    rng = np.random.default_rng(55)
    eng.surrogate_setup(4 * eng.n_atoms, seed=99)
    for k in range(4):          # plain-GB radii at four perturbed configurations
        eng.set_positions(np.concatenate([p + rng.normal(scale=0.05, size=p.shape) for p in pos]))
        eng.surrogate_collect()
    loss = eng.surrogate_train(1500, batch=8192, lr=2e-3, seed=5)
    assert loss[-20:].mean() < 1e-3 * loss[:20].mean()
    eng.surrogate_enable(check_every=1, tolerance=0.02)
    x = np.concatenate(pos)
    f_sur = eng.total_force(x, include_sasa=True, sasa_gamma_scale=0.25, surrogate=True)
    mode, dev = eng.surrogate_status()
    ref = oracle_forces_sasa(systems, pos, 0.25)
    errs = np.array([rel_rms(fk, rk) for fk, rk in zip(eng.split(f_sur), ref)])
    print("C5 surrogate: per-chain force rel RMS vs FP64 oracle max %.3e mean %.3e; "
          "radius deviation max %.4f; %d of 64 chains on the surrogate"
          % (errs.max(), errs.mean(), dev.max(), int(mode.sum())))
    assert errs.max() <= 0.01
    # fallback verdict per chain: on the surrogate exactly where the radii
    # stayed within the 2 % tolerance of plain GB
    assert np.array_equal(mode.astype(bool), dev <= 0.02)
    # plain GB (the fallback and the judge) stays at the 1e-5 bar
    f_plain = eng.total_force(x, include_sasa=True, sasa_gamma_scale=0.25)
    assert max(rel_rms(fk, rk) for fk, rk in zip(eng.split(f_plain), ref)) < FORCE_RMS_TOL
