"""GPU parity: the HIP engine against the FP64 oracle on identical inputs.

Bar (from the engine's contract): FP32 pair arithmetic with int64
fixed-point accumulation must stay within 1e-5 relative force RMS of the
FP64 oracle; FP64 paths (bonded, SHAKE, integrator, SASA) within 1e-9.
"""
import copy
import math

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


@pytest.fixture(scope="module")
def E():
    from origami_amd import engine
    return engine


def test_library_is_native(E):
    import ctypes
    lib = E.lib()
    assert isinstance(lib, ctypes.CDLL)
    assert lib._name.endswith("liborigami_gpu.so")


@pytest.mark.parametrize("mode_a", [False, True])
def test_total_force_c1(E, c1_system, mode_a):
    s = c1_system
    orc = O.Oracle(s)
    cut = O.MODE_A if mode_a else O.MODE_B
    fref, terms = orc.total_force(s.pos, terms=True, **cut)
    eng = E.Engine([s])
    f = eng.total_force(mode_a=mode_a)
    err = rel_rms(f, fref)
    print("C1 mode_a=%s rel force RMS %.3e" % (mode_a, err))
    assert err < FORCE_RMS_TOL
    # Newton's third law: every term is pairwise
    assert np.abs(f.sum(0)).max() < 1e-6 * np.abs(fref).max()


def test_born_radii_and_integral(E, c2_system):
    s = c2_system
    orc = O.Oracle(s)
    integ_ref, radii_ref, clamped = orc.born_radii(s.pos, math.inf)
    eng = E.Engine([s])
    eng.total_force()
    radii, integ = eng.born_radii()
    assert np.array_equal(radii, s.gb_rho - 0.09)       # R_i == rho~_i exactly
    assert np.allclose(radii, radii_ref, rtol=0, atol=0)
    rel = np.abs(integ - integ_ref) / np.abs(integ_ref)
    print("descreening integral max rel err %.3e" % rel.max())
    assert rel.max() < 1e-4
    assert eng.status()["clamped"][0] == 0


@pytest.mark.parametrize("mode_a", [False, True])
def test_total_force_c2(E, c2_system, mode_a):
    s = c2_system
    orc = O.Oracle(s)
    cut = O.MODE_A if mode_a else O.MODE_B
    fref = orc.total_force(s.pos, **cut)
    eng = E.Engine([s])
    f = eng.total_force(mode_a=mode_a)
    err = rel_rms(f, fref)
    print("C2 mode_a=%s rel force RMS %.3e" % (mode_a, err))
    assert err < FORCE_RMS_TOL


@pytest.fixture(scope="module")
def relaxed(E, c1_system, c2_system, c3_system):
    """The seeded chains after 300 L-BFGS steps (all pairs).  Dynamics straight
    from the extended builder output runs away within a few steps at dt = 2 fs
    (it leaves ~1e4 kJ/mol/A on strained rings), on the oracle as well."""
    from origami_amd.minimize import minimize
    out = {}
    for name, s in (("c1", c1_system), ("c2", c2_system), ("c3", c3_system)):
        eng = E.Engine([s])
        minimize(eng, max_steps=300, mode_a=False)
        r = copy.copy(s)
        r.pos = eng.positions()
        out[name] = r
    return out


def jittered(s, seed, scale=0.02):
    """Generic geometry: the builder's exact symmetries removed."""
    return s.pos + np.random.default_rng(seed).normal(scale=scale, size=s.pos.shape)


def test_sasa_forces_c2_generic_geometry(E, c2_system):
    s = c2_system
    pos = jittered(s, 2)
    orc = O.Oracle(s)
    fs, stats = orc.sasa_forces(pos, 0.25)
    eng = E.Engine([s])
    f0 = eng.total_force(pos)
    f1 = eng.total_force(pos, include_sasa=True, sasa_gamma_scale=0.25)
    err = np.abs((f1 - f0) - fs).max()
    print("SASA max abs diff %.3e (max |F| %.3e), stats %s" % (err, np.abs(fs).max(), stats))
    assert err < 1e-8       # FP64 kernel; int64 fixed-point resolution 2.3e-10
    fref = orc.total_force(pos) + fs
    assert rel_rms(f1, fref) < FORCE_RMS_TOL


def dense_packing(n, seed, density=0.12, min_sep=1.1):
    """n points at a dense protein-like packing (atoms per A^3) in a ball:
    interior atoms overlap 80-100 neighbours, more caps than the two-pass
    arrangement's first pass holds (80), so they go through its second (MAXC)
    pass."""
    rng = np.random.default_rng(seed)
    R = (3 * n / (4 * math.pi * density)) ** (1 / 3)
    pts = []
    while len(pts) < n:
        p = rng.uniform(-R, R, 3)
        if np.linalg.norm(p) > R:
            continue
        if pts and np.min(np.linalg.norm(np.asarray(pts) - p, axis=1)) < min_sep:
            continue
        pts.append(p)
    return np.asarray(pts)


def test_sasa_dense_packing_second_pass(E, c1_system):
    s = c1_system
    pos = dense_packing(s.n_atoms, 7)
    rad = s.sasa_radius
    d = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    over = ((d < rad[:, None] + rad[None]) & (d > 0))[s.sasa_gamma != 0].sum(1)
    assert (over > 80).sum() >= 10
    orc = O.Oracle(s)
    fs, stats = orc.sasa_forces(pos, 0.25)
    total, per_atom = orc.sasa_area(pos)
    e_ref = float(np.sum(s.sasa_gamma * per_atom))
    eng = E.Engine([s])
    f0 = eng.total_force(pos)
    got = {}
    for passes in (2, 1):       # 80 caps + second pass over the rest; one pass of 128
        eng.set_launch(sasa_passes=passes)
        f1 = eng.total_force(pos, include_sasa=True, sasa_gamma_scale=0.25)
        err = np.abs((f1 - f0) - fs).max()
        print("dense SASA, %d pass(es): max abs diff %.3e (max |F| %.3e), %d atoms over 80 "
              "neighbours, stats %s" % (passes, err, np.abs(fs).max(), (over > 80).sum(), stats))
        assert stats["atoms_with_arcs"] > 20
        assert err < 1e-8
        e_gpu = eng.total_energy(pos, include_sasa=True)[0]["sasa"]
        assert abs(e_gpu - e_ref) < 1e-9 * max(1.0, abs(e_ref))
        got[passes] = (f1, e_gpu)
    assert np.array_equal(got[1][0], got[2][0]) and got[1][1] == got[2][1]


def test_sasa_total_force_c2_built_geometry(E, c2_system):
    # The freshly built chain is exactly symmetric, so a handful of cap
    # triple points sit on origami's 1e-12 topology tolerances and single
    # SASA terms there depend on last-ulp rounding; the total force still
    # meets the bar.
    s = c2_system
    orc = O.Oracle(s)
    fref = orc.total_force(s.pos) + orc.sasa_forces(s.pos, 0.25)[0]
    eng = E.Engine([s])
    f1 = eng.total_force(include_sasa=True, sasa_gamma_scale=0.25)
    err = rel_rms(f1, fref)
    print("C2 + SASA (built geometry) rel force RMS %.3e" % err)
    assert err < FORCE_RMS_TOL


def test_energies_c2(E, c2_system):
    s = c2_system
    orc = O.Oracle(s)
    for mode_a in (False, True):
        cut = O.MODE_A if mode_a else O.MODE_B
        ref = orc.total_energy(s.pos, cut_nb=cut["cut_nb"], cut_descreen=cut["cut_descreen"])
        eng = E.Engine([s])
        got = eng.total_energy(mode_a=mode_a)[0]
        print("mode_a=%s" % mode_a, {k: (got[k], ref[k]) for k in ref if k in got})
        for k in ("bond", "angle", "dihedral", "improper", "lj", "coulomb", "gb_self", "gb_pair",
                  "total"):
            assert energy_close(got[k], ref[k]), (k, got[k], ref[k])


def test_uneven_batch_matches_single(E):
    rng = S.Xoshiro256pp(4)
    systems = []
    for k in range(5):
        L = 10 + int(rng.next_f64() * 60)
        systems.append(S.resolve(S.build_chain(S.random_sequence(rng, L))))
    eng = E.Engine(systems)
    pos = [jittered(s, 10 + k) for k, s in enumerate(systems)]
    f = eng.total_force(np.concatenate(pos), include_sasa=True, sasa_gamma_scale=0.25)
    for s, p, fs_ in zip(systems, pos, eng.split(f)):
        orc = O.Oracle(s)
        fref = orc.total_force(p) + orc.sasa_forces(p, 0.25)[0]
        assert rel_rms(fs_, fref) < FORCE_RMS_TOL


def test_bit_reproducible(E, c2_system):
    eng = E.Engine([c2_system])
    a = eng.total_force(include_sasa=True, sasa_gamma_scale=0.25)
    b = eng.total_force(include_sasa=True, sasa_gamma_scale=0.25)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("shake", [False, True])
def test_baoab_step_matches_oracle(E, c1_system, shake):
    s = c1_system
    orc = O.Oracle(s)
    eng = E.Engine([s])
    v0 = E.maxwell_boltzmann(s, 310.0, S.Xoshiro256pp(5))
    eng.set_velocities(v0)
    dt = 2.0 if shake else 1.0
    seed = 0x1234567890ABCDEF
    eng.dynamics_setup([310.0], dt_fs=dt, friction_ps_inv=2.0, seed=seed, constrain_h_bonds=shake)
    f0 = eng.forces()
    eng.run(1)
    x1, v1, f1 = eng.positions(), eng.velocities(), eng.forces()
    xi = O.philox_normals(seed, 0, s.n_atoms, 0)
    fo = orc.total_force(s.pos)
    assert rel_rms(f0, fo) < FORCE_RMS_TOL
    # integrate the oracle from the same state with the engine's own start
    # forces and noise, so the step itself is what is compared
    px, pv, pf, fails = orc.baoab_step(s.pos, v0, f0, xi, dt, 310.0, 2.0, shake)
    assert fails == 0
    assert np.abs(x1 - px).max() < 1e-9
    # velocities before the closing half-kick (which uses each side's own F(x1))
    kick = (S.ACCEL_FACTOR / s.mass * 0.5 * dt)[:, None]
    assert np.abs((v1 - f1 * kick) - (pv - pf * kick)).max() < 1e-9
    assert rel_rms(f1, pf) < FORCE_RMS_TOL


def test_thermostat_holds_temperature(E):
    s = S.build_system(["Ala", "Ala", "Ala"])
    eng, summ = E.run_langevin(s, 2000, dt_fs=1.0, temperature_k=310.0, friction_ps_inv=2.0,
                               seed=7)
    assert not summ.diverged
    assert abs(summ.temperature_mean_k - 310.0) < 60.0
    x = eng.positions()
    assert np.isfinite(x).all() and np.abs(x).max() < 1000.0


@pytest.fixture(scope="module")
def c1_minimised(E, c1_system):
    from origami_amd.minimize import minimize
    import copy
    eng = E.Engine([c1_system])
    res = minimize(eng, max_steps=300, mode_a=False)[0]
    s = copy.copy(c1_system)
    s.pos = eng.positions()
    return s, res


def test_lbfgs_lowers_energy(c1_minimised):
    s, res = c1_minimised
    print(res)
    assert res.final_energy < res.initial_energy - 1000.0
    assert res.line_search_failures <= 2


def test_shake_dt2_no_failures(E, c1_minimised):
    c1_system, _ = c1_minimised
    # dynamics/tests/shake_dt2.rs: 1000 steps at dt = 2 fs, T within 80 K
    eng, summ = E.run_langevin(c1_system, 1000, dt_fs=2.0, temperature_k=310.0,
                               friction_ps_inv=2.0, seed=11, constrain_h_bonds=True)
    assert not summ.diverged
    assert summ.shake_failures == 0
    assert abs(summ.temperature_mean_k - 310.0) < 80.0
    # constraints hold
    x = eng.positions()
    d = np.linalg.norm(x[c1_system.constraints[:, 0]] - x[c1_system.constraints[:, 1]], axis=1)
    assert np.abs(d ** 2 - c1_system.constraint_len ** 2).max() < 1e-6


def test_remd_runs_and_swaps(E, c3_system):
    import copy
    from origami_amd.minimize import minimize
    e0 = E.Engine([c3_system])
    minimize(e0, max_steps=300, mode_a=False)
    s = copy.copy(c3_system)
    s.pos = e0.positions()
    temps = E.geometric_ladder(280.0, 395.0, 8)
    eng, summ = E.run_remd(s, temps, total_time_fs=4000.0, swap_interval_fs=200.0,
                           dt_fs=2.0, friction_ps_inv=2.0, seed=3, constrain_h_bonds=True)
    assert sum(summ.swap_attempts) > 0
    assert max(summ.acceptance_ratios()) > 0.1
    assert np.isfinite(eng.positions()).all()


def test_trp_cage_minimise_acceptance(E):
    # dynamics/tests/trp_cage_minimise.rs:12-55: extended Trp-cage, L-BFGS with
    # max_steps 500, gradient_tol 5.0, max_step_a 0.1 and origami's minimiser
    # energy (10 Å cutoffs) -> below 1000 kJ/mol in fewer than 500 steps with
    # the bond term under 200 kJ/mol.
    from origami_amd.minimize import minimize
    s = S.build_system([S.FROM_ONE_LETTER[c] for c in "NLYIQWLKDGGPSSGRPPPS"])
    eng = E.Engine([s])
    res = minimize(eng, max_steps=500, gradient_tol=5.0, max_step_a=0.1, mode_a=True)[0]
    e = eng.total_energy(mode_a=True)[0]
    print(res, e)
    assert res.initial_energy > 40000.0          # +47 506 kJ/mol extended (README.md:34-36)
    assert res.final_energy < 1000.0
    assert res.steps < 500
    assert e["bond"] < 200.0


def test_graph_replay_matches_eager(E, relaxed):
    """Captured-step replay (og_run with graphs) against direct launches, with
    eager work and synchronisations between replays, every replay checked."""
    c1_system = relaxed["c1"]
    c3_system = relaxed["c3"]
    def make():
        eng = E.Engine([c1_system, c3_system, c3_system])
        v = np.concatenate([E.maxwell_boltzmann(s, 300.0, S.Xoshiro256pp(5 + k))
                            for k, s in enumerate([c1_system, c3_system, c3_system])])
        eng.set_velocities(v)
        eng.dynamics_setup([300.0, 320.0, 340.0], dt_fs=1.0, friction_ps_inv=2.0, seed=77,
                           constrain_h_bonds=True)
        return eng
    eager, graph = make(), make()
    eager.use_graphs(False)
    graph.use_graphs(True)
    for rep in range(4):
        eager.run(8)
        graph.run(8)
        assert np.array_equal(eager.positions(), graph.positions()), rep
        assert np.array_equal(eager.velocities(), graph.velocities()), rep
        fa, fb = eager.total_force(), graph.total_force()          # eager work between replays
        assert np.array_equal(fa, fb)
    graph.swap(1, 2)
    eager.swap(1, 2)
    eager.run(5)
    graph.run(5)
    assert np.array_equal(eager.positions(), graph.positions())


def test_autotune_keeps_results(E, c2_system):
    """Launch widths leave forces bit-identical; the strip length (the force
    pass's tile shape) changes only the FP32 rounding of per-strip sums."""
    eng = E.Engine([c2_system])
    s0 = eng.strip_tiles()
    f0 = eng.total_force(include_sasa=True, sasa_gamma_scale=0.25)
    ch = eng.autotune(include_sasa=True)
    print("autotune", ch)
    assert ch["pair_blocks"] > 0 and ch["sasa_blocks"] > 0 and ch["strip_tiles"] in (2, 4, 8)
    eng.strip_tiles(s0)
    f1 = eng.total_force(include_sasa=True, sasa_gamma_scale=0.25)
    assert np.array_equal(f0, f1)
    fref = O.Oracle(c2_system).total_force(c2_system.pos) + O.Oracle(c2_system).sasa_forces(
        c2_system.pos, 0.25)[0]
    for sj in (1, 2, 3, 8):
        eng.strip_tiles(sj)
        f = eng.total_force(include_sasa=True, sasa_gamma_scale=0.25)
        assert rel_rms(f, f0) < 1e-6, sj
        assert rel_rms(f, fref) < FORCE_RMS_TOL, sj
    with pytest.raises(E.EngineError, match="strip length"):
        eng.strip_tiles(9)


def test_sasa_side_stream_matches_serial(E, relaxed):
    """SASA on its own stream beside the pair kernels (the default) against the
    one-stream order that the per-kernel profile uses: forces, energies and
    captured SASA dynamics bitwise equal."""
    c1_system = relaxed["c1"]
    c2_system = relaxed["c2"]
    systems = [c1_system, c2_system]
    ovl, ser = E.Engine(systems), E.Engine(systems)
    ser.profile(True)
    for eng in (ovl, ser):
        eng.set_positions(np.concatenate([jittered(s, 11) for s in systems]))
    fa = ovl.total_force(include_sasa=True, sasa_gamma_scale=0.25)
    fb = ser.total_force(include_sasa=True, sasa_gamma_scale=0.25)
    assert np.array_equal(fa, fb)
    ea = ovl.total_energy(include_sasa=True)
    eb = ser.total_energy(include_sasa=True)
    assert ea == eb
    for eng in (ovl, ser):
        v = np.concatenate([E.maxwell_boltzmann(s, 300.0, S.Xoshiro256pp(21 + k))
                            for k, s in enumerate(systems)])
        eng.set_velocities(v)
        eng.dynamics_setup([300.0, 330.0], dt_fs=2.0, friction_ps_inv=2.0, seed=5,
                           constrain_h_bonds=True, include_sasa=True, sasa_gamma_scale=0.25)
    ovl.use_graphs(True)
    for rep in range(3):
        ovl.run(4)                      # captured steps, two streams
        ser.run(4)                      # profiled: direct launches, one stream
        assert np.array_equal(ovl.positions(), ser.positions()), rep
        assert np.array_equal(ovl.velocities(), ser.velocities()), rep


def test_energies_leave_dynamics_forces(E, relaxed):
    """og_energies between og_run chunks (as REMD does for its swaps, with
    other flags than the dynamics) must not change the forces the next step
    starts from: trajectories bitwise equal to a run without the calls, and
    the energies equal to those of a force-and-energy evaluation."""
    c1_system = relaxed["c1"]
    c2_system = relaxed["c2"]
    systems = [c1_system, c2_system]

    def make():
        eng = E.Engine(systems)
        eng.set_positions(np.concatenate([jittered(s, 60 + k) for k, s in enumerate(systems)]))
        v = np.concatenate([E.maxwell_boltzmann(s, 300.0, S.Xoshiro256pp(80 + k))
                            for k, s in enumerate(systems)])
        eng.set_velocities(v)
        eng.dynamics_setup([300.0, 320.0], dt_fs=2.0, friction_ps_inv=2.0, seed=3,
                           constrain_h_bonds=True, include_sasa=True, sasa_gamma_scale=0.25)
        return eng
    a, b = make(), make()
    for rep in range(3):
        a.run(3)
        b.run(3)
        ea = a.total_energy(mode_a=True, include_sasa=True)
        assert np.array_equal(a.positions(), b.positions()), rep
        assert np.array_equal(a.velocities(), b.velocities()), rep
    ref = E.Engine(systems)
    _, e_ref = ref.total_force(a.positions(), mode_a=True, include_sasa=True, energies=True)
    ea = a.total_energy(mode_a=True, include_sasa=True)
    for x, y in zip(ea, e_ref):
        for k in y:
            assert x[k] == y[k], k


def test_sasa_energy_matches_oracle(E, c1_system, c2_system):
    """PowerSasa energy sum_i gamma_i A_i (powersasa_energy, no gamma scale) on
    the GPU against the oracle's per-atom accessible areas, generic geometry,
    uneven batch of two chains."""
    systems = [c1_system, c2_system]
    pos = [jittered(s, 5) for s in systems]
    eng = E.Engine(systems)
    e = eng.total_energy(np.concatenate(pos), include_sasa=True)
    for k, s in enumerate(systems):
        total, per_atom = O.Oracle(s).sasa_area(pos[k])
        ref = float((s.sasa_gamma * per_atom).sum())
        print("system %d SASA energy gpu %.10f oracle %.10f kJ/mol (area %.3f A^2)"
              % (k, e[k]["sasa"], ref, total))
        assert abs(e[k]["sasa"] - ref) <= 1e-9 * max(1.0, abs(ref))
    # forces are unchanged by the energy request
    f0 = eng.total_force(np.concatenate(pos), include_sasa=True, sasa_gamma_scale=0.25)
    f1, _ = eng.total_force(np.concatenate(pos), include_sasa=True, sasa_gamma_scale=0.25, energies=True)
    assert np.array_equal(f0, f1)


@pytest.mark.parametrize("graphs", [False, True])
def test_noise_ahead_matches_single_steps(E, relaxed, graphs):
    """Without SASA each step's O-step normals are formed during the force
    evaluation of the step before (one og_run call of several steps); a run of
    single-step calls forms every step's at the start of its call, the form
    that test_baoab_step_matches_oracle pins to the oracle's Philox normals.
    Both bitwise equal over steps that cover both halves of the buffer."""
    c1_system = relaxed["c1"]
    c3_system = relaxed["c3"]
    systems = [c1_system, c3_system]

    def make():
        eng = E.Engine(systems)
        v = np.concatenate([E.maxwell_boltzmann(s, 300.0, S.Xoshiro256pp(31 + k))
                            for k, s in enumerate(systems)])
        eng.set_velocities(v)
        eng.dynamics_setup([300.0, 325.0], dt_fs=2.0, friction_ps_inv=2.0, seed=99,
                           constrain_h_bonds=True)
        return eng
    many, single = make(), make()
    many.use_graphs(graphs)
    for rep in range(2):
        many.run(5)
        for _ in range(5):
            single.run(1)
        assert np.array_equal(many.positions(), single.positions()), rep
        assert np.array_equal(many.velocities(), single.velocities()), rep


def test_noise_paths_agree(E, relaxed):
    """Dynamics with SASA switched on at zero surface tension (no SASA force;
    the cluster step forms its own normals) against dynamics without SASA (the
    normals read from the precomputed buffer): the same trajectory up to the
    rounding of the O-step's multiply-add."""
    c1_system = relaxed["c1"]
    c3_system = relaxed["c3"]
    systems = [c1_system, c3_system]
    engs = []
    for sasa in (False, True):
        eng = E.Engine(systems)
        v = np.concatenate([E.maxwell_boltzmann(s, 300.0, S.Xoshiro256pp(41 + k))
                            for k, s in enumerate(systems)])
        eng.set_velocities(v)
        eng.dynamics_setup([300.0, 325.0], dt_fs=2.0, friction_ps_inv=2.0, seed=123,
                           constrain_h_bonds=True, include_sasa=sasa, sasa_gamma_scale=0.0)
        eng.run(6)
        engs.append(eng)
    a, b = engs
    assert np.abs(a.positions() - b.positions()).max() < 1e-9
    assert np.abs(a.velocities() - b.velocities()).max() < 1e-9


def test_gather_forms_agree(E, c2_system):
    """The bonded gather sums a thread per atom on batches of 20 000 atoms and
    more and from an LDS copy of each block's slots below that; both add in
    slot order, so a system's forces in a 16-copy batch (21 008 atoms) equal
    its forces alone bit for bit (tile pairs never straddle systems and every
    sum is fixed point)."""
    pos = [jittered(c2_system, 60 + k) for k in range(16)]
    big = E.Engine([c2_system] * 16)
    assert big.n_atoms >= 20000
    f = big.total_force(np.concatenate(pos))
    for k in (0, 7, 15):
        one = E.Engine([c2_system])
        assert np.array_equal(big.split(f)[k], one.total_force(pos[k])), k


@pytest.mark.parametrize("mode_a", [False, True])
def test_split_tile_pairs_small_system(E, c3_system, mode_a):
    """Systems of at most four tiles (here 171 atoms, three tiles) run each
    off-diagonal tile pair as two 32-step work items: forces, Born radii and
    energies against the oracle, at generic geometry, alone and in a batch
    beside a larger (unsplit) system."""
    s = c3_system
    pos = jittered(s, 70)
    orc = O.Oracle(s)
    cut = O.MODE_A if mode_a else O.MODE_B
    fref = orc.total_force(pos, **cut)
    eng = E.Engine([s])
    assert eng.info()["tile_pairs"] == 6
    f = eng.total_force(pos, mode_a=mode_a)
    err = rel_rms(f, fref)
    print("171 atoms mode_a=%s rel force RMS %.3e" % (mode_a, err))
    assert err < FORCE_RMS_TOL
    ref = orc.total_energy(pos, cut_nb=cut["cut_nb"], cut_descreen=cut["cut_descreen"])
    got = eng.total_energy(pos, mode_a=mode_a)[0]
    for k in ("lj", "coulomb", "gb_self", "gb_pair"):
        assert energy_close(got[k], ref[k]), (k, got[k], ref[k])
    big = S.seeded_chain(1, 20)                     # 300+ atoms: whole tile pairs
    both = E.Engine([big, s])
    fb = both.total_force(np.concatenate([big.pos, pos]), mode_a=mode_a)
    assert np.array_equal(both.split(fb)[1], f)


@pytest.mark.parametrize("mode_a", [False, True])
def test_diagonal_tile_far_exclusions(E, mode_a):
    """Exclusion and 1-4 pairs more than 32 atoms apart inside one 64-atom
    tile: the diagonal-tile force kernel evaluates such a pair on the lane of
    its higher atom, which must carry the pair's mask bit too (a C(i)-CD(i+1)
    1-3 pair across a large residue followed by Pro spans 33)."""
    s = S.build_system(["Trp", "Pro", "Trp", "Pro", "Trp", "Pro"])
    far = lambda p: [(int(a), int(b)) for a, b in np.asarray(p).reshape(-1, 2)
                     if a // 64 == b // 64 and b - a > 32]
    assert far(s.excl_pairs) and far(s.pair14)
    pos = jittered(s, 12)
    orc = O.Oracle(s)
    cut = O.MODE_A if mode_a else O.MODE_B
    fref = orc.total_force(pos, **cut)
    eng = E.Engine([s])
    f = eng.total_force(pos, mode_a=mode_a)
    err = rel_rms(f, fref)
    print("Trp-Pro x3 mode_a=%s rel force RMS %.3e, far pairs %s / %s"
          % (mode_a, err, far(s.excl_pairs), far(s.pair14)))
    assert err < FORCE_RMS_TOL
    ref = orc.total_energy(pos, cut_nb=cut["cut_nb"], cut_descreen=cut["cut_descreen"])
    got = eng.total_energy(pos, mode_a=mode_a)[0]
    for k in ("lj", "coulomb", "gb_pair", "total"):
        assert energy_close(got[k], ref[k]), (k, got[k], ref[k])


def test_cutoff_a_sets_lj_coulomb_cutoff(E, c2_system):
    """og_forces' cutoff_a (total_force_with_scratch's cutoff_a, forces.rs:82):
    under OG_MODE_A it is the LJ / reaction-field Coulomb cutoff, while the GB
    pair force stays at 10 A and the descreening sum at 20 A."""
    s = c2_system
    pos = jittered(s, 31)
    orc = O.Oracle(s)
    eng = E.Engine([s])
    for rc in (8.0, 12.5):
        fref = orc.total_force(pos, cut_nb=rc, cut_gb=10.0, cut_descreen=20.0)
        f = eng.total_force(pos, mode_a=True, cutoff_a=rc)
        assert rel_rms(f, fref) < FORCE_RMS_TOL, rc
        ref = orc.total_energy(pos, cut_nb=rc, cut_descreen=20.0)
        got = eng.total_energy(pos, mode_a=True, cutoff_a=rc)[0]
        for k in ("lj", "coulomb"):
            assert energy_close(got[k], ref[k]), (rc, k, got[k], ref[k])
    for bad in (0.0, -1.0, float("inf"), float("nan")):
        with pytest.raises(E.EngineError, match="cutoff_a"):
            eng.total_force(pos, mode_a=True, cutoff_a=bad)
    # without OG_MODE_A every pair counts and cutoff_a is not read
    assert np.array_equal(eng.total_force(pos, cutoff_a=float("nan")), eng.total_force(pos))


def test_fixed_point_range_is_reported(E, c3_system):
    """Two atoms 0.3 A apart give pair forces far beyond the int64 fixed-point
    range: the evaluation reports OG_ERR_NONFINITE with the system's range bit
    instead of returning wrapped sums, every kernel exits normally, and the
    engine evaluates normally again afterwards."""
    s = c3_system
    eng = E.Engine([s])
    f_ok = eng.total_force()
    pos = s.pos.copy()
    heavy = np.flatnonzero(s.mass > 2.0)
    a, b = int(heavy[0]), int(heavy[-1])
    pos[b] = pos[a] + np.array([0.3, 0.0, 0.0])
    with pytest.raises(E.EngineError, match="fixed-point range"):
        eng.total_force(pos)
    e = eng.total_energy(pos, allow_nonfinite=True)[0]
    assert np.isnan(e["lj"])
    assert np.array_equal(eng.total_force(s.pos), f_ok)


def test_sasa_capacity_stops_the_system(E, c1_system):
    """A packing where SASA atoms overlap more than 128 neighbours exceeds the
    arrangement's capacity: the evaluation reports it, and in dynamics that
    system stops at the step where it happened while the others run on."""
    s = c1_system
    pos = dense_packing(s.n_atoms, 9, density=0.16, min_sep=1.2)
    rad = s.sasa_radius
    d = np.linalg.norm(pos[:, None] - pos[None], axis=2)
    over = ((d < rad[:, None] + rad[None]) & (d > 0))[s.sasa_gamma != 0].sum(1)
    assert (over > 128).sum() >= 1
    single = E.Engine([s])
    with pytest.raises(E.EngineError, match="capacity"):
        single.total_force(pos, include_sasa=True, sasa_gamma_scale=0.25)

    def make():
        eng = E.Engine([s, s])
        eng.set_velocities(np.concatenate([E.maxwell_boltzmann(s, 300.0, S.Xoshiro256pp(3 + k))
                                           for k in range(2)]))
        eng.dynamics_setup([300.0, 300.0], dt_fs=1.0, friction_ps_inv=2.0, seed=8,
                           include_sasa=True, sasa_gamma_scale=0.25)
        eng.set_positions(np.concatenate([pos, s.pos]))
        return eng
    one, five = make(), make()
    one.run(1, allow_stopped=True)
    with pytest.raises(E.EngineError, match="stopped"):
        five.run(5)
    st = five.status()["diverged"]
    print("status bits after 5 steps:", st)
    assert st[0] & E.STATUS_SASA and st[1] == 0
    x1, x5 = one.split(one.positions()), five.split(five.positions())
    assert np.array_equal(x1[0], x5[0])             # stopped at the first step
    assert not np.array_equal(x1[1], x5[1])         # the other system ran on
    assert np.isfinite(x5[1]).all()


def test_swap_refuses_other_topology(E, c3_system):
    s = c3_system
    t = copy.deepcopy(s)
    t.charge = t.charge.copy()
    t.charge[0] += 0.01
    eng = E.Engine([s, t, s])
    eng.dynamics_setup([300.0, 310.0, 320.0], dt_fs=1.0)
    with pytest.raises(E.EngineError, match="topology"):
        eng.swap(0, 1)
    eng.swap(0, 2)


def test_remd_stopped_replica_is_skipped(E, c3_system):
    """remd.rs:382-389, :430: a replica whose velocities blow up stops, pairs
    with it are not attempted, and the other replicas run on."""
    import copy as cp
    from origami_amd.minimize import minimize
    e0 = E.Engine([c3_system])
    minimize(e0, max_steps=200, mode_a=False)
    s = cp.copy(c3_system)
    s.pos = e0.positions()
    temps = E.geometric_ladder(280.0, 395.0, 6)
    v = [E.maxwell_boltzmann(s, t, S.Xoshiro256pp(50 + k), remove_com=False)
         for k, t in enumerate(temps)]
    v[2] = v[2] * 1e150
    eng, summ = E.run_remd(s, temps, total_time_fs=400.0, swap_interval_fs=50.0, dt_fs=2.0,
                           friction_ps_inv=2.0, seed=4, constrain_h_bonds=True,
                           velocities=np.concatenate(v))
    print(summ)
    assert summ.diverged == [False, False, True, False, False, False]
    assert summ.swap_attempts[1] == 0 and summ.swap_attempts[2] == 0
    assert summ.swap_attempts[0] > 0 and summ.swap_attempts[3] > 0 and summ.swap_attempts[4] > 0
    xs = eng.split(eng.positions())
    for k in (0, 1, 3, 4, 5):
        assert np.isfinite(xs[k]).all()


@pytest.mark.parametrize("mode_a", [False, True])
def test_persistent_replica_steps_match_launches(E, relaxed, mode_a):
    """The persistent replica-step kernel (a block per system, many steps per
    launch) against captured steps and direct launches: positions,
    velocities, forces and per-step kinetic energies bit for bit, over
    several calls (chunk boundaries), an uneven batch and a REMD-style ladder."""
    c1, c3 = relaxed["c1"], relaxed["c3"]
    for systems in ([c3] * 8, [c1, c3, c3, c1]):
        def make(mode):
            eng = E.Engine(systems)
            v = np.concatenate([E.maxwell_boltzmann(s, 300.0, S.Xoshiro256pp(60 + k))
                                for k, s in enumerate(systems)])
            eng.set_velocities(v)
            eng.dynamics_setup([280.0 + 10.0 * k for k in range(len(systems))], dt_fs=2.0,
                               friction_ps_inv=2.0, seed=31, constrain_h_bonds=True, mode_a=mode_a)
            eng.use_graphs(mode)
            return eng
        per, graph, direct = make("persistent"), make(True), make(False)
        for n in (7, 1, 600):
            print("persistent vs captured vs direct: %d systems, %d steps, mode_a=%s"
                  % (len(systems), n, mode_a), flush=True)
            per.run(n)
            graph.run(n)
            direct.run(n)
            for other in (graph, direct):
                assert np.array_equal(per.positions(), other.positions()), n
                assert np.array_equal(per.velocities(), other.velocities()), n
                assert np.array_equal(per.forces(), other.forces()), n
        ke_p = per.run(5, record_ke=True)
        ke_d = direct.run(5, record_ke=True)
        assert np.array_equal(ke_p, ke_d)
        st = per.status()
        assert not st["diverged"].any() and st["shake_failures"].sum() == 0
