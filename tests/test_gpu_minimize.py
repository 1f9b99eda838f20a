"""Device-resident batched L-BFGS (og_minimize) against the host restatement
of origami's minimiser (origami_amd/minimize.py, lbfgs.rs:61-254 +
line_search.rs:52-101) on the same engine forces and FP64 energies."""
import copy

import numpy as np
import pytest

from origami_amd import system as S
from origami_amd.minimize import minimize, minimize_device

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from origami_amd import engine
    return engine


@pytest.mark.parametrize("mode_a", [False, True])
def test_device_iterates_match_host(E, c1_system, mode_a):
    """The iterate sequence: after k = 1, 2, 5, 10, 25 steps from the same
    start both minimisers stand at the same coordinates and report the same
    steps, energies and line-search failures."""
    s = c1_system
    worst = 0.0
    for k in (1, 2, 5, 10, 25):
        print("og_minimize vs host L-BFGS: %d steps, mode_a=%s" % (k, mode_a), flush=True)
        eh, ed = E.Engine([s]), E.Engine([s])
        rh = minimize(eh, max_steps=k, mode_a=mode_a)[0]
        rd = minimize_device(ed, max_steps=k, mode_a=mode_a)[0]
        d = float(np.abs(eh.positions() - ed.positions()).max())
        worst = max(worst, d)
        assert d < 1e-9, (k, d)
        assert rh.steps == rd.steps and rh.line_search_failures == rd.line_search_failures
        assert rh.converged == rd.converged
        assert abs(rh.initial_energy - rd.initial_energy) <= 1e-9 * abs(rh.initial_energy)
        assert abs(rh.final_energy - rd.final_energy) <= 1e-9 * abs(rh.final_energy) + 1e-9
        assert abs(rh.max_force - rd.max_force) <= 1e-9 * abs(rh.max_force)
    print("worst coordinate difference %.3e A" % worst)


def test_device_batch_matches_single_systems(E, c1_system, c3_system):
    """Systems minimise independently in one batch: each ends where it ends
    alone (a system that converges early stops while the others go on)."""
    systems = [c1_system, c3_system, c1_system]
    eb = E.Engine(systems)
    rb = minimize_device(eb, max_steps=60, mode_a=True)
    xb = eb.split(eb.positions())
    for k, s in enumerate(systems):
        e1 = E.Engine([s])
        r1 = minimize_device(e1, max_steps=60, mode_a=True)[0]
        assert rb[k].steps == r1.steps and rb[k].final_energy == r1.final_energy, k
        assert np.array_equal(xb[k], e1.positions()), k
    assert rb[0] == rb[2]


def test_device_trp_cage_acceptance(E):
    # dynamics/tests/trp_cage_minimise.rs:12-55 on the device minimiser
    s = S.build_system([S.FROM_ONE_LETTER[c] for c in "NLYIQWLKDGGPSSGRPPPS"])
    eng = E.Engine([s])
    res = minimize_device(eng, max_steps=500, gradient_tol=5.0, max_step_a=0.1, mode_a=True)[0]
    e = eng.total_energy(mode_a=True)[0]
    print(res, e["bond"])
    assert res.initial_energy > 40000.0
    assert res.final_energy < 1000.0 and res.steps < 500 and e["bond"] < 200.0
    assert abs(res.final_energy - e["total"]) <= 1e-9 * abs(e["total"]) + 1e-9


def test_device_minimiser_leaves_consistent_state(E, c2_system):
    """Forces after og_minimize are the forces at the final coordinates (the
    next dynamics step starts from them), and the energy it reports is the
    energy there."""
    eng = E.Engine([c2_system])
    res = minimize_device(eng, max_steps=30, mode_a=False)[0]
    f_after = eng.forces()
    f_fresh = eng.total_force()
    assert np.array_equal(f_after, f_fresh)
    assert res.final_energy < res.initial_energy
    with pytest.raises(E.EngineError, match="history"):
        minimize_device(eng, max_steps=5, history=0)
