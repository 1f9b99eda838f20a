"""og_run_noalloc (per-step kinetic energies in the library's static scratch
array, no device allocation) against og_run on the same engine state: two
calls in a row, the second with a smaller input, every output bitwise equal."""
import copy

import numpy as np
import pytest

from origami_amd import system as S

pytestmark = pytest.mark.gpu

NOALLOC_KE_MAX = 65536        # OG_NOALLOC_KE_MAX, include/origami_gpu.h


@pytest.fixture(scope="module")
def E():
    from origami_amd import engine
    return engine


@pytest.fixture(scope="module")
def chains(E):
    """Two relaxed seeded chains (20 and 10 residues)."""
    from origami_amd.minimize import minimize
    out = []
    for seed, n in ((1, 20), (3, 10)):
        s = S.seeded_chain(seed, n)
        eng = E.Engine([s])
        minimize(eng, max_steps=200, mode_a=False)
        r = copy.copy(s)
        r.pos = eng.positions()
        out.append(r)
    return out


def make(E, systems, sasa):
    eng = E.Engine(systems)
    eng.set_velocities(np.concatenate([E.maxwell_boltzmann(s, 300.0, S.Xoshiro256pp(70 + k))
                                       for k, s in enumerate(systems)]))
    eng.dynamics_setup([300.0 + 15.0 * k for k in range(len(systems))], dt_fs=2.0,
                       friction_ps_inv=2.0, seed=11, constrain_h_bonds=True, include_sasa=sasa,
                       sasa_gamma_scale=0.25)
    return eng


@pytest.mark.parametrize("sasa", [False, True])
def test_noalloc_run_matches_run(E, chains, sasa):
    systems = chains + [chains[1]]
    ref, na = make(E, systems, sasa), make(E, systems, sasa)
    for n_steps in (6, 2):              # the second call with a smaller input
        print("og_run_noalloc vs og_run: %d systems, %d steps, sasa=%s" % (len(systems), n_steps, sasa),
              flush=True)
        ke_ref = ref.run(n_steps, record_ke=True)
        ke_na = na.run(n_steps, record_ke=True, noalloc=True)
        assert ke_na.shape == (n_steps, len(systems))
        assert np.array_equal(ke_na, ke_ref)
        assert (ke_na > 0).all()
        assert np.array_equal(na.positions(), ref.positions())
        assert np.array_equal(na.velocities(), ref.velocities())
        assert np.array_equal(na.forces(), ref.forces())
    na.run(3, noalloc=True)             # without kinetic energies
    ref.run(3)
    assert np.array_equal(na.positions(), ref.positions())


def test_noalloc_rejects_oversized_request(E, chains):
    eng = make(E, [chains[1]] * 2, False)
    x0 = eng.positions()
    with pytest.raises(E.EngineError, match="exceeds"):
        eng.run(NOALLOC_KE_MAX // 2 + 1, record_ke=True, noalloc=True)
    assert np.array_equal(eng.positions(), x0)        # rejected before any step
