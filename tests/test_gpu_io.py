"""Trajectory frames on the device and the reference's own systems on the GPU.

  * og_frames_setup / og_frames_read: the positions after every `every`-th
    step, kept in a device ring (langevin.rs save_every, :205-220, :332-341),
    equal to the positions a run in chunks of `every` steps reaches, and a
    run that would overflow the ring is refused before any step.
  * The public PDB entries of the reference's io tests (1UAO chignolin, 1L2Y
    Trp-cage, 2F4K villin; README.md:406-410) through og_forces with
    origami's integrator cutoffs (10 A LJ / reaction-field Coulomb, 10 A GB,
    20 A descreening) against the FP64 oracle.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from origami_amd import pdb as P
from origami_amd import system as S

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FORCE_RMS_TOL = 1e-5


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="module")
def E():
    from origami_amd import engine
    return engine


@pytest.fixture(scope="module")
def fixtures():
    out = {}
    for name in ("1UAO_chignolin", "1L2Y_model1", "2F4K_villin_hp35"):
        res = P.read_pdb(os.path.join(GOLDEN, name + ".pdb"))
        out[name] = (res, P.system_from_residues(res))
    return out


@pytest.mark.parametrize("name", ["1UAO_chignolin", "1L2Y_model1", "2F4K_villin_hp35"])
def test_reference_systems_mode_a_parity(E, fixtures, name):
    res, s = fixtures[name]
    print("fixture %s: %d atoms" % (name, s.n_atoms), flush=True)
    orc = O.Oracle(s)
    eng = E.Engine([s])
    f = eng.total_force(mode_a=True)
    fref = orc.total_force(s.pos, **O.MODE_A)
    err = rel_rms(f, fref)
    print("%s Mode A force rel RMS %.3e" % (name, err))
    assert err < FORCE_RMS_TOL
    got = eng.total_energy(mode_a=True)[0]
    ref = orc.total_energy(s.pos, cut_nb=10.0, cut_descreen=20.0)
    for k in ("bond", "angle", "dihedral", "improper", "lj", "coulomb", "gb_self", "gb_pair"):
        assert abs(got[k] - ref[k]) <= 1e-9 * abs(ref[k]) + 1e-6, (name, k, got[k], ref[k])
    # with SASA (generic geometry: NMR / X-ray coordinates)
    fs = eng.total_force(mode_a=True, include_sasa=True, sasa_gamma_scale=1.0)
    fsref = fref + orc.sasa_forces(s.pos, 1.0)[0]
    assert rel_rms(fs, fsref) < FORCE_RMS_TOL


def test_frames_ring_matches_chunked_run(E, fixtures):
    _, s = fixtures["1UAO_chignolin"]
    systems = [s, s]

    def make():
        eng = E.Engine(systems)
        eng.set_velocities(np.concatenate([E.maxwell_boltzmann(s, 300.0, S.Xoshiro256pp(9 + k))
                                           for k in range(2)]))
        eng.dynamics_setup([300.0, 330.0], dt_fs=2.0, friction_ps_inv=2.0, seed=17,
                           constrain_h_bonds=True, mode_a=True)
        return eng
    ring, chunks = make(), make()
    x0 = ring.positions()
    ring.frames_setup(5, 5)
    ring.run(20)                         # frames at steps 5, 10, 15, 20
    steps, xs = ring.read_frames()
    assert steps == [0, 5, 10, 15, 20]
    assert np.array_equal(xs[0], x0)
    for k in range(1, 5):
        chunks.run(5)
        assert np.array_equal(xs[k], chunks.positions()), k
    assert np.array_equal(xs[4], ring.positions())
    assert ring.read_frames()[0] == []
    # 6 frames would exceed a ring of 5: refused before any step
    x = ring.positions()
    with pytest.raises(E.EngineError, match="exceed"):
        ring.run(30)
    assert np.array_equal(ring.positions(), x)
    ring.run(25)
    assert ring.read_frames()[0] == [25, 30, 35, 40, 45]


def test_langevin_frames_to_pdb_trajectory(E, fixtures):
    res, s = fixtures["1L2Y_model1"]
    eng, summ = E.run_langevin(s, 60, dt_fs=2.0, temperature_k=300.0, friction_ps_inv=2.0, seed=5,
                               constrain_h_bonds=True, mode_a=True, save_every=10, frame_chunk=2)
    assert not summ.diverged
    assert [st for st, _ in summ.frames] == [0, 10, 20, 30, 40, 50, 60]
    assert np.array_equal(summ.frames[-1][1], eng.positions())
    text = P.write_pdb_trajectory(res, [x for _, x in summ.frames], "1L2Y 300 K")
    back = P.read_pdb_trajectory(text)
    assert len(back) == 7
    assert P.rmsd_ca(res, back[0]) < 1e-3            # step 0 is the NMR model (8.3f)
    drift = [P.rmsd_ca(res, f) for f in back]
    print("Calpha RMSD from the NMR model every 20 fs:", ["%.3f" % d for d in drift])
    assert 0.0 < drift[-1] < 2.0
