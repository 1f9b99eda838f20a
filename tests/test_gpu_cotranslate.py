"""Co-translational growth on the GPU engine (dynamics/src/cotranslate.rs)
and the external exit tunnel in the engine's dynamics."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from origami_amd import cotranslate as CT
from origami_amd import system as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def E():
    from origami_amd import engine
    return engine


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_uniform_ribosome_emission_times():
    # cotranslate.rs:404-416
    r = CT.UniformRibosome(["Ala", "Gly", "Ser"], 500.0)
    assert [r.emission_time_fs(i) for i in range(3)] == [0.0, 500.0, 1000.0]


def test_tunnel_forces_in_dynamics(E, c1_system):
    """The first step's forces (og_dynamics_setup) carry the tunnel: the
    oracle's total force (origami's integrator cutoffs) plus its tunnel force,
    on a chain lying partly outside a narrow tunnel."""
    s = c1_system
    tun = CT.CylindricalTunnel(axis_origin=(-5.0, 0.0, 0.0), axis_direction=(1.0, 0.0, 0.0),
                               radius_a=3.0, length_a=40.0, k_confine=25.0)
    eng = E.Engine([s])
    eng.set_tunnel(tun)
    eng.dynamics_setup([300.0], dt_fs=1.0, friction_ps_inv=2.0, seed=1, mode_a=True)
    f = eng.forces()
    orc = O.Oracle(s)
    ft = O.tunnel_force(s.pos, tun.axis_origin, tun.axis_direction, tun.radius_a, tun.length_a,
                        tun.k_confine)
    assert np.abs(ft).max() > 1.0                   # the tunnel acts on this geometry
    fref = orc.total_force(s.pos, **O.MODE_A) + ft
    assert rel_rms(f, fref) < 1e-5
    # og_forces itself is origami's total_force (no external potential)
    assert rel_rms(eng.total_force(mode_a=True), orc.total_force(s.pos, **O.MODE_A)) < 1e-5
    with pytest.raises(E.EngineError, match="unit vector"):
        eng.set_tunnel(CT.CylindricalTunnel(axis_direction=(0.0, 0.0, 2.0)))


def test_cotranslate_ala3_grows_to_three_residues():
    # cotranslate.rs:468-500
    r = CT.UniformRibosome(["Ala", "Ala", "Ala"], 50.0)
    res = CT.run_cotranslate(r, dt_fs=1.0, temperature_k=310.0, friction_ps_inv=2.0, seed=1,
                             tail_steps=50, save_every=10)
    assert len(res.residues) == 3
    assert len(res.frames) > 5
    assert all(f.residue_count <= 3 for f in res.frames)
    assert all(np.isfinite(p).all() for r_ in res.residues for p in r_.pos)


def test_cotranslate_chain_grows_through_the_tunnel():
    seq = [S.FROM_ONE_LETTER[c] for c in "GYDPETGTWG"]
    tun = CT.CylindricalTunnel(axis_origin=(-10.0, 0.0, 0.0), axis_direction=(1.0, 0.0, 0.0),
                               radius_a=8.0, length_a=30.0, k_confine=50.0)
    r = CT.UniformRibosome(seq, 100.0)
    res = CT.run_cotranslate(r, dt_fs=1.0, temperature_k=310.0, friction_ps_inv=2.0, seed=7,
                             tail_steps=200, tunnel=tun, save_every=20)
    counts = [f.residue_count for f in res.frames]
    times = [f.time_fs for f in res.frames]
    print("frames %d, residue counts %s, minimiser steps %s" % (len(res.frames), sorted(set(counts)),
                                                               res.minimise_steps))
    assert counts == sorted(counts) and counts[0] == 1 and counts[-1] == 10
    assert all(b > a for a, b in zip(times, times[1:]))
    assert abs(times[-1] - (900.0 + 200.0)) < 1e-9
    final = S.resolve(res.residues)
    assert final.n_atoms == S.build_system(seq).n_atoms
    x = final.pos
    assert np.isfinite(x).all()
    # atoms within the tunnel's axial range stay near its radius
    along = x[:, 0] - tun.axis_origin[0]
    inside = (along >= 0.0) & (along <= tun.length_a)
    radial = np.linalg.norm(x[:, 1:], axis=1)
    assert inside.any() and radial[inside].max() < tun.radius_a + 2.0
    assert all(np.isfinite(f.instantaneous_temperature_k) for f in res.frames)
