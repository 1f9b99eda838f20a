"""GPU: the learned Born-radius surrogate against its CPU oracle and against
plain GB (the arbiter), through the C ABI.

Tolerances: descriptors are exact (integer counts, exact bin edges);
forward radii within 1e-3 relative and gradients within 2e-2 relative L2 of
the oracle's bfloat16-emulating FP64 restatement (the GPU accumulates in FP32
in its own order, so a bfloat16 rounding of an intermediate can land on the
other side of a tie); trained radii within 2 % of plain GB (the fallback
threshold) and surrogate forces within 1 % relative RMS of plain GB.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import surrogate as SO
from origami_amd import system as S

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="module")
def E():
    from origami_amd import engine
    return engine


@pytest.fixture(scope="module")
def pair(E, c1_system, c2_system):
    """Two systems in one engine (uneven batch)."""
    eng = E.Engine([c1_system, c2_system])
    eng.surrogate_setup(200000, seed=7)
    return eng


def jittered(s, seed, scale=0.3):
    return s.pos + np.random.default_rng(seed).normal(scale=scale, size=s.pos.shape)


def test_descriptors_exact(E, pair, c1_system, c2_system):
    systems = [c1_system, c2_system]
    pos = np.concatenate([jittered(s, 10 + k) for k, s in enumerate(systems)])
    x = pair.surrogate_descriptors(pos)
    mass = np.concatenate([s.mass for s in systems])
    rho = np.concatenate([s.gb_rho for s in systems])
    ref = SO.descriptors(pos, mass, rho, pair.offsets)
    print("descriptor rows", x.shape, "nonzero", int((ref != 0).sum()))
    assert np.array_equal(x.astype(np.float64), ref)


def test_forward_matches_oracle(E, pair, c1_system, c2_system):
    p = SO.init_params(3)
    pair.surrogate_params(p.astype(np.float32))
    got = pair.surrogate_params()
    assert np.array_equal(got, p.astype(np.float32))
    pos = np.concatenate([c1_system.pos, c2_system.pos])
    r = pair.surrogate_radii(pos)
    x = pair.surrogate_descriptors(pos)
    y, _ = SO.forward(p, x.astype(np.float64))
    ref = SO.radius(y)
    err = np.abs(r - ref) / np.abs(ref)
    print("forward max rel err %.3e" % err.max())
    assert err.max() < 1e-3


def test_gradient_matches_oracle(E, pair, c2_system):
    p = SO.init_params(4)
    pair.surrogate_params(p.astype(np.float32))
    x = pair.surrogate_descriptors()[:512].astype(np.float64)
    rng = np.random.default_rng(1)
    t = rng.uniform(1.1, 1.8, size=len(x))
    g, loss = pair.surrogate_gradient(x.astype(np.float32), t)
    gref, lref = SO.gradient(p, x, t)
    print("loss gpu %.6f oracle %.6f" % (loss, lref))
    assert abs(loss - lref) <= 1e-3 * lref
    blocks = {"W1": (SO.OFF_W1, SO.OFF_B1), "b1": (SO.OFF_B1, SO.OFF_W2), "W2": (SO.OFF_W2, SO.OFF_B2),
              "b2": (SO.OFF_B2, SO.OFF_W3), "w3": (SO.OFF_W3, SO.OFF_B3), "b3": (SO.OFF_B3, SO.P)}
    for k, (a, b) in blocks.items():
        e = np.linalg.norm(g[a:b] - gref[a:b]) / np.linalg.norm(gref[a:b])
        print("grad %-3s rel L2 err %.3e" % (k, e))
        assert e < 2e-2, k


def test_train_fallback_and_forces(E, c1_system, c2_system):
    """Train on plain-GB radii over perturbed conformations, then use it."""
    systems = [c1_system, c2_system]
    eng = E.Engine(systems)
    eng.surrogate_setup(100000, seed=11)
    base = np.concatenate([s.pos for s in systems])
    rows = 0
    for k in range(12):
        eng.set_positions(base + np.random.default_rng(100 + k).normal(scale=0.3, size=base.shape))
        rows = eng.surrogate_collect()
    assert rows == 12 * eng.n_atoms
    pos = base + np.random.default_rng(7).normal(scale=0.3, size=base.shape)
    eng.total_force(pos)
    r_plain = eng.born_radii()[0]
    r0 = eng.surrogate_radii(pos)
    err0 = np.abs(r0 - r_plain) / r_plain
    loss = eng.surrogate_train(400, batch=2048, lr=2e-3, seed=5)
    r1 = eng.surrogate_radii(pos)
    err1 = np.abs(r1 - r_plain) / r_plain
    print("loss %.4e -> %.4e; radius max rel err untrained %.3e trained %.3e"
          % (loss[:10].mean(), loss[-10:].mean(), err0.max(), err1.max()))
    assert loss[-10:].mean() < 0.05 * loss[:10].mean()
    assert err1.max() < 0.02 and err1.max() < 0.1 * err0.max()
    # forces with surrogate radii vs plain GB and vs the FP64 oracle
    eng.surrogate_enable(check_every=1, tolerance=0.02)
    f_sur = eng.total_force(pos, surrogate=True)
    mode, dev = eng.surrogate_status()
    f_plain = eng.total_force(pos)
    fo = np.concatenate([O.Oracle(s).total_force(p) for s, p in zip(systems, eng.split(pos))])
    print("surrogate mode", mode, "deviation", dev, "force rel RMS vs plain %.3e vs oracle %.3e"
          % (rel_rms(f_sur, f_plain), rel_rms(f_sur, fo)))
    assert np.all(mode == 1) and np.all(dev < 0.02)
    assert rel_rms(f_sur, f_plain) < 0.01 and rel_rms(f_sur, fo) < 0.01
    # fallback: a network that predicts R = 1.5 everywhere is > 2 % off
    p = eng.surrogate_params()
    p[:] = 0.0
    eng.surrogate_params(p)
    eng.total_force(pos, surrogate=True)          # check call: flags both systems
    mode, dev = eng.surrogate_status()
    assert np.all(mode == 0) and np.all(dev > 0.02)
    f_fb = eng.total_force(pos, surrogate=True)   # now plain GB for both
    assert np.array_equal(f_fb, f_plain)


def test_surrogate_dynamics_runs(E, c1_system):
    from origami_amd.minimize import minimize
    eng = E.Engine([c1_system])
    minimize(eng, max_steps=60, mode_a=False)
    eng.surrogate_setup(20000, seed=1)
    eng.set_velocities(E.maxwell_boltzmann(c1_system, 310.0, S.Xoshiro256pp(1)))
    eng.dynamics_setup([310.0], dt_fs=0.5, friction_ps_inv=5.0, seed=2)
    for _ in range(5):
        eng.run(10)
        eng.surrogate_collect()
    eng.surrogate_train(400, batch=1024, lr=2e-3, seed=1)
    eng.surrogate_enable(check_every=10, tolerance=0.02)
    eng.dynamics_setup([310.0], dt_fs=0.5, friction_ps_inv=5.0, seed=2, surrogate=True)
    eng.run(50)
    assert np.isfinite(eng.positions()).all()
    mode, dev = eng.surrogate_status()
    print("after 50 surrogate steps: mode", mode, "deviation", dev)
    assert mode[0] == 1
