"""Co-translational chain growth on the GPU engine.

Follows origami's run_cotranslate (dynamics/src/cotranslate.rs:174-310): a
ribosome emits residues on a schedule; after each emission the residue is
appended with the NeRF builder from the chain's current coordinates, the
grown chain is relaxed by L-BFGS (200 steps, |g|inf < 10, dE < 0.01 -> 0.1,
0.1 Å steps), the new atoms get Maxwell-Boltzmann velocities from the
driver's xoshiro256++ stream (initialise_velocities_for_new_atoms,
langevin.rs:417-447), and BAOAB Langevin runs until the next emission (a
trailing slice of tail_steps after the last) with origami's integrator
cutoffs, no SHAKE (run_slice, cotranslate.rs:312-389) and, optionally, the
cylindrical exit tunnel as an external force on every step.

Each slice runs on a device engine built for the grown chain (origami builds
a fresh ForceScratch per slice as well, cotranslate.rs:330); the relaxation
is the device L-BFGS, the slice the engine's BAOAB with its frame ring.
The O-step noise is the engine's counter-based Philox stream, keyed per
slice by (seed, residue count) so that no two slices draw the same normals;
trajectories therefore match origami's statistically, not draw for draw.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import engine as E
from . import system as S
from .minimize import minimize_device


class UniformRibosome:
    """Residue i emitted at i x interval_fs (cotranslate.rs:60-84)."""

    def __init__(self, sequence: Sequence[str], interval_fs: float):
        self.sequence = list(sequence)
        self.interval_fs = max(float(interval_fs), 0.0)

    def emission_time_fs(self, i: int) -> float:
        return i * self.interval_fs


@dataclass
class CylindricalTunnel:
    """Soft cylindrical confinement (cotranslate.rs:105-129 defaults)."""
    axis_origin: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    axis_direction: Tuple[float, float, float] = (0.0, 0.0, 1.0)
    radius_a: float = 12.0
    length_a: float = 80.0
    k_confine: float = 50.0


@dataclass
class CotranslateFrame:
    time_fs: float
    residue_count: int
    positions: np.ndarray            # [N, 3] of the chain at that time
    instantaneous_temperature_k: float


@dataclass
class CotranslateResult:
    residues: List[S.Residue]        # final chain (coordinates updated)
    frames: List[CotranslateFrame] = field(default_factory=list)
    minimise_steps: List[int] = field(default_factory=list)


def _velocities_for_new_atoms(system: S.ResolvedSystem, old: np.ndarray, temperature_k: float,
                              rng: S.Xoshiro256pp) -> np.ndarray:
    """initialise_velocities_for_new_atoms (langevin.rs:417-447): existing
    atoms keep theirs, new ones draw x, y, z normals in atom order."""
    n_old = len(old)
    kbt = S.BOLTZMANN_KJ * temperature_k
    v = np.zeros((system.n_atoms, 3))
    v[:n_old] = old
    for i in range(n_old, system.n_atoms):
        sigma = math.sqrt(kbt * S.ACCEL_FACTOR / system.mass[i])
        v[i, 0] = sigma * rng.gaussian()
        v[i, 1] = sigma * rng.gaussian()
        v[i, 2] = sigma * rng.gaussian()
    return v


def _set_positions(residues: List[S.Residue], pos: np.ndarray):
    k = 0
    for r in residues:
        for j in range(len(r.pos)):
            r.pos[j] = tuple(float(c) for c in pos[k])
            k += 1


def run_cotranslate(ribosome: UniformRibosome, dt_fs: float = 1.0, temperature_k: float = 310.0,
                    friction_ps_inv: float = 2.0, seed: int = 0, tail_steps: int = 0,
                    tunnel: Optional[CylindricalTunnel] = None, save_every: int = 0,
                    include_sasa: bool = False, sasa_gamma_scale: float = 1.0) -> CotranslateResult:
    seq = ribosome.sequence
    out = CotranslateResult(residues=[])
    if not seq:
        return out
    residues = out.residues
    velocities = np.zeros((0, 3))
    rng = S.Xoshiro256pp(seed)
    clock = 0.0
    for i, aa in enumerate(seq):
        clock = max(clock, ribosome.emission_time_fs(i))
        S.append_residue(residues, aa)
        system = S.resolve(residues)
        eng = E.Engine([system])
        res = minimize_device(eng, max_steps=200, gradient_tol=10.0, energy_tol=0.1, max_step_a=0.1,
                              mode_a=True)[0]
        out.minimise_steps.append(res.steps)
        x = eng.positions()
        _set_positions(residues, x)
        system.pos = x
        velocities = _velocities_for_new_atoms(system, velocities, temperature_k, rng)
        if i + 1 < len(seq):
            target = ribosome.emission_time_fs(i + 1)
        else:
            target = clock + tail_steps * dt_fs
        steps = int(round(max(target - clock, 0.0) / dt_fs))
        if steps == 0:
            ke = 0.5 * float((system.mass[:, None] * velocities ** 2).sum()) / S.ACCEL_FACTOR
            out.frames.append(CotranslateFrame(clock, i + 1, x.copy(),
                                               E.kinetic_temperature(ke, system.n_atoms)))
            continue
        eng.set_velocities(velocities)
        eng.set_tunnel(tunnel)
        slice_seed = (seed * 0x9E3779B97F4A7C15 + (i + 1)) & ((1 << 64) - 1)
        eng.dynamics_setup([temperature_k], dt_fs=dt_fs, friction_ps_inv=friction_ps_inv,
                           seed=slice_seed, constrain_h_bonds=False, mode_a=True,
                           include_sasa=include_sasa, sasa_gamma_scale=sasa_gamma_scale)
        kes = []
        if save_every > 0:
            chunk = 256
            eng.frames_setup(chunk + 1, save_every)
            eng.read_frames()                      # the set-up frame: origami emits none here
            done = 0
            while done < steps:
                n = min(steps - done, chunk * save_every)
                kes.append(eng.run(n, record_ke=True))
                done += n
                st, xs = eng.read_frames()
                ke_all = np.concatenate(kes)[:, 0]
                for s_, xf in zip(st, xs):
                    out.frames.append(CotranslateFrame(
                        clock + s_ * dt_fs, i + 1, xf,
                        E.kinetic_temperature(float(ke_all[s_ - 1]), system.n_atoms)))
        else:
            eng.run(steps)
        clock += steps * dt_fs
        x = eng.positions()
        velocities = eng.velocities()
        _set_positions(residues, x)
    return out
