"""SASA forces and areas on the 157 072-atom batch, pinned bit for bit.

The arrangement kernel's shortcuts (skip, cover and burial tests, the code
layout of its loops) only decide which work is left out, so they must leave
the emitted arcs and vertices, and therefore the SASA forces and areas,
exactly as they are.  This test fixes the SHA-256 of

  * the SASA forces: og_forces with OG_SASA minus og_forces without it
    (both are int64 fixed-point sums converted to FP64 at the end, and the
    pair and bonded parts are the same integers in both calls, so the
    difference is the SASA part exactly), FP64 [N, 3] bytes;
  * the per-system SASA energies (sum of gamma_i A_i, og_energies with
    OG_SASA), FP64 [64] bytes;

on the 64 seeded chains (10-300 residues) with seeded 0.02 Å jitter, in
both arrangement pass forms.  The digests were taken from the arrangement
kernel before its loop and walk restructuring; any change to the emitted
arrangement changes them.
"""
import hashlib
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SASA_FORCE_SHA256 = "52d648634e785ef297b08a149f920f00377f73cfdc7e9cfef532d4478e1ea1cf"
SASA_ENERGY_SHA256 = "855dde84ae82951a81f2aa17bd4ffb70a1257a946efbb6824928eb41166cb6fa"


@pytest.fixture(scope="module")
def batch():
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    from origami_amd import engine as E
    systems, _ = bench.build_config4()
    rng = np.random.default_rng(44)
    x = np.concatenate([s.pos + rng.normal(scale=0.02, size=s.pos.shape) for s in systems])
    return E, systems, x


@pytest.mark.parametrize("passes, blocks_per_cu", [(0, 4), (2, 6)])
def test_sasa_forces_and_areas_bit_identical(batch, passes, blocks_per_cu):
    E, systems, x = batch
    eng = E.Engine(systems)
    assert eng.n_atoms == 157072
    eng.set_launch(sasa_blocks=blocks_per_cu * 256, sasa_passes=passes)
    f_sasa = eng.total_force(x, include_sasa=True, sasa_gamma_scale=0.25) - eng.total_force(x)
    e_sasa = np.array([e["sasa"] for e in eng.total_energy(x, include_sasa=True)])
    print("SASA |F| sum %.6f, energy sum %.6f" % (np.abs(f_sasa).sum(), e_sasa.sum()))
    assert hashlib.sha256(f_sasa.tobytes()).hexdigest() == SASA_FORCE_SHA256
    assert hashlib.sha256(e_sasa.tobytes()).hexdigest() == SASA_ENERGY_SHA256
