"""Python host mirror of origami's force/dynamics interface over the C ABI.

The names follow origami's own entry points so parity tests read like
origami's tests:

  ``Engine``                       ForceScratch (energy/src/scratch.rs:105-262)
  ``Engine.total_force``           total_force_with_scratch (energy/src/forces.rs:78-128)
  ``Engine.total_energy``          total_energy_with_options (dynamics/src/energy_eval.rs:20-34)
  ``Engine.born_radii``            compute_born_radii_into_scratch (energy/src/gb.rs:203-342)
  ``run_langevin``                 dynamics/src/langevin.rs:128-367
  ``run_remd``                     dynamics/src/remd.rs:136-498

Everything computes on the GPU through ``liborigami_gpu.so``; there is no CPU
fallback.  Importing this module without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import system as S

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "liborigami_gpu.so")

MODE_A = 1
SASA = 2
ENERGY = 4
SURROGATE = 8
SASA_SR = 16

ENERGY_TERMS = ["bond", "angle", "dihedral", "improper", "lj", "coulomb", "gb_self", "gb_pair",
                "sasa", "kinetic"]
KERNEL_CLASSES = ["cluster", "bonded", "descreen", "radii", "force", "sasa", "back", "surrogate",
                  "other"]

SURROGATE_FEATURES = 161
SURROGATE_PARAMS = 41345

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class SystemDesc(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("pos", _dp), ("charge", _dp), ("rmin_half", _dp),
                ("epsilon", _dp), ("rmin_half_14", _dp), ("epsilon_14", _dp), ("gb_rho", _dp),
                ("gb_scale", _dp), ("sasa_radius", _dp), ("sasa_gamma", _dp), ("mass", _dp),
                ("n_excl", C.c_int32), ("excl", _ip), ("n_14", C.c_int32), ("pair14", _ip),
                ("n_bonds", C.c_int32), ("bonds", _ip), ("bond_params", _dp),
                ("n_angles", C.c_int32), ("angles", _ip), ("angle_params", _dp),
                ("n_dihedrals", C.c_int32), ("dihedrals", _ip), ("dihedral_params", _dp),
                ("n_impropers", C.c_int32), ("impropers", _ip), ("improper_params", _dp),
                ("n_constraints", C.c_int32), ("constraints", _ip), ("constraint_len", _dp)]


class DynamicsParams(C.Structure):
    _fields_ = [("dt_fs", C.c_double), ("friction_ps_inv", C.c_double), ("seed", C.c_uint64),
                ("constrain_h_bonds", C.c_int32), ("flags", C.c_uint32),
                ("sasa_gamma_scale", C.c_double), ("cutoff_a", C.c_double)]


# origami's DEFAULT_CUTOFF_A (energy/src/nonbonded.rs:37): the LJ / Coulomb
# cutoff its integrator passes; used here under mode_a
DEFAULT_CUTOFF_A = 10.0

# per-system status bits (og_get_status)
STATUS_NONFINITE = 1
STATUS_RANGE = 2
STATUS_SASA = 4


class EngineError(RuntimeError):
    pass


class Tunnel(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("direction", C.c_double * 3), ("radius_a", C.c_double),
                ("length_a", C.c_double), ("k_confine", C.c_double)]


class MinimizeParams(C.Structure):
    _fields_ = [("max_steps", C.c_int32), ("history", C.c_int32), ("gradient_tol", C.c_double),
                ("energy_tol", C.c_double), ("max_step_a", C.c_double), ("alpha0", C.c_double),
                ("contraction", C.c_double), ("c1", C.c_double), ("min_alpha", C.c_double),
                ("flags", C.c_uint32), ("cutoff_a", C.c_double), ("sasa_gamma_scale", C.c_double)]


class MinimizeResult(C.Structure):
    _fields_ = [("steps", C.c_int32), ("converged", C.c_int32),
                ("line_search_failures", C.c_int32), ("stopped", C.c_int32),
                ("initial_energy", C.c_double), ("final_energy", C.c_double),
                ("max_force", C.c_double)]


_lib = None


def lib():
    """Load the HIP engine library; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise EngineError("HIP engine library not built: %s (run python -m origami_amd.build)"
                              % _LIB_PATH)
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.og_last_error.restype = C.c_char_p
        L.og_engine_create.argtypes = [C.POINTER(SystemDesc), C.c_int32, C.POINTER(vp)]
        L.og_engine_destroy.argtypes = [vp]
        L.og_engine_destroy.restype = None
        L.og_engine_info.argtypes = [vp, C.POINTER(C.c_int64)]
        for nm in ("og_set_positions", "og_get_positions", "og_set_velocities",
                   "og_get_velocities", "og_get_forces"):
            getattr(L, nm).argtypes = [vp, _dp]
        L.og_forces.argtypes = [vp, _dp, _dp, C.c_uint32, C.c_double, C.c_double, _dp]
        L.og_energies.argtypes = [vp, _dp, C.c_uint32, C.c_double, _dp]
        L.og_born_radii.argtypes = [vp, _dp, _dp]
        L.og_dynamics_setup.argtypes = [vp, C.POINTER(DynamicsParams), _dp]
        L.og_run.argtypes = [vp, C.c_int32, _dp]
        L.og_run_noalloc.argtypes = [vp, C.c_int32, _dp]
        L.og_frames_setup.argtypes = [vp, C.c_int32, C.c_int32]
        L.og_strip_tiles.argtypes = [vp, C.c_int32, _ip]
        L.og_set_tunnel.argtypes = [vp, C.POINTER(Tunnel)]
        L.og_minimize.argtypes = [vp, C.POINTER(MinimizeParams), C.POINTER(MinimizeResult)]
        L.og_frames_read.argtypes = [vp, _dp, C.POINTER(C.c_uint32), C.c_int32, C.POINTER(C.c_int32)]
        L.og_set_temperature.argtypes = [vp, C.c_int32, C.c_double]
        L.og_swap_systems.argtypes = [vp, C.c_int32, C.c_int32]
        L.og_get_status.argtypes = [vp, _ip, _ip, _ip]
        L.og_profile_enable.argtypes = [vp, C.c_int32]
        L.og_profile_read.argtypes = [vp, _dp, C.POINTER(C.c_int64)]
        L.og_use_graphs.argtypes = [vp, C.c_int32]
        L.og_autotune.argtypes = [vp, C.c_uint32, _ip, _dp]
        L.og_set_launch.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
        _fp = C.POINTER(C.c_float)
        L.og_surrogate_setup.argtypes = [vp, C.c_int64, C.c_uint64]
        L.og_surrogate_collect.argtypes = [vp, C.c_uint32, C.POINTER(C.c_int64)]
        L.og_surrogate_train.argtypes = [vp, C.c_int32, C.c_int32, C.c_double, C.c_uint64, _fp]
        L.og_surrogate_enable.argtypes = [vp, C.c_int32, C.c_double]
        L.og_surrogate_radii.argtypes = [vp, _dp, _dp]
        L.og_surrogate_descriptors.argtypes = [vp, _dp, _fp]
        L.og_surrogate_params.argtypes = [vp, _fp, _fp]
        L.og_surrogate_gradient.argtypes = [vp, _fp, _dp, C.c_int32, _fp, _dp]
        L.og_surrogate_status.argtypes = [vp, _ip, _fp]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise EngineError("origami engine error %d: %s" % (rc, lib().og_last_error().decode()))


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Engine:
    """A device-resident batch of one or more systems (chains / replicas)."""

    def __init__(self, systems: Sequence[S.ResolvedSystem]):
        if isinstance(systems, S.ResolvedSystem):
            systems = [systems]
        L = lib()
        self.systems = list(systems)
        self.n_systems = len(self.systems)
        self.sizes = [s.n_atoms for s in self.systems]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n_atoms = int(self.offsets[-1])
        self._keep = []
        descs = (SystemDesc * self.n_systems)()

        def d(x, w=None):
            a = np.ascontiguousarray(x, dtype=np.float64)
            self._keep.append(a)
            return _d(a)

        def i(x):
            a = np.ascontiguousarray(x, dtype=np.int32)
            self._keep.append(a)
            return _i(a)

        for k, s in enumerate(self.systems):
            D = descs[k]
            D.n_atoms = s.n_atoms
            D.pos, D.charge = d(s.pos), d(s.charge)
            D.rmin_half, D.epsilon = d(s.rmin_half), d(s.epsilon)
            D.rmin_half_14, D.epsilon_14 = d(s.rmin_half_14), d(s.epsilon_14)
            D.gb_rho, D.gb_scale = d(s.gb_rho), d(s.gb_scale)
            D.sasa_radius, D.sasa_gamma, D.mass = d(s.sasa_radius), d(s.sasa_gamma), d(s.mass)
            D.n_excl, D.excl = len(s.excl_pairs), i(s.excl_pairs)
            D.n_14, D.pair14 = len(s.pair14), i(s.pair14)
            D.n_bonds, D.bonds, D.bond_params = len(s.bonds), i(s.bonds), d(s.bond_params)
            D.n_angles, D.angles, D.angle_params = len(s.angles), i(s.angles), d(s.angle_params)
            D.n_dihedrals, D.dihedrals, D.dihedral_params = (len(s.dihedrals), i(s.dihedrals),
                                                             d(s.dihedral_params))
            D.n_impropers, D.impropers, D.improper_params = (len(s.impropers), i(s.impropers),
                                                             d(s.improper_params))
            D.n_constraints, D.constraints, D.constraint_len = (len(s.constraints),
                                                                i(s.constraints),
                                                                d(s.constraint_len))
        h = C.c_void_p()
        _check(L.og_engine_create(descs, self.n_systems, C.byref(h)))
        self._h = h
        self._keep = []
        self.temperatures = np.zeros(self.n_systems)

    def close(self):
        if getattr(self, "_h", None):
            lib().og_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- info / state -------------------------------------------------------
    def info(self):
        out = (C.c_int64 * 8)()
        _check(lib().og_engine_info(self._h, out))
        keys = ["atoms", "systems", "tiles", "tile_pairs", "masked_tile_pairs", "clusters",
                "pairs", "sasa_work_atoms"]
        return dict(zip(keys, [int(x) for x in out]))

    def _state(self, getter):
        out = np.zeros((self.n_atoms, 3))
        _check(getter(self._h, _d(out)))
        return out

    def positions(self):
        return self._state(lib().og_get_positions)

    def velocities(self):
        return self._state(lib().og_get_velocities)

    def forces(self):
        return self._state(lib().og_get_forces)

    def set_positions(self, xyz):
        a = np.ascontiguousarray(xyz, dtype=np.float64).reshape(self.n_atoms, 3)
        _check(lib().og_set_positions(self._h, _d(a)))

    def set_velocities(self, v):
        a = np.ascontiguousarray(v, dtype=np.float64).reshape(self.n_atoms, 3)
        _check(lib().og_set_velocities(self._h, _d(a)))

    def split(self, arr):
        return [arr[self.offsets[k]:self.offsets[k + 1]] for k in range(self.n_systems)]

    # -- forces / energies ----------------------------------------------------
    def total_force(self, xyz=None, mode_a=False, include_sasa=False, sasa_gamma_scale=1.0,
                    energies=False, surrogate=False, cutoff_a=DEFAULT_CUTOFF_A):
        """mode_a: origami's integrator cutoffs with cutoff_a for LJ / Coulomb
        (total_force_with_scratch's cutoff_a, forces.rs:82); otherwise all pairs."""
        flags = ((MODE_A if mode_a else 0) | (SASA if include_sasa else 0)
                 | (SURROGATE if surrogate else 0))
        f = np.zeros((self.n_atoms, 3))
        e = np.zeros((self.n_systems, len(ENERGY_TERMS))) if energies else None
        x = None
        if xyz is not None:
            x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(self.n_atoms, 3)
        _check(lib().og_forces(self._h, _d(x) if x is not None else None, _d(f), flags,
                               float(cutoff_a), float(sasa_gamma_scale),
                               _d(e) if e is not None else None))
        if energies:
            return f, [dict(zip(ENERGY_TERMS, row)) for row in e]
        return f

    def total_energy(self, xyz=None, mode_a=False, surrogate=False, include_sasa=False,
                     cutoff_a=DEFAULT_CUTOFF_A, allow_nonfinite=False, sasa_method="powersasa"):
        """Per-system energy terms.  allow_nonfinite: a system whose sums left
        the fixed-point range gets NaN terms instead of an exception.
        include_sasa with sasa_method "powersasa" (analytical areas,
        powersasa_energy) or "shrake_rupley" (960-dot areas, sasa_energy: the
        term origami's REMD swaps use)."""
        if include_sasa and sasa_method not in ("powersasa", "shrake_rupley"):
            raise ValueError("sasa_method must be 'powersasa' or 'shrake_rupley'")
        sasa_flag = 0
        if include_sasa:
            sasa_flag = SASA if sasa_method == "powersasa" else SASA_SR
        e = np.zeros((self.n_systems, len(ENERGY_TERMS)))
        x = None
        if xyz is not None:
            x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(self.n_atoms, 3)
        rc = lib().og_energies(self._h, _d(x) if x is not None else None,
                               (MODE_A if mode_a else 0) | (SURROGATE if surrogate else 0)
                               | sasa_flag, float(cutoff_a), _d(e))
        if not (allow_nonfinite and rc == -4):
            _check(rc)
        out = []
        for row in e:
            dct = dict(zip(ENERGY_TERMS, row.tolist()))
            dct["gb"] = dct["gb_self"] + dct["gb_pair"]
            dct["total"] = (dct["bond"] + dct["angle"] + dct["dihedral"] + dct["improper"]
                            + dct["lj"] + dct["coulomb"] + dct["gb"] + dct["sasa"])
            out.append(dct)
        return out

    def born_radii(self):
        r = np.zeros(self.n_atoms)
        integ = np.zeros(self.n_atoms)
        _check(lib().og_born_radii(self._h, _d(r), _d(integ)))
        return r, integ

    # -- dynamics ---------------------------------------------------------------
    def dynamics_setup(self, temperatures, dt_fs=1.0, friction_ps_inv=1.0, seed=0,
                       constrain_h_bonds=False, mode_a=False, include_sasa=False,
                       sasa_gamma_scale=1.0, surrogate=False, cutoff_a=DEFAULT_CUTOFF_A):
        t = np.ascontiguousarray(np.broadcast_to(np.asarray(temperatures, dtype=np.float64),
                                                 (self.n_systems,)))
        self.temperatures = t.copy()
        p = DynamicsParams(dt_fs, friction_ps_inv, seed & ((1 << 64) - 1),
                           1 if constrain_h_bonds else 0,
                           (MODE_A if mode_a else 0) | (SASA if include_sasa else 0)
                           | (SURROGATE if surrogate else 0),
                           sasa_gamma_scale, cutoff_a)
        self.dt_fs = dt_fs
        _check(lib().og_dynamics_setup(self._h, C.byref(p), _d(t)))

    def run(self, n_steps, record_ke=False, allow_stopped=False, noalloc=False):
        """n_steps BAOAB steps of every system.  A system whose velocities turn
        non-finite (or whose forces leave the fixed-point range, or whose SASA
        exceeds capacity) stops at that step while the others keep running
        (remd.rs:382-389); that raises unless allow_stopped, and status()
        names the stopped systems.  noalloc: og_run_noalloc (no device
        allocation; per-step kinetic energies in the library's fixed scratch
        array of OG_NOALLOC_KE_MAX entries, direct launches)."""
        ke = np.zeros((n_steps, self.n_systems)) if record_ke and n_steps > 0 else None
        fn = lib().og_run_noalloc if noalloc else lib().og_run
        rc = fn(self._h, int(n_steps), _d(ke) if ke is not None else None)
        if rc == -4:
            if allow_stopped:
                return ke
            raise EngineError("diverged: " + lib().og_last_error().decode())
        _check(rc)
        return ke

    def use_graphs(self, on=True):
        """True: captured steps; False: direct launches; "persistent": the
        persistent replica-step kernel (one block per system for many steps;
        systems of up to 512 atoms, no SASA or surrogate); None: the engine's
        choice (direct launches, the default)."""
        mode = 2 if on is None else (3 if on == "persistent" else (1 if on else 0))
        _check(lib().og_use_graphs(self._h, mode))

    def set_launch(self, pair_blocks=0, sasa_blocks=0, sasa_passes=0):
        """Launch widths (0 keeps the current one) and the SASA arrangement's
        pass form (0 by LDS fit, 1 one pass of 128 caps, 2 passes 80 + 128)."""
        _check(lib().og_set_launch(self._h, int(pair_blocks), int(sasa_blocks), int(sasa_passes)))

    def autotune(self, mode_a=False, include_sasa=False, surrogate=False):
        ch = np.zeros(2, np.int32)
        ms = np.zeros(2)
        _check(lib().og_autotune(self._h, (MODE_A if mode_a else 0) | (SASA if include_sasa else 0)
                                 | (SURROGATE if surrogate else 0),
                                 _i(ch), _d(ms)))
        return {"pair_blocks": int(ch[0]), "sasa_blocks": int(ch[1]), "pair_ms": float(ms[0]),
                "sasa_ms": float(ms[1]), "strip_tiles": self.strip_tiles()}

    def strip_tiles(self, n=0):
        """The force pass's strip length (j tiles per wave on systems above
        four tiles); n > 0 sets it.  Returns the current value."""
        out = np.zeros(1, np.int32)
        _check(lib().og_strip_tiles(self._h, int(n), _i(out)))
        return int(out[0])

    def set_temperature(self, system, t):
        _check(lib().og_set_temperature(self._h, int(system), float(t)))
        self.temperatures[system] = t

    def swap(self, a, b):
        _check(lib().og_swap_systems(self._h, int(a), int(b)))

    def set_tunnel(self, tunnel=None):
        """External cylindrical exit tunnel for the dynamics (an object with
        axis_origin, axis_direction, radius_a, length_a, k_confine, e.g.
        cotranslate.CylindricalTunnel); None removes it.  Set before
        dynamics_setup."""
        if tunnel is None:
            _check(lib().og_set_tunnel(self._h, None))
            return
        t = Tunnel((C.c_double * 3)(*map(float, tunnel.axis_origin)),
                   (C.c_double * 3)(*map(float, tunnel.axis_direction)),
                   float(tunnel.radius_a), float(tunnel.length_a), float(tunnel.k_confine))
        _check(lib().og_set_tunnel(self._h, C.byref(t)))

    def frames_setup(self, capacity, every):
        """Keep every system's positions after every `every`-th step in a
        device ring of `capacity` frames (the current positions are the
        first frame); capacity = every = 0 switches frames off."""
        _check(lib().og_frames_setup(self._h, int(capacity), int(every)))
        self._frame_cap = int(capacity)

    def read_frames(self):
        """(steps, positions [n_frames, N, 3]) of the unread frames, oldest
        first."""
        cap = getattr(self, "_frame_cap", 0)
        if cap == 0:
            return [], np.zeros((0, self.n_atoms, 3))
        xyz = np.zeros((cap, self.n_atoms, 3))
        steps = (C.c_uint32 * cap)()
        n = C.c_int32()
        _check(lib().og_frames_read(self._h, _d(xyz), steps, cap, C.byref(n)))
        return [int(steps[k]) for k in range(n.value)], xyz[:n.value]

    def status(self):
        n = self.n_systems
        div, fails, cl = (np.zeros(n, np.int32) for _ in range(3))
        _check(lib().og_get_status(self._h, _i(div), _i(fails), _i(cl)))
        return {"diverged": div, "shake_failures": fails, "clamped": cl}

    # -- learned Born-radius surrogate -------------------------------------------
    def surrogate_setup(self, capacity_rows, seed=0):
        _check(lib().og_surrogate_setup(self._h, int(capacity_rows), int(seed) & ((1 << 64) - 1)))

    def surrogate_collect(self, mode_a=False):
        n = C.c_int64(0)
        _check(lib().og_surrogate_collect(self._h, MODE_A if mode_a else 0, C.byref(n)))
        return int(n.value)

    def surrogate_train(self, n_steps, batch=4096, lr=1e-3, seed=0):
        loss = np.zeros(max(int(n_steps), 1), np.float32)
        _check(lib().og_surrogate_train(self._h, int(n_steps), int(batch), float(lr),
                                        int(seed) & ((1 << 64) - 1),
                                        loss.ctypes.data_as(C.POINTER(C.c_float))))
        return loss[:int(n_steps)]

    def surrogate_enable(self, check_every=100, tolerance=0.02):
        _check(lib().og_surrogate_enable(self._h, int(check_every), float(tolerance)))

    def surrogate_radii(self, xyz=None):
        r = np.zeros(self.n_atoms)
        x = None if xyz is None else np.ascontiguousarray(xyz, dtype=np.float64).reshape(self.n_atoms, 3)
        _check(lib().og_surrogate_radii(self._h, _d(x) if x is not None else None, _d(r)))
        return r

    def surrogate_descriptors(self, xyz=None):
        out = np.zeros((self.n_atoms, SURROGATE_FEATURES), np.float32)
        x = None if xyz is None else np.ascontiguousarray(xyz, dtype=np.float64).reshape(self.n_atoms, 3)
        _check(lib().og_surrogate_descriptors(self._h, _d(x) if x is not None else None,
                                              out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def surrogate_params(self, new=None):
        got = np.zeros(SURROGATE_PARAMS, np.float32)
        st = None
        if new is not None:
            st = np.ascontiguousarray(new, dtype=np.float32)
            assert st.shape == (SURROGATE_PARAMS,)
        _check(lib().og_surrogate_params(self._h, got.ctypes.data_as(C.POINTER(C.c_float)),
                                         st.ctypes.data_as(C.POINTER(C.c_float)) if st is not None else None))
        return got

    def surrogate_gradient(self, x, r_target):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(r_target, dtype=np.float64)
        g = np.zeros(SURROGATE_PARAMS, np.float32)
        loss = C.c_double(0.0)
        _check(lib().og_surrogate_gradient(self._h, x.ctypes.data_as(C.POINTER(C.c_float)), _d(t),
                                           int(len(t)), g.ctypes.data_as(C.POINTER(C.c_float)),
                                           C.byref(loss)))
        return g, float(loss.value)

    def surrogate_status(self):
        mode = np.zeros(self.n_systems, np.int32)
        dev = np.zeros(self.n_systems, np.float32)
        _check(lib().og_surrogate_status(self._h, _i(mode), dev.ctypes.data_as(C.POINTER(C.c_float))))
        return mode, dev

    # -- profiling --------------------------------------------------------------
    def profile(self, on=True):
        _check(lib().og_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        ms = np.zeros(len(KERNEL_CLASSES))
        cnt = np.zeros(len(KERNEL_CLASSES), np.int64)
        _check(lib().og_profile_read(self._h, _d(ms), cnt.ctypes.data_as(C.POINTER(C.c_int64))))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}


# ---------------------------------------------------------------------------
# Drivers
# ---------------------------------------------------------------------------
def maxwell_boltzmann(sys: S.ResolvedSystem, temperature_k: float, rng: S.Xoshiro256pp,
                      remove_com: bool = True):
    """Initial velocities (langevin.rs:487-513; remd.rs:206 uses no COM removal)."""
    kbt = S.BOLTZMANN_KJ * temperature_k
    v = np.zeros((sys.n_atoms, 3))
    for i, m in enumerate(sys.mass):
        sigma = math.sqrt(kbt * S.ACCEL_FACTOR / m)
        v[i, 0] = sigma * rng.gaussian()
        v[i, 1] = sigma * rng.gaussian()
        v[i, 2] = sigma * rng.gaussian()
    if remove_com:
        p = (v * sys.mass[:, None]).sum(0)
        v -= p / sys.mass.sum()
    return v


def kinetic_temperature(ke_kj, n_atoms):
    return 2.0 * ke_kj / (3.0 * n_atoms * S.BOLTZMANN_KJ)


@dataclass
class LangevinSummary:
    steps_run: int
    temperature_mean_k: float
    temperature_stddev_k: float
    equipartition_ratio: float
    final_kinetic_energy_kj_mol: float
    atoms_count: int
    diverged: bool
    shake_failures: int
    # (step, positions [N, 3]) every save_every steps from step 0 (the
    # frames the reference hands its callback, langevin.rs:205-220, :332-341)
    frames: List = field(default_factory=list)


def run_langevin(sys: S.ResolvedSystem, steps: int, dt_fs=1.0, temperature_k=310.0,
                 friction_ps_inv=1.0, seed=0, randomise_initial_velocities=True,
                 constrain_h_bonds=False, include_sasa=False, sasa_gamma_scale=1.0,
                 mode_a=False, engine: Optional[Engine] = None, save_every=0, frame_chunk=256):
    """BAOAB Langevin for one system on the GPU (langevin.rs:128-367).
    save_every > 0 keeps a frame every save_every steps (and step 0) in the
    engine's device frame ring, read back once per frame_chunk frames."""
    eng = engine or Engine([sys])
    if randomise_initial_velocities:
        eng.set_velocities(maxwell_boltzmann(sys, temperature_k, S.Xoshiro256pp(seed)))
    eng.dynamics_setup([temperature_k], dt_fs=dt_fs, friction_ps_inv=friction_ps_inv, seed=seed,
                       constrain_h_bonds=constrain_h_bonds, mode_a=mode_a,
                       include_sasa=include_sasa, sasa_gamma_scale=sasa_gamma_scale)
    frames = []
    if save_every > 0:
        eng.frames_setup(frame_chunk + 1, save_every)
    diverged = False
    try:
        if save_every > 0:
            kes = []
            done = 0
            while done < steps:
                n = min(steps - done, frame_chunk * save_every)
                kes.append(eng.run(n, record_ke=True))
                done += n
                st_, xs = eng.read_frames()
                frames.extend(zip(st_, xs))
            ke = np.concatenate(kes)
        else:
            ke = eng.run(steps, record_ke=True)
    except EngineError:
        diverged = True
        ke = np.zeros((0, 1))
        if save_every > 0:
            st_, xs = eng.read_frames()
            frames.extend(zip(st_, xs))
    st = eng.status()
    t = kinetic_temperature(ke[:, 0], sys.n_atoms) if len(ke) else np.zeros(0)
    ke_final = float(ke[-1, 0]) if len(ke) else 0.0
    dof = 3.0 * sys.n_atoms
    return eng, LangevinSummary(
        steps_run=len(t), temperature_mean_k=float(t.mean()) if len(t) else 0.0,
        temperature_stddev_k=float(t.std()) if len(t) else 0.0,
        equipartition_ratio=ke_final / (0.5 * dof * S.BOLTZMANN_KJ * temperature_k),
        final_kinetic_energy_kj_mol=ke_final, atoms_count=sys.n_atoms,
        diverged=diverged or bool(st["diverged"][0]), shake_failures=int(st["shake_failures"][0]),
        frames=frames)


@dataclass
class RemdSummary:
    n_replicas: int
    atoms_count: int
    swap_attempts: List[int]
    swap_accepts: List[int]
    temperatures: List[float]
    final_potential_energy: List[float] = field(default_factory=list)
    diverged: List[bool] = field(default_factory=list)

    def acceptance_ratios(self):
        return [a / t if t else 0.0 for a, t in zip(self.swap_accepts, self.swap_attempts)]


def geometric_ladder(t_min, t_max, n):
    if n == 1:
        return [t_min]
    return [t_min * (t_max / t_min) ** (k / (n - 1)) for k in range(n)]


def swap_energies(eng: Engine, include_sasa: bool):
    """remd::potential_energy of every replica (remd.rs:529-539)."""
    return [e["total"] for e in eng.total_energy(mode_a=True, include_sasa=include_sasa,
                                                 sasa_method="shrake_rupley", allow_nonfinite=True)]


def run_remd(sys: S.ResolvedSystem, temperatures: Sequence[float], total_time_fs: float,
             swap_interval_fs: float, dt_fs=1.0, friction_ps_inv=2.0, seed=0,
             constrain_h_bonds=False, include_sasa=False, sasa_gamma_scale=1.0, mode_a=True,
             engine: Optional[Engine] = None, velocities=None):
    """Parallel tempering with every replica in one device batch (remd.rs:136-498).

    Forces use origami's integrator cutoffs by default (mode_a: remd.rs:209-214
    passes DEFAULT_CUTOFF_A); swap energies are remd::potential_energy
    (remd.rs:529-539): bonded + 10 A LJ / reaction-field Coulomb + GB, plus,
    with include_sasa, the Shrake-Rupley SASA energy (960 dots).

    velocities: initial velocities of all replicas, (n_replicas * N, 3) in
    ladder order (a restart); by default Maxwell-Boltzmann draws from each
    replica's seed without COM removal (remd.rs:204-206).

    Replicas advance together in the batched step; swap decisions use the
    replicas' potential energies, origami's swap RNG (seed + 0xDEADBEEF) and
    Metropolis rule on alternating adjacent pairs, then exchange positions,
    velocities and forces with the sqrt(T_a/T_b) velocity rescale.  A replica
    that diverges stops while the others keep running, and pairs with a
    stopped replica are not attempted (remd.rs:382-389, :430).
    """
    temps = sorted(float(t) for t in temperatures)
    n_rep = len(temps)
    eng = engine or Engine([sys] * n_rep)
    if velocities is None:
        v0 = []
        for i, t in enumerate(temps):
            rng = S.Xoshiro256pp((seed + i * 0xC0FFEE) & ((1 << 64) - 1))
            v0.append(maxwell_boltzmann(sys, t, rng, remove_com=False))
        velocities = np.concatenate(v0)
    eng.set_velocities(velocities)
    eng.dynamics_setup(temps, dt_fs=dt_fs, friction_ps_inv=friction_ps_inv, seed=seed,
                       constrain_h_bonds=constrain_h_bonds, mode_a=mode_a,
                       include_sasa=include_sasa, sasa_gamma_scale=sasa_gamma_scale)
    steps_per_swap = int(round(swap_interval_fs / dt_fs))
    total_steps = int(round(total_time_fs / dt_fs))
    n_rounds = -(-total_steps // max(steps_per_swap, 1))
    attempts = [0] * (n_rep - 1)
    accepts = [0] * (n_rep - 1)
    swap_rng = S.Xoshiro256pp((seed + 0xDEADBEEF) & ((1 << 64) - 1))
    for rnd in range(n_rounds):
        n = min(steps_per_swap, total_steps - rnd * steps_per_swap)
        if n <= 0:
            break
        eng.run(n, allow_stopped=True)
        stopped = eng.status()["diverged"] != 0
        energies = swap_energies(eng, include_sasa)
        for pair in range(rnd % 2, n_rep - 1, 2):
            a, b = pair, pair + 1
            if stopped[a] or stopped[b]:
                continue
            attempts[pair] += 1
            beta_a = 1.0 / (S.BOLTZMANN_KJ * temps[a])
            beta_b = 1.0 / (S.BOLTZMANN_KJ * temps[b])
            delta = (beta_a - beta_b) * (energies[a] - energies[b])
            p_acc = 1.0 if delta >= 0.0 else math.exp(delta)
            if swap_rng.next_f64() < p_acc:
                accepts[pair] += 1
                eng.swap(a, b)
                energies[a], energies[b] = energies[b], energies[a]
    final = swap_energies(eng, include_sasa)
    stopped = [bool(x) for x in eng.status()["diverged"]]
    return eng, RemdSummary(n_rep, sys.n_atoms, attempts, accepts, temps, final, stopped)
