"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the engine itself never does.  The shared library is built by
``oracle/Makefile`` into ``oracle/_build/liborigami_oracle.so``.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liborigami_oracle.so")
_lib = None

INF = math.inf
MODE_A = dict(cut_nb=10.0, cut_gb=10.0, cut_descreen=20.0)   # origami integrator cutoffs
MODE_B = dict(cut_nb=INF, cut_gb=INF, cut_descreen=INF)      # all pairs


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = ([C.c_int] + [dp] * 10 + [C.c_int, ip, C.c_int, ip] +
                                 [C.c_int, ip, dp] * 5)
        L.orc_destroy.argtypes = [C.c_void_p]
        for nm in ("orc_nonbonded_forces", "orc_nonbonded_energy"):
            getattr(L, nm).argtypes = [C.c_void_p, dp, C.c_double, dp]
        L.orc_pairwise_descreening.restype = C.c_double
        L.orc_pairwise_descreening.argtypes = [C.c_double] * 3
        L.orc_born_radii.restype = C.c_int
        L.orc_born_radii.argtypes = [C.c_void_p, dp, C.c_double, dp, dp]
        L.orc_gb_forces.argtypes = [C.c_void_p, dp, dp, C.c_double, dp]
        L.orc_gb_energy.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_bonded_forces.argtypes = [C.c_void_p, dp, dp]
        L.orc_bonded_energy.argtypes = [C.c_void_p, dp, dp]
        L.orc_total_force.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, dp, dp, dp]
        L.orc_total_energy.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, dp]
        L.orc_shake.restype = C.c_int
        L.orc_shake.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_int]
        L.orc_baoab_step.restype = C.c_int
        L.orc_baoab_step.argtypes = [C.c_void_p, dp, dp, dp, dp, C.c_double, C.c_double,
                                     C.c_double, C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_kinetic_energy.restype = C.c_double
        L.orc_kinetic_energy.argtypes = [C.c_void_p, dp]
        L.orc_xoshiro_u64.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_uint64)]
        L.orc_splitmix64.argtypes = [C.c_uint64, C.c_int, C.POINTER(C.c_uint64)]
        L.orc_xoshiro_from_state.argtypes = [C.POINTER(C.c_uint64), C.c_int, C.POINTER(C.c_uint64)]
        L.orc_philox4x32.argtypes = [C.c_uint32] * 6 + [C.POINTER(C.c_uint32)]
        L.orc_philox_normals.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32,
                                         C.POINTER(C.c_double)]
        L.orc_psa_forces.argtypes = [C.c_int, dp, dp, dp, dp, dp]
        L.orc_psa_atom_gradient.restype = C.c_int
        L.orc_psa_atom_gradient.argtypes = [C.c_int, dp, dp, C.c_int, C.c_int, ip, dp]
        L.orc_psa_area.restype = C.c_double
        L.orc_psa_area.argtypes = [C.c_int, dp, dp, dp]
        L.orc_tunnel_force.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_sr_dots.argtypes = [C.c_int, dp]
        L.orc_sr_energy.restype = C.c_double
        L.orc_sr_energy.argtypes = [C.c_int, dp, dp, dp, C.c_int, dp]
        L.orc_psa_area_from_caps.restype = C.c_double
        L.orc_psa_area_from_caps.argtypes = [C.c_int, dp, dp, C.c_double]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """FP64 restatement of origami's force/energy/integrator path for one system."""

    def __init__(self, sys):
        L = lib()
        self.sys = sys
        self.n = sys.n_atoms
        keep = []

        def d(x):
            a, p = _d(x)
            keep.append(a)
            return p

        def i(x):
            a, p = _i(x)
            keep.append(a)
            return p

        self._h = L.orc_create(
            self.n, d(sys.charge), d(sys.rmin_half), d(sys.epsilon), d(sys.rmin_half_14),
            d(sys.epsilon_14), d(sys.gb_rho), d(sys.gb_scale), d(sys.mass), d(sys.sasa_radius),
            d(sys.sasa_gamma), len(sys.excl_pairs), i(sys.excl_pairs), len(sys.pair14),
            i(sys.pair14), len(sys.bonds), i(sys.bonds), d(sys.bond_params), len(sys.angles),
            i(sys.angles), d(sys.angle_params), len(sys.dihedrals), i(sys.dihedrals),
            d(sys.dihedral_params), len(sys.impropers), i(sys.impropers),
            d(sys.improper_params), len(sys.constraints), i(sys.constraints),
            d(sys.constraint_len))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    def _pos(self, pos):
        return _d(np.asarray(pos, dtype=np.float64).reshape(-1, 3))

    # -- forces ----------------------------------------------------------
    def nonbonded_forces(self, pos, cutoff):
        p, pp = self._pos(pos)
        f, fp = _d(np.zeros((self.n, 3)))
        lib().orc_nonbonded_forces(self._h, pp, cutoff, fp)
        return f

    def born_radii(self, pos, cutoff):
        p, pp = self._pos(pos)
        integ, ip_ = _d(np.zeros(self.n))
        radii, rp = _d(np.zeros(self.n))
        clamped = lib().orc_born_radii(self._h, pp, cutoff, ip_, rp)
        return integ, radii, clamped

    def gb_forces(self, pos, radii, cutoff):
        p, pp = self._pos(pos)
        r, rp = _d(radii)
        f, fp = _d(np.zeros((self.n, 3)))
        lib().orc_gb_forces(self._h, pp, rp, cutoff, fp)
        return f

    def bonded_forces(self, pos):
        p, pp = self._pos(pos)
        f, fp = _d(np.zeros((self.n, 3)))
        lib().orc_bonded_forces(self._h, pp, fp)
        return f

    def total_force(self, pos, cut_nb=INF, cut_gb=INF, cut_descreen=INF, terms=False):
        p, pp = self._pos(pos)
        f, fp = _d(np.zeros((self.n, 3)))
        radii, rp = _d(np.zeros(self.n))
        t, tp = _d(np.zeros((3, self.n, 3)))
        lib().orc_total_force(self._h, pp, cut_nb, cut_gb, cut_descreen, fp, rp, tp)
        if terms:
            return f, {"bonded": t[0], "nonbonded": t[1], "gb": t[2], "radii": radii}
        return f

    def sasa_forces(self, pos, gamma_scale=1.0):
        p, pp = self._pos(pos)
        r, rp = _d(self.sys.sasa_radius)
        g, gp = _d(self.sys.sasa_gamma * gamma_scale)
        f, fp = _d(np.zeros((self.n, 3)))
        st, sp = _d(np.zeros(6))
        lib().orc_psa_forces(self.n, pp, rp, gp, fp, sp)
        return f, dict(zip(["work_atoms", "atoms_with_arcs", "caps", "arcs", "vertices",
                            "cap_pairs"], st.tolist()))

    def sasa_area(self, pos):
        p, pp = self._pos(pos)
        r, rp = _d(self.sys.sasa_radius)
        per, perp = _d(np.zeros(self.n))
        total = lib().orc_psa_area(self.n, pp, rp, perp)
        return total, per

    # -- energies --------------------------------------------------------
    def nonbonded_energy(self, pos, cutoff):
        p, pp = self._pos(pos)
        out, op = _d(np.zeros(4))
        lib().orc_nonbonded_energy(self._h, pp, cutoff, op)
        return {"lj": out[0], "coulomb": out[1], "pairs": out[2], "one_four": out[3]}

    def gb_energy(self, pos, radii):
        p, pp = self._pos(pos)
        r, rp = _d(radii)
        out, op = _d(np.zeros(3))
        lib().orc_gb_energy(self._h, pp, rp, op)
        return {"total": out[0], "self": out[1], "pair": out[2]}

    def bonded_energy(self, pos):
        p, pp = self._pos(pos)
        out, op = _d(np.zeros(4))
        lib().orc_bonded_energy(self._h, pp, op)
        return dict(zip(["bond", "angle", "dihedral", "improper"], out.tolist()))

    def total_energy(self, pos, cut_nb=INF, cut_descreen=INF):
        p, pp = self._pos(pos)
        out, op = _d(np.zeros(10))
        lib().orc_total_energy(self._h, pp, cut_nb, cut_descreen, op)
        return dict(zip(["bond", "angle", "dihedral", "improper", "lj", "coulomb", "gb",
                         "gb_self", "gb_pair", "total"], out.tolist()))

    # -- integrator ------------------------------------------------------
    def shake(self, pos, ref, tol_sq=1e-6, max_iters=64):
        p, pp = _d(np.array(pos, dtype=np.float64).reshape(-1, 3))
        r, rp = self._pos(ref)
        it = lib().orc_shake(self._h, pp, rp, tol_sq, max_iters)
        return p, it

    def baoab_step(self, pos, vel, forces, xi, dt, temperature, gamma_ps, use_shake,
                   cut_nb=INF, cut_gb=INF, cut_descreen=INF):
        """In-place on copies; returns (pos, vel, forces, shake_failures)."""
        p, pp = _d(np.array(pos, dtype=np.float64).reshape(-1, 3))
        v, vp = _d(np.array(vel, dtype=np.float64).reshape(-1, 3))
        f, fp = _d(np.array(forces, dtype=np.float64).reshape(-1, 3))
        x, xp = _d(np.asarray(xi, dtype=np.float64).reshape(-1, 3))
        fails = lib().orc_baoab_step(self._h, pp, vp, fp, xp, dt, temperature, gamma_ps,
                                     int(use_shake), cut_nb, cut_gb, cut_descreen)
        return p, v, f, fails

    def kinetic_energy(self, vel):
        v, vp = self._pos(vel)
        return lib().orc_kinetic_energy(self._h, vp)


def pairwise_descreening(r, rho_i_tilde, s_rho_j_tilde):
    return lib().orc_pairwise_descreening(r, rho_i_tilde, s_rho_j_tilde)


def xoshiro_u64(seed, count):
    out = (C.c_uint64 * count)()
    lib().orc_xoshiro_u64(seed, count, out)
    return [int(x) for x in out]


def splitmix64(seed, count):
    out = (C.c_uint64 * count)()
    lib().orc_splitmix64(seed, count, out)
    return [int(x) for x in out]


def xoshiro_from_state(state, count):
    st = (C.c_uint64 * 4)(*state)
    out = (C.c_uint64 * count)()
    lib().orc_xoshiro_from_state(st, count, out)
    return [int(x) for x in out]


def philox4x32(key, ctr):
    out = (C.c_uint32 * 4)()
    lib().orc_philox4x32(key[0], key[1], ctr[0], ctr[1], ctr[2], ctr[3], out)
    return [int(x) for x in out]


def philox_normals(seed, system, n_atoms, step):
    """The engine's O-step noise for one system and step, [n_atoms, 3]."""
    out, op = _d(np.zeros((n_atoms, 3)))
    lib().orc_philox_normals(seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF, system, n_atoms, step, op)
    return out


def psa_atom_gradient(pos, radii, atom, neighbours):
    n = len(radii)
    p, pp = _d(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    r, rp = _d(radii)
    nb, nbp = _i(neighbours)
    g, gp = _d(np.zeros((n, 3)))
    na = lib().orc_psa_atom_gradient(n, pp, rp, atom, len(nb), nbp, gp)
    return g, na


def psa_area(pos, radii):
    n = len(radii)
    p, pp = _d(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    r, rp = _d(radii)
    per, perp = _d(np.zeros(n))
    total = lib().orc_psa_area(n, pp, rp, perp)
    return total, per


def psa_area_from_caps(axes, cos_a, radius):
    a, ap = _d(np.asarray(axes, dtype=np.float64).reshape(-1, 3))
    c, cp = _d(cos_a)
    return lib().orc_psa_area_from_caps(len(c), ap, cp, radius)


def sr_dots(n):
    out, op = _d(np.zeros((n, 3)))
    lib().orc_sr_dots(n, op)
    return out


def sr_energy(pos, radii, gamma, n_dots=960):
    """Shrake-Rupley energy and per-atom areas (sasa.rs:113-178)."""
    p, pp = _d(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    r, rp = _d(radii)
    g, gp = _d(gamma)
    area, ap = _d(np.zeros(len(p)))
    e = lib().orc_sr_energy(len(p), pp, rp, gp, n_dots, ap)
    return e, area


def tunnel_force(pos, origin, direction, radius, length, k):
    """Forces of the cylindrical exit tunnel (cotranslate.rs:131-155)."""
    p, pp = _d(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    o, op = _d(origin)
    d, dp_ = _d(direction)
    f, fp = _d(np.zeros((len(p), 3)))
    lib().orc_tunnel_force(len(p), pp, op, dp_, radius, length, k, fp)
    return f

