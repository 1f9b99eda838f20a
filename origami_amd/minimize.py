"""Batched L-BFGS energy minimisation on the GPU engine.

Follows origami's minimiser for every system of a batch independently
(systems advance in lock-step so each energy/force evaluation is one batched
GPU call):

  lbfgs                 dynamics/src/lbfgs.rs:61-210   (m = 10, SD fallback)
  two_loop_recursion    dynamics/src/lbfgs.rs:212-254
  backtracking          dynamics/src/line_search.rs:52-101  (Armijo c1 = 1e-4, x0.5)
  minimize defaults     dynamics/src/minimize.rs:27-38 (500 steps, |g|inf < 1, dE < 0.01,
                        max step 0.1 Å)

Forces come from the FP32-pair / fixed-point force path, energies (line
search) from the FP64 energy pass; the two-loop recursion runs in FP64 on
the host.  ``mode_a`` selects origami's 10 Å cutoffs (its minimiser's
setting, energy_eval.rs:20-50) or all pairs.
"""
from __future__ import annotations

from collections import deque
from dataclasses import dataclass
from typing import List

import numpy as np


@dataclass
class LbfgsResult:
    steps: int
    final_energy: float
    initial_energy: float
    max_force: float
    converged: bool
    line_search_failures: int


def _two_loop(g, hs, hy, hr):
    m = len(hs)
    q = g.copy()
    alpha = [0.0] * m
    for i in range(m - 1, -1, -1):
        a = hr[i] * float(np.dot(hs[i], q))
        alpha[i] = a
        q -= a * hy[i]
    s_y = float(np.dot(hs[-1], hy[-1]))
    y_y = float(np.dot(hy[-1], hy[-1]))
    gamma = s_y / y_y if y_y > 0.0 else 1.0
    r = gamma * q
    for i in range(m):
        beta = hr[i] * float(np.dot(hy[i], r))
        r += (alpha[i] - beta) * hs[i]
    return -r


def minimize_device(engine, max_steps=500, gradient_tol=1.0, energy_tol=0.01, max_step_a=0.1,
                    history=10, mode_a=True, alpha0=1.0, contraction=0.5, c1=1e-4,
                    min_alpha=1e-12, include_sasa=False, sasa_gamma_scale=1.0,
                    cutoff_a=10.0) -> List[LbfgsResult]:
    """The same minimiser run on the device (og_minimize): every system's
    L-BFGS state, history and line search stay in device memory; the host
    only sequences the phases.  Same arguments and results as minimize()."""
    import ctypes as C
    from . import engine as E
    p = E.MinimizeParams(int(max_steps), int(history), gradient_tol, energy_tol, max_step_a, alpha0,
                         contraction, c1, min_alpha,
                         (E.MODE_A if mode_a else 0) | (E.SASA if include_sasa else 0), cutoff_a,
                         sasa_gamma_scale)
    out = (E.MinimizeResult * engine.n_systems)()
    E._check(E.lib().og_minimize(engine._h, C.byref(p), out))
    return [LbfgsResult(int(r.steps), float(r.final_energy), float(r.initial_energy),
                        float(r.max_force), bool(r.converged), int(r.line_search_failures))
            for r in out]


def minimize(engine, max_steps=500, gradient_tol=1.0, energy_tol=0.01, max_step_a=0.1,
             history=10, mode_a=True, alpha0=1.0, contraction=0.5, c1=1e-4,
             min_alpha=1e-12) -> List[LbfgsResult]:
    """Minimise every system of ``engine`` in place; returns one result each."""
    ns = engine.n_systems
    off = engine.offsets

    def energies(x):
        return np.array([e["total"] for e in engine.total_energy(x, mode_a=mode_a)])

    def grads(x):
        return -engine.total_force(x, mode_a=mode_a)

    def flat(a, k):
        return a[off[k]:off[k + 1]].reshape(-1)

    x = engine.positions()
    e_init = energies(x)
    e_prev = e_init.copy()
    hs = [deque() for _ in range(ns)]
    hy = [deque() for _ in range(ns)]
    hr = [deque() for _ in range(ns)]
    active = np.ones(ns, bool)
    converged = np.zeros(ns, bool)
    steps = np.zeros(ns, np.int64)
    ls_fail = np.zeros(ns, np.int64)
    max_force = np.full(ns, np.inf)

    def line_search(x, dirs, gflat, e0, a0s, who):
        """Batched Armijo backtracking.  Returns (alpha, new_energy) per system
        in ``who`` (None entries = failure)."""
        res = {k: None for k in who}
        pdg = {}
        alpha = {}
        pending = []
        for k in who:
            p_dot_g = float(np.dot(dirs[k], gflat[k]))
            if p_dot_g >= 0.0:
                continue
            pdg[k] = p_dot_g
            alpha[k] = a0s[k]
            pending.append(k)
        while pending:
            trial = x.copy()
            for k in pending:
                trial[off[k]:off[k + 1]] += (alpha[k] * dirs[k]).reshape(-1, 3)
            et = energies(trial)
            nxt = []
            for k in pending:
                if et[k] <= e0[k] + c1 * alpha[k] * pdg[k]:
                    res[k] = (alpha[k], float(et[k]))
                else:
                    alpha[k] *= contraction
                    if alpha[k] >= min_alpha:
                        nxt.append(k)
            pending = nxt
        return res

    for step in range(max_steps):
        if not active.any():
            break
        g = grads(x)
        gflat = [flat(g, k).copy() for k in range(ns)]
        for k in np.where(active)[0]:
            steps[k] = step
            max_force[k] = float(np.abs(gflat[k]).max())
            if max_force[k] < gradient_tol:
                converged[k] = True
                active[k] = False
        who = list(np.where(active)[0])
        if not who:
            break
        dirs, a0s = {}, {}
        for k in who:
            d = -gflat[k] if not hs[k] else _two_loop(gflat[k], hs[k], hy[k], hr[k])
            md = float(np.abs(d).max())
            dirs[k] = d
            a0s[k] = min(alpha0, max_step_a / md) if md > 0.0 else alpha0
        e_now = energies(x)
        res = line_search(x, dirs, gflat, e_now, a0s, who)
        sd_who = [k for k in who if res[k] is None]
        sd_res = {}
        if sd_who:
            sd_dirs, sd_a0 = {}, {}
            for k in sd_who:
                ls_fail[k] += 1
                d = -gflat[k]
                md = float(np.abs(d).max())
                sd_dirs[k] = d
                sd_a0[k] = min(alpha0, max_step_a / md) if md > 0.0 else alpha0
            sd_res = line_search(x, sd_dirs, gflat, e_now, sd_a0, sd_who)
        lbfgs_moved = []
        for k in who:
            if res[k] is None:
                r2 = sd_res.get(k)
                if r2 is None:
                    ls_fail[k] += 1
                    active[k] = False
                    continue
                x[off[k]:off[k + 1]] += (r2[0] * sd_dirs[k]).reshape(-1, 3)
                hs[k].clear(), hy[k].clear(), hr[k].clear()
                e_prev[k] = r2[1]
                continue
            x[off[k]:off[k + 1]] += (res[k][0] * dirs[k]).reshape(-1, 3)
            lbfgs_moved.append(k)
        if lbfgs_moved:
            g_after = grads(x)
            for k in lbfgs_moved:
                s_vec = res[k][0] * dirs[k]
                y_vec = flat(g_after, k) - gflat[k]
                s_dot_y = float(np.dot(s_vec, y_vec))
                if s_dot_y > 1e-12:
                    if len(hs[k]) == history:
                        hs[k].popleft(), hy[k].popleft(), hr[k].popleft()
                    hs[k].append(s_vec)
                    hy[k].append(y_vec)
                    hr[k].append(1.0 / s_dot_y)
                e_new = res[k][1]
                if abs(e_prev[k] - e_new) < energy_tol and max_force[k] < gradient_tol:
                    converged[k] = True
                    active[k] = False
                    steps[k] = step + 1
                e_prev[k] = e_new
    engine.set_positions(x)
    e_fin = energies(x)
    return [LbfgsResult(int(steps[k]), float(e_fin[k]), float(e_init[k]), float(max_force[k]),
                        bool(converged[k]), int(ls_fail[k])) for k in range(ns)]
