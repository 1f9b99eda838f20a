#!/usr/bin/env python3
"""Benchmark: aggregate ns/day and pair-interactions/s on one MI355X.

Workload (default): 64 independent chains of seeded random sequence with
lengths drawn uniformly from 10..300 residues (about 157k atoms, 2.5e8 atom
pairs per pass), each at its own bath temperature in 280..395 K; all-pairs
LJ + Coulomb + GB (descreening + pair force) plus analytical SASA forces
(gamma scale 0.25), BAOAB at dt = 2 fs with SHAKE on X-H bonds, friction
2/ps.  One "step" is one BAOAB step of every chain (forces included).  The
chains are built extended and relaxed by a short batched L-BFGS before the
clock starts (untimed), so that dt = 2 fs is stable.

Prints ONE JSON line.  The timed run steps with SASA on its own stream beside
the all-pairs chain; og_run replays captured hipGraphs for batches under
20 000 atoms (the REMD batch) and launches directly above (the 64-chain
batch, where direct launches measured 2-3 % faster).  Kernel timing for the
roofline figure comes from a separate run with HIP events around every
launch on one stream (kernels one at a time).  The CPU baseline times the
FP64 oracle on a bounded sample of the same workload, on the host cores this
process may use.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 4]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "ns/day + pair-interactions/s (GB+LJ+SASA) on 1 GPU; surrogate vs plain"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
# Algorithmic work per unordered pair (counting convention: +,-,*,/,sqrt,exp,ln = 1; fma = 2)
FLOPS_DESCREEN = 53
FLOPS_FORCE = 54


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_config4(seed=4, n_chains=64):
    from origami_amd import system as S
    rng = S.Xoshiro256pp(seed)
    specs = []
    for _ in range(n_chains):
        L = 10 + int(rng.next_f64() * 291.0)
        T = 280.0 + 115.0 * rng.next_f64()
        specs.append((L, T))
    systems = [S.resolve(S.build_chain(S.random_sequence(rng, L))) for L, _ in specs]
    return systems, [T for _, T in specs]


def build_config2(seed=2):
    from origami_amd import system as S
    return [S.seeded_chain(seed, 80)], [310.0]


def build_config3(seed=3, n_rep=8):
    from origami_amd import engine as E
    from origami_amd import system as S
    s = S.seeded_chain(seed, 10)
    return [s] * n_rep, E.geometric_ladder(280.0, 395.0, n_rep)


def _oracle_step(O, s, gamma_scale):
    orc = O.Oracle(s)
    f = orc.total_force(s.pos)
    f += orc.sasa_forces(s.pos, gamma_scale)[0]
    return f


def cpu_baseline(systems, temps, dt_fs, gamma_scale, budget_s=12.0, budget_1core_s=6.0):
    """Time the FP64 oracle on a bounded sample of the batch: full force
    evaluations (bonded + LJ/Coulomb + GB + SASA, all pairs) of chains taken
    in a fixed shuffled order until the time budget is used, scaled by pair
    count to one step of the whole batch.  Once on every host core this
    process may use (one chain per thread; the oracle's C calls release the
    interpreter lock) and once on a single core."""
    import concurrent.futures as cf
    import threading
    from oracle import oracle as O
    order = [int(k) for k in np.random.default_rng(0).permutation(len(systems))]
    total_pairs = sum(s.n_atoms * (s.n_atoms - 1) // 2 for s in systems)
    pairs = lambda k: systems[k].n_atoms * (systems[k].n_atoms - 1) // 2

    def ns_day(step_s):
        return len(systems) * dt_fs * 1e-6 / step_s * 86400.0

    # single core
    done1, pairs1, t1 = [], 0, 0.0
    for k in order:
        t0 = time.perf_counter()
        _oracle_step(O, systems[k], gamma_scale)
        t1 += time.perf_counter() - t0
        pairs1 += pairs(k)
        done1.append(k)
        if t1 >= budget_1core_s:
            break
    step1 = t1 * total_pairs / max(pairs1, 1)
    # all cores: workers take chains from the shuffled order until the budget is spent
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))
    lock = threading.Lock()
    state = {"next": 0, "pairs": 0, "done": 0}
    t_start = time.perf_counter()

    def worker():
        while True:
            with lock:
                if state["next"] >= len(order) or time.perf_counter() - t_start >= budget_s:
                    return
                k = order[state["next"]]
                state["next"] += 1
            _oracle_step(O, systems[k], gamma_scale)
            with lock:
                state["pairs"] += pairs(k)
                state["done"] += 1

    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        for fut in [ex.submit(worker) for _ in range(cores)]:
            fut.result()
    wall = time.perf_counter() - t_start
    step_n = wall * total_pairs / max(state["pairs"], 1)
    return {"value": ns_day(step_n), "unit": "ns/day", "cores": cores, "kind": "port",
            "value_1core": ns_day(step1),
            "sample": "FP64 oracle force evaluation (bonded + LJ/Coulomb + GB + SASA, all pairs) of "
                      "%d of %d chains on %d threads (%d pairs, %.1f s wall), scaled by pair count to one "
                      "step of the batch (%.2f s/step); single core: %d chains, %.1f s (%.2f s/step)"
                      % (state["done"], len(systems), cores, state["pairs"], wall, step_n,
                         len(done1), t1, step1)}


def surrogate_leg(eng, systems, temps, args, dt_fs, gamma_scale, include_sasa, ns_plain, pairs):
    """Learned Born radii: collect plain-GB radii along the batch's own
    trajectory, train on the GPU, run the same batch with the surrogate in
    place of the descreening pass (plain-GB check every 100 steps, 2 %
    fallback) and report speed and error against plain GB and the FP64 oracle."""
    from oracle import oracle as O
    from origami_amd import engine as E
    n_snap = 7
    eng.surrogate_setup(n_snap * eng.n_atoms, seed=99)
    t0 = time.time()
    for _ in range(n_snap):
        eng.run(10)
        rows = eng.surrogate_collect()
    t_collect = time.time() - t0
    t0 = time.time()
    loss = eng.surrogate_train(args.train_steps, batch=8192, lr=2e-3, seed=5)
    t_train = time.time() - t0
    log("surrogate: %d rows in %.1f s, %d Adam steps in %.2f s, loss %.3e -> %.3e"
        % (rows, t_collect, args.train_steps, t_train, loss[:20].mean(), loss[-20:].mean()))
    eng.surrogate_enable(check_every=100, tolerance=0.02)
    eng.dynamics_setup(temps, dt_fs=dt_fs, friction_ps_inv=2.0, seed=2025,
                       constrain_h_bonds=True, include_sasa=include_sasa,
                       sasa_gamma_scale=gamma_scale, surrogate=True)
    # launch widths for the evaluation with learned radii (its kernels share
    # the CUs with the SASA arrangement differently from the descreening pass)
    tune_s = eng.autotune(include_sasa=include_sasa, surrogate=True)
    log("autotune (surrogate):", tune_s)
    eng.run(args.warmup)
    t0 = time.perf_counter()
    eng.run(args.steps)
    wall = time.perf_counter() - t0
    eng.profile(True)
    eng.run(args.profile_steps)
    prof = eng.profile_read()
    eng.profile(False)
    mode, dev = eng.surrogate_status()
    ns_sur = len(systems) * dt_fs * args.steps * 1e-6 / wall * 86400.0
    k_ms = {k: v[0] / max(v[1], 1) for k, v in prof.items() if v[1] > 0}
    # errors at the final state: surrogate vs plain GB (whole batch) ...
    pos = eng.positions()
    r_plain = (eng.total_force(pos), eng.born_radii()[0])[1]
    r_sur = eng.surrogate_radii(pos)
    f_sur, e_sur = eng.total_force(pos, surrogate=True, energies=True)
    f_pl, e_pl = eng.total_force(pos, energies=True)
    frms = float(np.sqrt(((f_sur - f_pl) ** 2).sum() / (f_pl ** 2).sum()))
    gb_sur = np.array([d["gb_self"] + d["gb_pair"] for d in eng.total_energy(pos, surrogate=True)])
    gb_pl = np.array([d["gb_self"] + d["gb_pair"] for d in eng.total_energy(pos)])
    # ... and vs the FP64 oracle on the four smallest chains
    small = np.argsort([s.n_atoms for s in systems])[:4]
    parts = eng.split(pos)
    fs_parts = eng.split(f_sur)
    num = den = 0.0
    for k in small:
        fo = O.Oracle(systems[k]).total_force(parts[k])
        num += float(((fs_parts[k] - fo) ** 2).sum())
        den += float((fo ** 2).sum())
    return {
        "value": ns_sur, "unit": "ns/day", "speedup_vs_plain": ns_sur / ns_plain, "autotune": tune_s,
        "ms_per_step": wall * 1e3 / max(args.steps, 1),
        "pair_interactions_per_s": pairs / ((k_ms.get("force", 0) + k_ms.get("surrogate", 0)
                                             + k_ms.get("descreen", 0)) * 1e-3),
        "kernel_ms_per_step": {k: v[0] / max(args.profile_steps, 1) for k, v in prof.items() if v[1] > 0},
        "training": {"rows": rows, "adam_steps": args.train_steps, "batch": 8192,
                     "collect_s": t_collect, "train_s": t_train,
                     "loss_first": float(loss[:20].mean()), "loss_last": float(loss[-20:].mean()),
                     "source": "plain-GB radii of every atom at 7 snapshots of this batch's own "
                               "trajectory (10 steps apart)"},
        "systems_on_surrogate": int(mode.sum()), "systems_fallen_back": int(len(mode) - mode.sum()),
        "max_radius_deviation_last_check": float(dev.max()),
        "radius_max_rel_err_vs_plain": float(np.max(np.abs(r_sur - r_plain) / r_plain)),
        "force_rel_rms_vs_plain_gb": frms,
        "gb_energy_max_rel_err_vs_plain": float(np.max(np.abs(gb_sur - gb_pl) / np.abs(gb_pl))),
        "force_rel_rms_vs_fp64_oracle_4_smallest_chains": float(np.sqrt(num / den)),
        "note": "plain GB's Born radii equal rho - 0.09 for every atom (the descreening integral is "
                "never positive), so the surrogate replaces an all-pairs pass whose result is "
                "a per-element constant; the speedup is bounded by the descreening pass's share "
                "of the step",
    }


def pmc_probe(E):
    """Workload for rocprofv3 --pmc runs (FETCH_SIZE / WRITE_SIZE per dispatch)."""
    import torch
    n = 128 * 1024 * 1024                       # 512 MiB of float32
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    systems, _ = build_config4()
    eng = E.Engine(systems)
    # the autotuner's usual choice for this batch on 256 CUs: 3 arrangement
    # blocks per CU, one pass holding 128 caps per atom, force strips of 8
    # j tiles
    eng.set_launch(sasa_blocks=3 * 256, sasa_passes=0)
    eng.strip_tiles(8)
    for _ in range(3):
        eng.total_force(include_sasa=True, sasa_gamma_scale=0.25)
    print(json.dumps({"pmc_probe": "copy 536870912 bytes read + written; 3 force evaluations",
                      "atoms": eng.n_atoms}), flush=True)
    return 0


# The reference's published timings (README.md:406-410, release build on an
# Apple M3 Pro): total_force per step with the 10 Å integrator cutoffs, and
# simulated time per wall day at dt = 2 fs with SHAKE.
REFERENCE_SYSTEMS = [("1UAO_chignolin", 134, 0.27, 500.0), ("1L2Y_model1", 300, 1.12, 150.0),
                     ("2F4K_villin_hp35", 520, 2.76, 60.0)]


def fixtures_leg(E, n_calls=200, n_steps=2000):
    """og_forces (origami's cutoffs, host buffers in and out: the entry-point
    replacement of total_force_with_scratch) and device-resident dynamics
    (dt 2 fs, SHAKE, 300 K) on the reference's own PDB systems."""
    from origami_amd import pdb as P
    from origami_amd import system as S
    from origami_amd.minimize import minimize_device
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden")
    out = []
    for name, atoms, ref_ms, ref_ns_day in REFERENCE_SYSTEMS:
        s = P.system_from_residues(P.read_pdb(os.path.join(golden, name + ".pdb")))
        eng = E.Engine([s])
        x = s.pos.copy()
        for _ in range(5):
            eng.total_force(x, mode_a=True)
        t0 = time.perf_counter()
        for _ in range(n_calls):
            eng.total_force(x, mode_a=True)
        ms_call = (time.perf_counter() - t0) * 1e3 / n_calls
        # dynamics from the relaxed entry (the deposited coordinates carry
        # close contacts that a 2 fs step turns into a blow-up)
        minimize_device(eng, max_steps=200, mode_a=True)
        eng.set_velocities(E.maxwell_boltzmann(s, 300.0, S.Xoshiro256pp(7)))
        eng.dynamics_setup([300.0], dt_fs=2.0, friction_ps_inv=2.0, seed=7, constrain_h_bonds=True,
                           mode_a=True)
        eng.run(100)
        t0 = time.perf_counter()
        eng.run(n_steps)
        ms_step = (time.perf_counter() - t0) * 1e3 / n_steps
        ns_day = 2.0 * 1e-6 / (ms_step * 1e-3) * 86400.0
        out.append({"system": name, "atoms": s.n_atoms, "og_forces_ms_per_call": ms_call,
                    "reference_total_force_ms_m3pro": ref_ms, "dynamics_ms_per_step": ms_step,
                    "dynamics_ns_per_day": ns_day, "reference_ns_per_day_m3pro": ref_ns_day})
        log("fixture %s: og_forces %.3f ms/call (reference %.2f), dynamics %.4f ms/step = %.0f ns/day "
            "(reference %.0f)" % (name, ms_call, ref_ms, ms_step, ns_day, ref_ns_day))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", type=int, default=4, choices=[2, 3, 4])
    ap.add_argument("--minimize-steps", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sasa", action="store_true", help="(diagnostic) leave SASA out")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the learned-radii leg (training + timed surrogate run)")
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--replicas", type=int, default=8, help="config 3: ladder size (8 or 64)")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="short workload for hardware-counter runs: a 512 MiB device copy (the "
                         "calibration kernel) and three force evaluations of the 64-chain batch")
    ap.add_argument("--step-mode", choices=["auto", "graph", "direct", "persistent"], default="auto",
                    help="how og_run issues a step: by batch size (auto), captured graphs, direct "
                         "launches, or the persistent replica-step kernel")
    ap.add_argument("--no-fixtures", action="store_true",
                    help="skip the timing of the reference's own PDB systems (1UAO, 1L2Y, 2F4K)")
    ap.add_argument("--profile-steps", type=int, default=20,
                    help="steps of the separate profiled run that gives the kernel breakdown")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU only; --gpus %d is not supported "
              "(whole chains and replicas would go to one engine per GPU)"
              % args.gpus, file=sys.stderr)
        return 2

    from origami_amd import engine as E
    from origami_amd.minimize import minimize_device
    if args.pmc_probe:
        return pmc_probe(E)

    gamma_scale = 0.25
    dt_fs = 2.0
    include_sasa = not args.no_sasa
    t0 = time.time()
    if args.config == 4:
        systems, temps = build_config4()
        workload = ("64 chains, seeded random sequences, 10-300 residues (uniform), extended "
                    "start + 100-step L-BFGS relax, own T in 280-395 K; all-pairs LJ + Coulomb "
                    "+ GB + PowerSasa SASA forces (gamma scale 0.25); BAOAB dt 2 fs + SHAKE X-H")
    elif args.config == 2:
        systems, temps = build_config2()
        workload = ("80-residue seeded chain, extended + L-BFGS; all-pairs LJ + Coulomb + GB + "
                    "PowerSasa (gamma scale 0.25); BAOAB dt 2 fs + SHAKE")
    else:
        systems, temps = build_config3(n_rep=args.replicas)
        include_sasa = False
        workload = ("REMD batch: %d replicas of a seeded 10-residue chain, 280-395 K geometric "
                    "ladder; all-pairs GB + LJ + Coulomb; BAOAB dt 2 fs + SHAKE; one captured "
                    "graph per step for all replicas" % args.replicas)
    log("built %d systems in %.1f s" % (len(systems), time.time() - t0))
    eng = E.Engine(systems)
    info = eng.info()
    log("engine:", info)
    t0 = time.time()
    res = minimize_device(eng, max_steps=args.minimize_steps, mode_a=False)
    relax_s = time.time() - t0
    log("L-BFGS relax on the device (%d steps max): %.2f s, energy %.3e -> %.3e" % (
        args.minimize_steps, relax_s, sum(r.initial_energy for r in res),
        sum(r.final_energy for r in res)))
    from origami_amd import system as S
    v = [E.maxwell_boltzmann(s, T, S.Xoshiro256pp(1000 + k)) for k, (s, T) in
         enumerate(zip(systems, temps))]
    eng.set_velocities(np.concatenate(v))
    eng.dynamics_setup(temps, dt_fs=dt_fs, friction_ps_inv=2.0, seed=2024,
                       constrain_h_bonds=True, include_sasa=include_sasa,
                       sasa_gamma_scale=gamma_scale)
    tune = eng.autotune(include_sasa=include_sasa)
    log("autotune:", tune)
    eng.use_graphs({"auto": None, "graph": True, "direct": False, "persistent": "persistent"}[args.step_mode])
    eng.run(args.warmup)
    # timed run: no per-kernel events (captured graphs below 20 000 atoms)
    t0 = time.perf_counter()
    eng.run(args.steps)
    wall = time.perf_counter() - t0
    # kernel breakdown from a separate run with HIP events around every launch
    eng.profile(True)
    eng.run(args.profile_steps)
    prof = eng.profile_read()
    eng.profile(False)
    st = eng.status()
    n_sys = len(systems)
    ms_per_step = wall * 1e3 / max(args.steps, 1)
    ns_day = n_sys * dt_fs * args.steps * 1e-6 / wall * 86400.0
    pairs = info["pairs"]
    k_ms = {k: v[0] / max(v[1], 1) for k, v in prof.items() if v[1] > 0}
    pair_ips = 2 * pairs / ((k_ms.get("descreen", 0) + k_ms.get("force", 0)) * 1e-3)
    # Roofline of the all-pairs force kernel (LJ + Coulomb + GB pair, the graded
    # FP32 kernel); the kernel class with the most time is reported beside it.
    totals = {k: v[0] for k, v in prof.items()}
    dom = max(totals, key=totals.get)
    flops = {"descreen": FLOPS_DESCREEN, "force": FLOPS_FORCE}
    ach = flops["force"] * pairs / (k_ms["force"] * 1e-3) / 1e12
    roof = {"kernel": "og_k_force_b (all-pairs LJ + Coulomb + GB pair force)", "bound": "valu",
            "roof": "FP32 vector ALU, 157.3 TFLOP/s (MI355X FP32 peak)", "achieved": ach,
            "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP32_PEAK_TFLOPS,
            "traffic": None, "flops_per_pair": FLOPS_FORCE, "pairs_per_launch": pairs,
            "avg_ms": k_ms["force"],
            "flops_note": ("algorithmic flops of the pair term for every pair; j tiles whose every pair "
                           "has exp(-r^2/(4 R_i R_j)) <= 2^-32 leave out the GB exponential and the second "
                           "reciprocal square root, which provably cannot change the FP32 result there")}
    # HBM bytes per launch of the force kernel, from the rocprofv3 --pmc runs
    # summarised in profiles/ (same kernel, same 157,072-atom batch)
    tf = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                      "hbm_traffic_64chains_157k_atoms.json")
    if os.path.exists(tf):
        tr = json.load(open(tf))
        if tr.get("atoms") == info["atoms"] and "og_k_force_b" in tr["kernels"]:
            kb = tr["kernels"]["og_k_force_b"]
            roof["traffic"] = kb["fetch_bytes"] + kb["write_bytes"]
            roof["traffic_note"] = ("HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated) from "
                                    "profiles/hbm_traffic_64chains_157k_atoms.md; algorithmic bytes "
                                    "13.2 MB, the rest is int64 fixed-point atomics")
    dominant = {"kernel_class": dom, "ms_per_step": totals[dom] / max(args.profile_steps, 1),
                "bound": "latency / divergence (FP64 PowerSasa arrangement)" if dom == "sasa" else None}
    # VALU activity of the SASA arrangement and the force kernel, from the
    # rocprofv3 --pmc run summarised in profiles/ (same batch)
    vf = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "valu_64chains_157k_atoms.json")
    if os.path.exists(vf):
        vj = json.load(open(vf))
        if vj.get("atoms") == info["atoms"]:
            arr = [v for k, v in vj["kernels"].items() if k.startswith("void og_k_sasa_arrange<false")]
            ka = max(arr, key=lambda v: v["valu_insts"]) if arr else None
            if ka and dom == "sasa":
                dominant["valu_busy"] = ka["valu_busy"]
                dominant["valu_lane_utilization"] = ka["valu_lane_utilization"]
                dominant["valu_insts_per_sasa_atom"] = ka["valu_insts"] / max(vj.get("sasa_atoms", 1), 1)
                dominant["valu_note"] = "og_k_sasa_arrange from profiles/valu_64chains_157k_atoms.md"
            kf = vj["kernels"].get("og_k_force_b")
            if kf:
                roof["valu_busy"] = kf["valu_busy"]
                roof["valu_lane_utilization"] = kf["valu_lane_utilization"]
    # FP32 pair kernels beside it
    pk = {}
    for k in ("descreen", "force"):
        if k in k_ms:
            ach = flops[k] * pairs / (k_ms[k] * 1e-3) / 1e12
            pk[k] = {"avg_ms": k_ms[k], "achieved_tflops": ach, "frac_fp32_peak": ach / FP32_PEAK_TFLOPS}
    out = {
        "metric": METRIC, "value": ns_day, "unit": "ns/day", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 pair arithmetic, int64 fixed-point accumulation, fp64 bonded/SASA/integrator",
        "data": "synthetic: seeded random-sequence chains from the NeRF builder, CHARMM36m parameters",
        "config": {"workload": workload, "config_index": args.config, "systems": n_sys,
                   "atoms": info["atoms"], "pairs_per_pass": pairs,
                   "sasa_work_atoms": info["sasa_work_atoms"] if include_sasa else 0,
                   "dt_fs": dt_fs, "sasa_gamma_scale": gamma_scale if include_sasa else 0.0,
                   "surrogate": False},
        "pair_interactions_per_s": pair_ips,
        "kernel_ms_per_step": {k: v[0] / max(args.profile_steps, 1) for k, v in prof.items() if v[1] > 0},
        "kernel_timing": "HIP events on the engine stream over a separate %d-step run with direct "
                         "launches, one kernel at a time (the timed run has SASA on a second "
                         "stream beside the pair kernels)" % args.profile_steps,
        "autotune": tune,
        "step_mode": args.step_mode,
        "relax": {"minimiser": "og_minimize (device-resident batched L-BFGS, all pairs)",
                  "max_steps": args.minimize_steps, "seconds": relax_s,
                  "steps_max_over_systems": max(r.steps for r in res)},
        "pair_kernels": pk,
        "roofline": roof,
        "dominant_kernel": dominant,
        "diverged_systems": int(st["diverged"].sum()),
        "shake_failures": int(st["shake_failures"].sum()),
    }
    if include_sasa:
        sasa_ms = k_ms.get("sasa")
        if sasa_ms:
            out["sasa_atoms_per_s"] = info["sasa_work_atoms"] / (sasa_ms * 1e-3)
    if not args.no_surrogate and args.config == 4:
        out["surrogate"] = surrogate_leg(eng, systems, temps, args, dt_fs, gamma_scale, include_sasa,
                                         ns_day, pairs)
    if not args.no_fixtures:
        out["reference_systems"] = fixtures_leg(E)
    if not args.no_cpu_baseline:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(systems, temps, dt_fs, gamma_scale)
        out["cpu_baseline"]["host_cpu"] = os.uname().machine + " x%d visible" % (os.cpu_count() or 0)
        log("cpu baseline %.1f s" % (time.time() - t0))
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
