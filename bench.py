"""Throughput of the LJ + Ewald MD step on one MI355X.

Workload: +-1 ions, N=262144, uniform random positions with minimum
separation 0.8 at rho=0.3 (L=95.603), LJ eps=sigma=1 rc=2.5, Ewald alpha=0.9,
k_electrostatic=2 (Bjerrum length 2), neighbour cutoff 3.5 (cutoff_extra_coef
1.12347), kmax=16 per axis (K=32768 as the reference truncates), velocity
Verlet dt=0.002.  One step = drift + wrap + ghosts + neighbour update +
LJ/Ewald_r pair forces + Ewald_k (S(k) and forces) + kick, all on the GPU with
inputs resident in HBM.

The headline value is mixed-precision atom-steps/s; the plain FP64 form is run
on the same input for the mixed/FP64 ratio, and the force RMS error of the
mixed form against FP64 is measured on identical positions.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X dense peaks (vendor figures)
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 157.3}   # FP64 / FP32 matrix (and vector) peaks
# Algorithmic FLOP convention per (k-vector, atom) term: S pass 16, force pass 22;
# per Ewald_r list pair 43, LJ pair +12, integration 20 per atom.
FLOP_SK, FLOP_KF, FLOP_PAIR, FLOP_LJ, FLOP_INT = 16, 22, 43, 12, 20


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_engine(system, precision, warmup, steps, dt):
    from caviar_amd import Engine
    e = Engine(system, precision=precision, timing=False)
    t0 = time.perf_counter()
    e.setup()
    log(f"{precision}: setup {time.perf_counter() - t0:.2f}s")
    e.step(dt, warmup)
    e.enable_timing(True)
    t0 = time.perf_counter()
    e.step(dt, steps)          # every step ends with a stream synchronisation
    t = time.perf_counter() - t0
    timings = {k: e.kernel_time(k) for k in ("pair", "ewald_k_sk", "ewald_k_force",
                                              "ewald_k_tables", "neighbor", "list", "integrate")}
    e.enable_timing(False)
    log(f"{precision}: {steps} steps in {t:.3f}s -> {system.n * steps / t:.4e} atom-steps/s; "
        + ", ".join(f"{k} {v[0]:.3f}ms" for k, v in timings.items()))
    return e, t, timings


def force_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum(axis=1).mean() / (b ** 2).sum(axis=1).mean()))


def cpu_baseline(system, threads):
    """Oracle (FP64 C restatement of the reference path) on the host cores, on a
    bounded sample of the same workload, extrapolated to one full step."""
    from oracle import Oracle
    n = system.n
    # pair part: LJ + Ewald_r over the full system (serial, like the reference)
    pair_sys = system.copy()
    pair_sys.ewald_k = None
    o = Oracle(pair_sys, threads=1)
    t0 = time.perf_counter()
    o.setup()                       # wrap, ghosts, cell + Verlet list, one force evaluation
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.forces()
    t_pair = time.perf_counter() - t0
    # Ewald_k sample: S(k) for ks k-vectors over all atoms, force pass for na atoms over all k
    ko = Oracle(system, threads=threads)
    nk = (2 * system.ewald_k.kx_max) * (2 * system.ewald_k.ky_max) * (2 * system.ewald_k.kz_max)
    ks, na = 256, 1024
    ts, tf = ko.ewald_k_sample(0, ks, 0, na)
    t_step = t_pair + ts * nk / ks + tf * n / na
    log(f"cpu: pair {t_pair:.2f}s (setup {t_setup:.2f}s), S(k) {ts:.2f}s/{ks}k, force {tf:.2f}s/{na} atoms "
        f"-> {t_step:.1f}s per step")
    return {"value": n / t_step, "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "sample": (f"one LJ+Ewald_r force evaluation over all {n} atoms (1 thread) + Ewald_k "
                       f"S(k) for {ks} of {nk} k-vectors over all atoms and the k-space force for "
                       f"{na} of {n} atoms over all k ({threads} threads), scaled to one full step")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=262144)
    ap.add_argument("--kmax", type=int, default=16)
    ap.add_argument("--dt", type=float, default=0.002)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", choices=("both", "mixed", "fp64"), default="both",
                    help="run one form only (for profiling); the JSON line needs both")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not part of this build)", file=sys.stderr)
        sys.exit(2)

    from caviar_amd import configs
    t0 = time.perf_counter()
    system = configs.ions(n=args.n, rho=0.3, kmax=args.kmax, seed=args.seed)
    log(f"workload built in {time.perf_counter() - t0:.1f}s: N={system.n} L={system.size[0]:.3f}")

    if args.precision != "both":
        run_engine(system, args.precision, args.warmup, args.steps, args.dt)
        return
    e64, t64, tm64 = run_engine(system, "fp64", args.warmup, args.steps, args.dt)
    e32, t32, tm32 = run_engine(system, "mixed", args.warmup, args.steps, args.dt)

    # force error of the mixed form on identical positions (initial and after the FP64 run)
    errs = []
    for pos in (system.pos, e64.get("pos")):
        for e in (e64, e32):
            e.set("pos", pos)
            e.recalculate(force_rebuild=True)
        errs.append(force_rms(e32.get("acc"), e64.get("acc")))
    ppa, lj_ppa = e64.list_stats()
    log(f"force rms err mixed vs fp64: {errs}; pairs/atom {ppa:.2f}, lj pairs/atom {lj_ppa:.2f}")

    n, K = system.n, args.steps
    value = n * K / t32
    value64 = n * K / t64
    nk = (2 * args.kmax) ** 3
    # dominant kernel of the mixed step and its algorithmic roofline
    fam = {"ewald_k_force": FLOP_KF * nk * n, "ewald_k_sk": FLOP_SK * nk * n}
    dom = max(fam, key=lambda k: tm32[k][0])
    dom_ms = tm32[dom][0]
    achieved = fam[dom] / (dom_ms * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_TFLOPS["mixed"],
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_TFLOPS["mixed"], 4), "traffic": None,
            "kernel": dom, "ms_per_launch": round(dom_ms, 4)}
    step_flop = n * (FLOP_PAIR * ppa + FLOP_LJ * lj_ppa + (FLOP_SK + FLOP_KF) * nk + FLOP_INT)

    out = {
        "metric": "atom-steps/s LJ+Ewald N=256k, mixed vs FP64 GPU at force RMS err <=1e-5",
        "value": round(value, 1),
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t32 / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32/fp64 mixed",
        "data": "synthetic",
        "config": {"workload": "charged LJ + Ewald ions, N=262144, rho=0.3, kmax=16, mixed precision",
                   "n_atoms": n, "kmax": args.kmax, "k_vectors": nk, "dt": args.dt,
                   "parallelism": "single GPU"},
        "fp64": {"value": round(value64, 1), "ms_per_step": round(1e3 * t64 / K, 4),
                 "mixed_over_fp64": round(value / value64, 3)},
        "force_rms_rel_err": max(errs),
        "algorithmic_tflops_per_step": {"mixed": round(step_flop / (t32 / K) / 1e12, 2),
                                        "fp64": round(step_flop / (t64 / K) / 1e12, 2)},
        "kernel_ms": {"mixed": {k: round(v[0], 4) for k, v in tm32.items()},
                      "fp64": {k: round(v[0], 4) for k, v in tm64.items()}},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        threads = max(1, min(threads, 16))
        out["cpu_baseline"] = cpu_baseline(system, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
