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

# MI355X peaks (vendor figures, dense): matrix cores by operand type, vector ALU
PEAK = {"fp64_mfma": 78.6, "fp32_mfma": 157.3, "fp16_mfma": 2500.0,
        "fp32_vector": 157.3, "fp64_vector": 78.6}          # TFLOP/s
HBM_PEAK_GBS = 8000.0
# Algorithmic FLOP convention per (k-vector, atom) term: S pass 16, force pass 22;
# per Ewald_r list pair 43, LJ pair +12, integration 20 per atom.
FLOP_SK, FLOP_KF, FLOP_PAIR, FLOP_LJ, FLOP_INT = 16, 22, 43, 12, 20
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_ions_N262144_kmax16.json")
TRAFFIC_SOURCE = ("read from profiles/pmc_traffic_ions_N262144_kmax16.json (rocprofv3 FETCH_SIZE and "
                  "WRITE_SIZE of the same launch at the same tiles, scaled by a copy of known size); "
                  "not measured in this run")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_engine(system, precision, warmup, steps, dt, pair_kernel="auto", pin=None):
    from caviar_amd import Engine
    e = Engine(system, precision=precision, timing=False)
    e.set_pair_kernel(pair_kernel)
    t0 = time.perf_counter()
    e.setup()
    log(f"{precision}: setup {time.perf_counter() - t0:.2f}s")
    t0 = time.perf_counter()
    if pin:
        tiles = e.set_tiles(*pin)
        log(f"{precision}: pinned Ewald_k tiles {tiles}")
    else:
        tiles = e.autotune()
        log(f"{precision}: autotuned Ewald_k tiles {tiles} in {time.perf_counter() - t0:.2f}s")
    e.step(dt, warmup)
    t0 = time.perf_counter()
    e.step(dt, steps)          # every step ends with a stream synchronisation
    t = time.perf_counter() - t0
    # per-kernel launch times from HIP events, in a separate pass so that the
    # event records do not weigh on the throughput above; Ewald_k runs in line
    # here (not on its concurrent stream) so each time is the kernel's own
    e.set_overlap(False)
    e.enable_timing(True)
    e.step(dt, min(steps, 10))
    timings = {k: e.kernel_time(k) for k in ("pair", "ewald_k_sk", "ewald_k_force",
                                              "ewald_k_tables", "neighbor", "list", "integrate")}
    e.enable_timing(False)
    e.set_overlap(True)
    log(f"{precision}: {steps} steps in {t:.3f}s -> {system.n * steps / t:.4e} atom-steps/s; "
        + ", ".join(f"{k} {v[0]:.3f}ms" for k, v in timings.items()))
    timings["tiles"] = tiles
    timings["pair_kernel"] = e.pair_kernel()
    return e, t, timings


def force_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum(axis=1).mean() / (b ** 2).sum(axis=1).mean()))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(system, threads):
    """Oracle (FP64 C restatement of the reference path) on the host cores, on a
    bounded sample of the same workload, extrapolated to one full step, on
    ``threads`` threads and on one.  The LJ + Ewald_r part (about 0.1% of the
    step) is timed over the whole system on one thread; Ewald_k, parallel over
    k in the S pass and over atoms in the force pass (no atomics), on samples:
    S(k) for ks k-vectors over all atoms and the k-space force of na atoms over
    all k.  On one core the oracle runs 1.43x faster than the reference itself
    (profiles/cpu_anchor_ions8000_kmax8.md)."""
    from oracle import Oracle
    n = system.n
    pair_sys = system.copy()
    pair_sys.ewald_k = None
    o = Oracle(pair_sys, threads=1)
    t0 = time.perf_counter()
    o.setup()                       # wrap, ghosts, cell + Verlet list, one force evaluation
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.forces()
    t_pair = time.perf_counter() - t0
    nk = (2 * system.ewald_k.kx_max) * (2 * system.ewald_k.ky_max) * (2 * system.ewald_k.kz_max)
    res = {}
    for th, ks, na in ((threads, 16 * 64, 16 * 512), (1, 64, 512)):
        ko = Oracle(system, threads=th)
        ts, tf = ko.ewald_k_sample(0, ks, 0, na)
        t_step = t_pair + ts * nk / ks + tf * n / na
        res[th] = (t_step, ts, tf, ks, na)
        log(f"cpu {th} threads: pair {t_pair:.2f}s (setup {t_setup:.2f}s), S(k) {ts:.2f}s/{ks}k, "
            f"force {tf:.2f}s/{na} atoms -> {t_step:.1f}s per step")
    t_all, ts, tf, ks, na = res[threads]
    t_one, ts1, tf1, ks1, na1 = res[1]
    return {"value": n / t_all, "unit": "atom-steps/s", "cores": threads, "kind": "port",
            "value_1_thread": n / t_one,
            "cpu": cpu_model(), "nproc": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "seconds_per_step": {"all_threads": round(t_all, 2), "one_thread": round(t_one, 2)},
            "sample": (f"one LJ+Ewald_r force evaluation over all {n} atoms (1 thread, "
                       f"{t_pair:.2f}s) + Ewald_k S(k) for {ks} of {nk} k-vectors over all atoms "
                       f"({ts:.2f}s) and the k-space force of {na} of {n} atoms over all k "
                       f"({tf:.2f}s) on {threads} threads; on 1 thread {ks1} k-vectors "
                       f"({ts1:.2f}s) and {na1} atoms ({tf1:.2f}s); each scaled to one full step")}



def copy_probe(mib):
    """hipMemcpy device-to-device of ``mib`` MiB through the HIP runtime (its copy
    kernel reads and writes exactly that many bytes; larger than the caches)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    nbytes = ctypes.c_size_t(mib << 20)
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, nbytes)
    hip.hipDeviceSynchronize()
    for _ in range(3):
        if hip.hipMemcpy(b, a, nbytes, 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: 3 x {mib} MiB device-to-device")


def lattice(kmax):
    nz = 2 * kmax + 1
    rpad = -(-((kmax + 1) * (2 * kmax + 1)) // 16) * 16
    sc = -(-(2 * nz) // 16) * 16
    # split S pass: the matrix cores take the nz != 0 columns (nz = 0 is summed on the VALU)
    sch = max(16, -(-(2 * (nz - 1)) // 16) * 16)
    kz = -(-(2 * nz) // 4) * 4
    nkh = max(1, -(-(2 * (nz - 1)) // 32))  # split force pass: z = 0 seeds the accumulators
    return rpad, sc, sch, kz, nkh


def separable_flop(kmax):
    """FLOP per atom of the separable-lattice algorithm the Ewald_k kernels run
    (exp(ik.r) = X(nx) Y(ny) Z(nz) over the half lattice nx >= 0): S pass, one
    complex product q X Y per row (6) and one complex multiply-add per (row,
    column) (8); force pass, two complex multiply-adds per (row, column) (U and
    nz U, 16) and an epilogue per row (P = X Y and Im(P U) times the k
    components, 22)."""
    rows, cols = (kmax + 1) * (2 * kmax + 1), 2 * kmax + 1
    return 6 * rows + 8 * rows * cols, 16 * rows * cols + 22 * rows


def kernel_rooflines(n, kmax, ppa, lj_ppa, tm32, tm64):
    """Per kernel family: launch time, the bound, and three rates of the MI355X
    unit it runs on: ``algorithmic`` (FLOP of the separable-lattice algorithm
    the Ewald_k kernels run, separable_flop; the pair kernel's per-pair count)
    with ``frac`` of the unit's peak; ``survey_rate`` (the per-(k-vector, atom)
    convention of 16 + 22 FLOP, which counts a sincos per term the kernels never
    evaluate and can therefore exceed the peak; no fraction); ``executed`` (MFMA
    FLOP the kernel issues, from its tile shapes: the split form's three
    products per term and the row padding; rocprofv3's
    SQ_INSTS_VALU_MFMA_MOPS_* x 512 gives the same totals)."""
    nk = (2 * kmax) ** 3
    rpad, sc, sch, kz, nkh = lattice(kmax)
    sep_sk, sep_kf = separable_flop(kmax)
    traffic = {}
    if os.path.exists(TRAFFIC_FILE):
        traffic = json.load(open(TRAFFIC_FILE)).get("bytes_per_launch", {})
    out = {}
    for form, tm in (("mixed", tm32), ("fp64", tm64)):
        split = form == "mixed"
        mfma = "fp16_mfma" if split else "fp64_mfma"
        exec_kf = (4 * rpad * 32 * nkh * 2 * 3 if split else 2 * 4 * rpad * kz) * n
        exec_sk = (2 * rpad * sch * 2 * 3 if split else 2 * rpad * sc * 2) * n
        rows = {
            "ewald_k_force": ("mfma", mfma, sep_kf * n, FLOP_KF * nk * n, exec_kf),
            "ewald_k_sk": ("mfma", mfma, sep_sk * n, FLOP_SK * nk * n, exec_sk),
            "pair": ("gather", "fp32_vector" if split else "fp64_vector",
                     n * (FLOP_PAIR * ppa + FLOP_LJ * lj_ppa), None, None),
        }
        fo = {}
        for k, (bound, unit, alg, survey, ex) in rows.items():
            ms = tm[k][0]
            ach = alg / (ms * 1e-3) / 1e12
            d = {"ms": round(ms, 4), "bound": bound, "unit": "TFLOP/s", "peak": PEAK[unit],
                 "peak_unit": unit, "achieved": round(ach, 2), "frac": round(ach / PEAK[unit], 4)}
            if survey is not None:
                d["survey_rate"] = round(survey / (ms * 1e-3) / 1e12, 2)
            if ex is not None:
                d["executed"] = round(ex / (ms * 1e-3) / 1e12, 2)
                d["executed_frac"] = round(ex / (ms * 1e-3) / 1e12 / PEAK[unit], 4)
            tb = traffic.get(f"{form}:{k}")
            if tb:
                d["traffic"] = tb
                d["traffic_GBs"] = round(tb / (ms * 1e-3) / 1e9, 1)
                d["traffic_source"] = TRAFFIC_SOURCE
            fo[k] = d
        for k in ("list", "neighbor", "ewald_k_tables", "integrate"):
            fo[k] = {"ms": round(tm[k][0], 4)}
        fo["list"]["bound"] = "fp64 compare + ballot, latency"
        fo["neighbor"]["bound"] = "hbm / launch"
        fo["ewald_k_tables"]["bound"] = "hbm write"
        fo["integrate"]["bound"] = "hbm"
        out[form] = fo
    return out


def run_batched(args):
    """The batched-replica workload: replicas with N ~ U[200, 4000] and density
    ~ U[0.2, 0.8], half +-1 / half neutral, kmax 6, mixed precision, one graph
    launch per step, the force-matching adjoint every ``adjoint_every`` steps
    (residual weights synthetic, seeded)."""
    from caviar_amd import Batch, configs
    t0 = time.perf_counter()
    systems = configs.replica_set(args.replicas, (200, 4000), (0.2, 0.8), kmax=6, seed=1)
    natoms = sum(s.n for s in systems)
    log(f"batched: {len(systems)} replicas, {natoms} atoms, built in {time.perf_counter() - t0:.1f}s")
    b = Batch(systems, precision="mixed")
    b.setup()
    tiles = b.autotune()
    rng = np.random.default_rng(7)
    weights = [rng.normal(size=(s.n, 3)) * 1e-3 for s in systems]
    b.step(args.dt, 5)
    steps, every = args.batch_steps, args.adjoint_every
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        k = min(every, steps - done)
        b.step(args.dt, k)
        done += k
        if every > 0 and done % every == 0:
            b.adjoint(weights)
    t = time.perf_counter() - t0
    # the same without the adjoint, for its share
    t1 = time.perf_counter()
    b.step(args.dt, 100)
    t_plain = (time.perf_counter() - t1) / 100
    res = {"metric": "atom-steps/s, batched replicas (mixed, adjoint every %d steps)" % every,
           "value": round(natoms * steps / t, 1), "unit": "atom-steps/s",
           "replicas": len(systems), "atoms": natoms, "steps": steps, "adjoint_every": every,
           "ms_per_step": round(1e3 * t / steps, 4),
           "ms_per_step_without_adjoint": round(1e3 * t_plain, 4),
           "graph_launch_per_step": True, "tiles": tiles,
           "config": {"workload": "half-charged LJ + Ewald replicas", "n_range": [200, 4000],
                      "rho_range": [0.2, 0.8], "kmax": 6, "k_vectors_per_replica": 1728}}
    log(f"batched: {res['value']:.4e} atom-steps/s, {res['ms_per_step']} ms/step "
        f"({res['ms_per_step_without_adjoint']} ms without the adjoint), tiles {tiles}")
    b.close()
    return res


def nve_drift(system, precision, steps, every, dt, equilibrate=0):
    """Total energy (kinetic + shifted-force LJ + Ewald_r + Ewald_k) every
    ``every`` steps of an NVE run, after ``equilibrate`` unsampled steps from
    the random start: relative drift and fluctuation, and the sampled series."""
    from caviar_amd import Engine
    e = Engine(system.copy(), precision=precision)
    e.setup()
    if equilibrate:
        e.step(dt, equilibrate)
    m = system.mass[system.types][:, None]
    t, E = [], []
    t0 = time.perf_counter()
    for s in range(0, steps + 1, every):
        if s:
            e.step(dt, every)
        v = e.get("vel")
        ke = 0.5 * float((m * v * v).sum())
        elj, er, ek = e.energy()
        t.append(s * dt)
        E.append(ke + elj + er + ek)
    E = np.array(E)
    slope = np.polyfit(np.array(t), E, 1)[0]
    res = {"steps": steps, "dt": dt, "equilibrate": equilibrate, "samples": len(E), "E0": float(E[0]),
           "rel_drift_end": float((E[-1] - E[0]) / abs(E[0])),
           "rel_drift_per_1000_steps": float(slope * 1000 * dt / abs(E[0])),
           "rel_fluctuation_rms": float(np.std(E - np.polyval(np.polyfit(t, E, 1), t)) / abs(E[0])),
           "seconds": round(time.perf_counter() - t0, 1),
           "E_rel": [float(v) for v in (E - E[0]) / abs(E[0])]}
    log(f"nve {precision}: " + str({k: v for k, v in res.items() if k != "E_rel"}))
    return res


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
    ap.add_argument("--snapshots", type=int, default=10,
                    help="seeded snapshots for the mixed-vs-FP64 force error")
    ap.add_argument("--replicas", type=int, default=512,
                    help="batched workload: replicas (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--adjoint-every", type=int, default=10)
    ap.add_argument("--pair-kernel", choices=("auto", "rows", "cells"), default="auto",
                    help="short-range pair kernel (auto = cell-staged where it applies)")
    ap.add_argument("--nve-steps", type=int, default=0,
                    help="also measure NVE energy drift over this many steps (both forms)")
    ap.add_argument("--nve-every", type=int, default=100)
    ap.add_argument("--nve-equilibrate", type=int, default=0,
                    help="unsampled steps from the random start before the NVE measurement")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the batched-replica workload and print its JSON")
    ap.add_argument("--tiles", type=lambda v: [int(t) for t in v.split(",")], default=None,
                    metavar="S_ROWS,S_CHUNK,F_TILES,F_WAVES",
                    help="with --precision mixed|fp64: pin the Ewald_k tiles instead of "
                         "tuning them (profiling runs that must see the same launches)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device with the HIP runtime (a kernel of "
                         "known traffic, to calibrate byte counters) and exit")
    args = ap.parse_args()
    if args.tiles is not None and (len(args.tiles) != 4 or args.precision == "both"):
        ap.error("--tiles takes four integers and needs --precision mixed or fp64")
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not part of this build)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.only_batched:
        print(json.dumps(run_batched(args)))
        return
    from caviar_amd import configs
    t0 = time.perf_counter()
    system = configs.ions(n=args.n, rho=0.3, kmax=args.kmax, seed=args.seed)
    log(f"workload built in {time.perf_counter() - t0:.1f}s: N={system.n} L={system.size[0]:.3f}")

    if args.precision != "both":
        run_engine(system, args.precision, args.warmup, args.steps, args.dt, args.pair_kernel,
                   args.tiles)
        return
    # one engine at a time: each context's two HIP streams (short-range passes,
    # concurrent Ewald_k) then get hardware queues of their own
    e64, t64, tm64 = run_engine(system, "fp64", args.warmup, args.steps, args.dt, args.pair_kernel)
    ppa, lj_ppa = e64.list_stats()
    e64.close()
    e32, t32, tm32 = run_engine(system, "mixed", args.warmup, args.steps, args.dt, args.pair_kernel)
    e32.close()

    # force error of the mixed form (at the tiles timed above) against FP64 on
    # identical positions: seeded snapshots
    from caviar_amd import Engine
    e64 = Engine(system.copy(), precision="fp64")
    e64.setup()
    e32 = Engine(system.copy(), precision="mixed")
    e32.setup()
    tl = tm32["tiles"]
    e32.set_tiles(tl["s_rows"], tl["s_chunk"], tl["f_tiles"], tl["f_waves"])
    errs = []
    for k in range(args.snapshots):
        snap = configs.ions(n=args.n, rho=0.3, kmax=args.kmax, seed=1000 + k).pos
        for e in (e64, e32):
            e.set("pos", snap)
            e.recalculate(force_rebuild=True)
        errs.append(force_rms(e32.get("acc"), e64.get("acc")))
    e64.close()
    e32.close()
    log(f"force rms err mixed vs fp64 over {len(errs)} snapshots: max {max(errs):.3e}, "
        f"mean {np.mean(errs):.3e}; pairs/atom {ppa:.2f}, lj pairs/atom {lj_ppa:.2f}")

    n, K = system.n, args.steps
    value = n * K / t32
    value64 = n * K / t64
    nk = (2 * args.kmax) ** 3
    kernels = kernel_rooflines(n, args.kmax, ppa, lj_ppa, tm32, tm64)
    # dominant kernel of the mixed step
    dom = max(("ewald_k_force", "ewald_k_sk", "pair"), key=lambda k: tm32[k][0])
    kd = kernels["mixed"][dom]
    ms = kd["ms"]
    alg = (FLOP_KF if dom == "ewald_k_force" else FLOP_SK) * nk * n if dom != "pair" else None
    ach = round(alg / (ms * 1e-3) / 1e12, 2) if alg else kd["achieved"]
    roof = {"bound": kd["bound"], "achieved": ach, "peak": kd["peak"], "unit": kd["unit"],
            "frac": round(ach / kd["peak"], 4), "traffic": kd.get("traffic"), "kernel": dom,
            "ms_per_launch": ms, "separable_achieved": kd["achieved"],
            "separable_frac": kd["frac"], "executed": kd.get("executed"),
            "executed_frac": kd.get("executed_frac"), "traffic_source": TRAFFIC_SOURCE,
            "note": "achieved = 22 FLOP per (k-vector, atom) x K x N per launch (the per-term "
                    "convention) / the measured launch time; separable_achieved = FLOP of the "
                    "separable-lattice algorithm the kernel runs; executed = MFMA FLOP issued (three "
                    "fp16 products per term, row padding); peak = the fp16 matrix peak it runs on"}
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
        "dtype": "f16 MFMA x3 hi/lo split (k-space) + f32 (pair, k tables) + f64 (sums, integration)",
        "data": "synthetic",
        "config": {"workload": "charged LJ + Ewald ions, N=262144, rho=0.3, kmax=16, mixed precision",
                   "n_atoms": n, "kmax": args.kmax, "k_vectors": nk, "dt": args.dt,
                   "parallelism": "single GPU"},
        "fp64": {"value": round(value64, 1), "ms_per_step": round(1e3 * t64 / K, 4),
                 "mixed_over_fp64": round(value / value64, 3)},
        "force_rms_rel_err": max(errs),
        "force_rms_rel_err_snapshots": {"count": len(errs), "max": max(errs),
                                        "mean": float(np.mean(errs)),
                                        "seeds": [1000 + k for k in range(len(errs))]},
        "list": {"pairs_per_atom": round(ppa, 3), "lj_pairs_per_atom": round(lj_ppa, 3)},
        "algorithmic_tflops_per_step": {"mixed": round(step_flop / (t32 / K) / 1e12, 2),
                                        "fp64": round(step_flop / (t64 / K) / 1e12, 2)},
        "kernel_ms": {"mixed": {k: round(v[0], 4) for k, v in tm32.items() if isinstance(v, tuple)},
                      "fp64": {k: round(v[0], 4) for k, v in tm64.items() if isinstance(v, tuple)}},
        "kernels": kernels,
        "tiles": {"mixed": tm32["tiles"], "fp64": tm64["tiles"]},
        "pair_kernel": tm32["pair_kernel"],
        "roofline": roof,
    }
    if args.nve_steps:
        out["nve"] = {f: nve_drift(system, f, args.nve_steps, args.nve_every, args.dt,
                                   args.nve_equilibrate)
                      for f in ("mixed", "fp64")}
    if args.replicas:
        out["batched"] = run_batched(args)
    if not args.no_cpu_baseline:
        # every host core the box gives this GPU (16), as OMP_NUM_THREADS states there
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        threads = max(1, min(threads, os.cpu_count() or 1, 16))
        out["cpu_baseline"] = cpu_baseline(system, threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
