"""CPU anchor for the oracle's timings: the reference itself (oracle/_ref/ref_harness,
built from the reference's sources by oracle/ref.mk) and the oracle, both on
one core, on the same 8000-ion LJ + Ewald input (rho 0.3, kmax 8), setup plus
a few velocity-Verlet steps.  Both run the same O(N K) Ewald_k algorithm, so
their per-step times should agree within a small factor; the ratio says how
far the oracle's timing (the CPU baseline bench.py reports) stands from the
reference's own on this host.

Runs where the reference tree and its built harness are present (the
development container):

    python oracle/anchor.py [--steps 5]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=5)
    args = ap.parse_args()
    import make_golden
    from caviar_amd import configs
    from oracle import Oracle

    harness = os.path.join(HERE, "_ref", "ref_harness")
    if not os.path.exists(harness):
        raise SystemExit("build the reference harness first: make -f oracle/ref.mk -j8")
    s = configs.ions(n=8000, kmax=8, seed=1)
    dt = 0.002
    with tempfile.TemporaryDirectory() as td:
        make_golden.write_xyz(os.path.join(td, "in.xyz"), s)
        with open(os.path.join(td, "in.casl"), "w") as f:
            f.write(make_golden.casl_for(s, "in.xyz", args.steps, dt))
        t0 = time.perf_counter()
        r = subprocess.run(["taskset", "-c", "0", harness, "in.casl", "run", "out.txt"], cwd=td,
                           capture_output=True, text=True)
        t_ref = time.perf_counter() - t0
        if r.returncode != 0:
            raise SystemExit(r.stderr[-2000:])
    o = Oracle(s, threads=1)
    t0 = time.perf_counter()
    o.setup()
    assert o.run(dt, args.steps) == 0
    t_orc = time.perf_counter() - t0
    evals = args.steps + 1
    out = {"input": "8000 ions, rho 0.3, kmax 8 (K=4096), LJ + Ewald_r + Ewald_k, dt 0.002",
           "force_evaluations": evals, "cpu": cpu_model(), "threads": 1,
           "reference_s": round(t_ref, 3), "oracle_s": round(t_orc, 3),
           "reference_atom_steps_per_s": round(s.n * evals / t_ref, 1),
           "oracle_atom_steps_per_s": round(s.n * evals / t_orc, 1),
           "oracle_over_reference_speed": round(t_ref / t_orc, 3),
           "note": "reference time includes reading the script and the XYZ input"}
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
