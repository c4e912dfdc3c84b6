"""Load the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from caviar_amd.system import (EwaldKSettings, EwaldRSettings, LJSettings, NeighborSettings,
                               System)

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names(pattern: str = "*") -> list:
    out = []
    for p in sorted(glob.glob(os.path.join(GOLDEN, pattern + ".npz"))):
        n = os.path.basename(p)[:-4]
        if n != "b2_published":
            out.append(n)
    return out


def load(name: str):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = json.loads(str(z["settings"]))
    nl = NeighborSettings(**d["neighbor"]) if "neighbor" in d else None
    lj = None
    if "lj" in d:
        lj = LJSettings(cutoff=d["lj"]["cutoff"], epsilon=np.array(d["lj"]["epsilon"]),
                        sigma=np.array(d["lj"]["sigma"]))
    er = None
    if "ewald_r" in d:
        e = d["ewald_r"]
        er = EwaldRSettings(alpha=e["alpha"], k_electrostatic=e["k_electrostatic"],
                            lambda_=None if e["lambda_"] is None else np.array(e["lambda_"]))
    ek = EwaldKSettings(**d["ewald_k"]) if "ewald_k" in d else None
    s = System(lower=d["lower"], upper=d["upper"], bc=d["bc"], ghost_cutoff=d["ghost_cutoff"],
               mass=d["mass"], charge=d["charge"], pos=z["in_pos"], vel=z["in_vel"],
               types=z["in_types"], neighbor=nl, lj=lj, ewald_r=er, ewald_k=ek)
    ref = {k[4:]: z[k] for k in z.files if k.startswith("ref_")}
    meta = {"mode": str(z["mode"]), "run_steps": int(z["run_steps"]), "dt": float(z["dt"])}
    return s, ref, meta


def published_b2():
    z = np.load(os.path.join(GOLDEN, "b2_published.npz"), allow_pickle=False)
    return {k: z[k] for k in z.files}


def rel_err(a, b) -> float:
    """Norm-wise relative error ||a-b|| / ||b||."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def sorted_pairs(p) -> np.ndarray:
    p = np.asarray(p, dtype=np.int64).reshape(-1, 2)
    if len(p) == 0:
        return p
    order = np.lexsort((p[:, 1], p[:, 0]))
    return p[order]
