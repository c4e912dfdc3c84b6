"""Load the golden fixtures written by tests/golden/make_golden.py."""
from __future__ import annotations

import glob
import json
import os

import numpy as np

from caviar_amd.system import (BerendsenSettings, ELCSettings, EwaldKSettings, EwaldRSettings,
                               LJListSettings, LJSettings, NeighborSettings, NVESettings,
                               ShortRangeSettings, System, ThermoSettings)

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

    def arrs(dd):
        return {k: (np.array(v) if isinstance(v, list) else v) for k, v in dd.items()}
    ljl = LJListSettings(**arrs(d["lj_list"])) if "lj_list" in d else None
    sr = None
    if "short_range" in d:
        sd = arrs(d["short_range"])
        sd["external_field"] = tuple(np.asarray(sd["external_field"]).tolist())
        sr = ShortRangeSettings(**sd)
    ref = {k[4:]: z[k] for k in z.files if k.startswith("ref_")}
    vel = z["in_vel"]
    if "random_velocity" in d:
        # the run starts from the velocities the reference's add_random_velocity
        # made (dumped by a setup of the same script; a setup fixture's own)
        vel = ref["vel_rv"] if "vel_rv" in ref else ref["vel"]
    s = System(lower=d["lower"], upper=d["upper"], bc=d["bc"], ghost_cutoff=d["ghost_cutoff"],
               mass=d["mass"], charge=d["charge"], pos=z["in_pos"], vel=vel,
               types=z["in_types"], mol=z["in_mol"] if "in_mol" in z.files else None,
               neighbor=nl, lj=lj, lj_list=ljl, short_range=sr, ewald_r=er, ewald_k=ek,
               elc=ELCSettings(**d["elc"]) if "elc" in d else None,
               thermo=ThermoSettings(**d["thermo"]) if "thermo" in d else None,
               berendsen=BerendsenSettings(**d["berendsen"]) if "berendsen" in d else None,
               nve=NVESettings(**d["nve"]) if "nve" in d else None,
               random_velocity=tuple(d["random_velocity"]) if "random_velocity" in d else None)
    if "writer" in d:
        s.meta["writer"] = d["writer"]
    meta = {"mode": str(z["mode"]), "run_steps": int(z["run_steps"]), "dt": float(z["dt"]),
            # large fixtures store the reference's list sorted, not in its order
            "pairs_sorted": bool(int(z["pairs_sorted"])) if "pairs_sorted" in z.files else False}
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


def numpy_ghosts(s, pos):
    """Ghost images of the owned positions ``pos`` of system ``s`` by the
    reference's rule (exchange_ghost_single_md_domain.cpp:67-215): strict tests
    against [lower + ghost_cutoff, upper - ghost_cutoff] times the boundary
    flag, images per owned atom in the order x, y, z, xy, xz, yz, xyz, each an
    integer shift times the box length subtracted.  Returns (positions, source
    owned index), ghost g being atom n + g of the Verlet list."""
    lo, hi, L, gc = s.lower, s.upper, s.size, s.ghost_cutoff
    side = np.zeros((len(pos), 3), dtype=np.int64)
    for d in range(3):
        v = np.where(pos[:, d] < lo[d] + gc, -1, np.where(pos[:, d] > hi[d] - gc, 1, 0))
        side[:, d] = v * int(s.bc[d])
    combos = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1)]
    srcs, ranks, imgs = [], [], []
    for rank, c in enumerate(combos):
        mask = np.ones(len(pos), dtype=bool)
        for d in range(3):
            if c[d]:
                mask &= side[:, d] != 0
        src = np.nonzero(mask)[0]
        img = pos[src].copy()
        for d in range(3):
            if c[d]:
                img[:, d] = img[:, d] - side[src, d] * L[d]
        srcs.append(src)
        ranks.append(np.full(len(src), rank))
        imgs.append(img)
    src = np.concatenate(srcs)
    rank = np.concatenate(ranks)
    img = np.concatenate(imgs).reshape(-1, 3)
    order = np.lexsort((rank, src))  # per owned atom, images in the combo order
    return img[order], src[order].astype(np.int32)


def exact_pairs(pos, gpos, rce):
    """The reference's Verlet-list membership as a set: (owned i, owned j > i)
    and (owned i, ghost n + g) with dr*dr < rce^2, dr*dr rounded as the
    reference rounds it ((dx*dx + dy*dy) + dz*dz, no FMA).  Candidates come from
    a k-d tree at a radius 1e-9 larger; the exact test is numpy's IEEE
    arithmetic."""
    from scipy.spatial import cKDTree
    n = len(pos)
    allp = np.concatenate([pos, gpos])
    cand = cKDTree(allp).query_pairs(rce * (1 + 1e-9), output_type="ndarray")
    if len(cand) == 0:
        return np.zeros((0, 2), dtype=np.int64)
    i, j = np.minimum(cand[:, 0], cand[:, 1]), np.maximum(cand[:, 0], cand[:, 1])
    keep = i < n                         # ghost-ghost pairs are not listed
    i, j = i[keep], j[keep]
    d = allp[i] - allp[j]
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    inside = r2 < rce * rce
    return np.stack([i[inside], j[inside]], axis=1)
