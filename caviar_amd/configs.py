"""Seeded synthetic workloads for the short-range + Ewald path.

Each builder returns a :class:`caviar_amd.System` for one of the workloads the
path is measured on, with the reference's object settings:

* ``lj_fcc``      - plain LJ fluid, N=4000 fcc at rho*=0.8442, T*=0.72
                    (Lj_cell_list, cell-built Verlet list, rc=2.5, skin 0.3);
* ``ions``        - charged LJ + Ewald: +-1 ions, uniform random positions with
                    minimum separation 0.8, rho=0.3, eps=sigma=1, rc_LJ=2.5,
                    Bjerrum length 2 (k_electrostatic=2), alpha=0.9,
                    neighbour cutoff 3.5, kmax per axis;
* ``b2_ewald``    - the reference's 100-ion Ewald known-answer input
                    (benchmark/b2-ewald), positions given by the caller.

The reference's own random distributor (unique/distribution.cpp:296-342) never
terminates, so positions come from seeded random sequential addition here.
"""
from __future__ import annotations

import numpy as np

from .system import (EwaldKSettings, EwaldRSettings, LJSettings, NeighborSettings,
                     System)


def maxwell_boltzmann(n: int, temperature: float, mass: np.ndarray, rng) -> np.ndarray:
    """Velocities ~ N(0, T/m) per component with the centre-of-mass velocity removed."""
    m = np.asarray(mass, dtype=np.float64).reshape(-1)
    v = rng.standard_normal((n, 3)) * np.sqrt(temperature / m)[:, None]
    p = (v * m[:, None]).sum(axis=0) / m.sum()
    return v - p[None, :]


def fcc_lattice(ncell: int, box: float) -> np.ndarray:
    a = box / ncell
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) * a
    g = np.arange(ncell) * a
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    return (cells[:, None, :] + base[None, :, :]).reshape(-1, 3)


def random_nonoverlap(n: int, box: float, dmin: float, seed: int, lower: float = 0.0) -> np.ndarray:
    """Random sequential addition in a periodic cube with minimum image
    separation ``dmin`` (batched, cKDTree-checked, deterministic for a seed)."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    acc = np.empty((0, 3))
    while len(acc) < n:
        need = n - len(acc)
        cand = rng.random((max(64, int(need * 1.3)), 3)) * box
        if len(acc):
            d, _ = cKDTree(acc, boxsize=box).query(cand, k=1)
            cand = cand[d >= dmin]
        if len(cand) > 1:
            pairs = cKDTree(cand, boxsize=box).query_pairs(dmin, output_type="ndarray")
            keep = np.ones(len(cand), dtype=bool)
            # drop the later point of every close pair (deterministic)
            for a, b in sorted(map(tuple, np.sort(pairs, axis=1))):
                if keep[a] and keep[b]:
                    keep[b] = False
            cand = cand[keep]
        acc = np.concatenate([acc, cand[:need]])
    # keep points strictly inside [0, box) and shift to the box origin
    acc = np.mod(acc, box)
    return acc + lower


def lj_fcc(ncell: int = 10, rho: float = 0.8442, temperature: float = 0.72, seed: int = 1,
           rc: float = 2.5, skin: float = 0.3, jitter: float = 0.0) -> System:
    """fcc LJ fluid; ``jitter`` > 0 displaces every site by N(0, jitter^2) per
    component (a perfect lattice has zero net force on every atom)."""
    n = 4 * ncell ** 3
    box = (n / rho) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    pos = fcc_lattice(ncell, box) + 0.5 * box / ncell * 0.25  # off the box faces
    if jitter > 0:
        pos = pos + np.random.default_rng(seed + 500).normal(scale=jitter, size=pos.shape)
    pos = np.mod(pos, box)
    types = np.zeros(n, dtype=np.int32)
    mass = np.array([1.0])
    vel = maxwell_boltzmann(n, temperature, mass[types], rng)
    coef = (rc + skin) / rc
    return System(lower=[0, 0, 0], upper=[box] * 3, bc=[1, 1, 1], ghost_cutoff=rc + skin + 0.1,
                  mass=mass, charge=[0.0], pos=pos, vel=vel, types=types,
                  neighbor=NeighborSettings(cutoff=rc, cutoff_extra_coef=coef),
                  lj=LJSettings(cutoff=rc, epsilon=np.ones((1, 1)), sigma=np.ones((1, 1))),
                  meta={"workload": "lj_fcc", "n": n, "rho": rho, "T": temperature, "seed": seed,
                        "dt": 0.005})


def ions(n: int = 8000, rho: float = 0.3, kmax: int = 8, seed: int = 1, alpha: float = 0.9,
         nl_cutoff: float = 3.5, rc_lj: float = 2.5, ke: float = 2.0, dmin: float = 0.8,
         temperature: float = 1.0, coef: float = 1.12347, with_lj: bool = True,
         with_r: bool = True, with_k: bool = True) -> System:
    """+-1 ions (types 0:+1, 1:-1; first half +1) in a periodic cube."""
    box = (n / rho) ** (1.0 / 3.0)
    pos = random_nonoverlap(n, box, dmin, seed)
    rng = np.random.default_rng(seed + 1000)
    types = np.zeros(n, dtype=np.int32)
    types[n // 2:] = 1
    mass = np.array([1.0, 1.0])
    vel = maxwell_boltzmann(n, temperature, mass[types], rng)
    rce = nl_cutoff * coef
    return System(lower=[0, 0, 0], upper=[box] * 3, bc=[1, 1, 1], ghost_cutoff=rce + 0.1,
                  mass=mass, charge=[1.0, -1.0], pos=pos, vel=vel, types=types,
                  neighbor=NeighborSettings(cutoff=nl_cutoff, cutoff_extra_coef=coef),
                  lj=(LJSettings(cutoff=rc_lj, epsilon=np.ones((2, 2)), sigma=np.ones((2, 2)))
                      if with_lj else None),
                  ewald_r=EwaldRSettings(alpha=alpha, k_electrostatic=ke) if with_r else None,
                  ewald_k=EwaldKSettings.cube(alpha, kmax, ke) if with_k else None,
                  meta={"workload": "ions", "n": n, "rho": rho, "kmax": kmax, "seed": seed,
                        "dt": 0.002})


def b2_ewald(pos: np.ndarray, charge_sign: np.ndarray, kmax: int = 25) -> System:
    """The reference's 100-ion Ewald known-answer input (benchmark/b2-ewald:
    L=10 periodic, alpha=1.112583061, real-space cutoff 4.9, k_electrostatic 1)."""
    types = np.where(np.asarray(charge_sign) < 0, 0, 1).astype(np.int32)
    n = len(pos)
    alpha = 1.112583061
    nl = NeighborSettings(cutoff=4.9)
    return System(lower=[0, 0, 0], upper=[10, 10, 10], bc=[1, 1, 1], ghost_cutoff=6.0,
                  mass=[1.0, 1.0], charge=[-1.0, 1.0], pos=pos, vel=np.zeros((n, 3)),
                  types=types, neighbor=nl, ewald_r=EwaldRSettings(alpha=alpha),
                  ewald_k=EwaldKSettings.cube(alpha, kmax),
                  meta={"workload": "b2_ewald", "n": n})


def half_charged(n: int, rho: float, kmax: int = 6, seed: int = 1, alpha: float = 0.9,
                 nl_cutoff: float = 3.5, rc_lj: float = 2.5, ke: float = 2.0, dmin: float = 0.8,
                 temperature: float = 1.0, coef: float = 1.12347) -> System:
    """One replica of the batched workload: half the atoms charged (+1 / -1 in
    equal numbers, types 0 and 1), half neutral (type 2), LJ eps = sigma = 1
    between every type pair, uniform random positions with minimum separation
    ``dmin`` in a periodic cube at number density ``rho``."""
    box = (n / rho) ** (1.0 / 3.0)
    pos = random_nonoverlap(n, box, dmin, seed)
    rng = np.random.default_rng(seed + 1000)
    types = np.full(n, 2, dtype=np.int32)
    nq = (n // 2) // 2
    types[:nq] = 0
    types[nq:2 * nq] = 1
    types = types[rng.permutation(n)]
    mass = np.array([1.0, 1.0, 1.0])
    vel = maxwell_boltzmann(n, temperature, mass[types], rng)
    rce = nl_cutoff * coef
    return System(lower=[0, 0, 0], upper=[box] * 3, bc=[1, 1, 1], ghost_cutoff=rce + 0.1,
                  mass=mass, charge=[1.0, -1.0, 0.0], pos=pos, vel=vel, types=types,
                  neighbor=NeighborSettings(cutoff=nl_cutoff, cutoff_extra_coef=coef),
                  lj=LJSettings(cutoff=rc_lj, epsilon=np.ones((3, 3)), sigma=np.ones((3, 3))),
                  ewald_r=EwaldRSettings(alpha=alpha, k_electrostatic=ke),
                  ewald_k=EwaldKSettings.cube(alpha, kmax, ke),
                  meta={"workload": "half_charged", "n": n, "rho": rho, "kmax": kmax, "seed": seed,
                        "dt": 0.002})


def replica_set(count: int = 512, n_range=(200, 4000), rho_range=(0.2, 0.8), kmax: int = 6,
                seed: int = 1) -> list:
    """The batched-replica workload: ``count`` half-charged systems with N and
    density drawn uniformly from the ranges (seeded)."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(n_range[0], n_range[1] + 1, size=count)
    rhos = rng.uniform(rho_range[0], rho_range[1], size=count)
    return [half_charged(int(n), float(r), kmax=kmax, seed=seed * 100003 + i)
            for i, (n, r) in enumerate(zip(ns, rhos))]
