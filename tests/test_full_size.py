"""The benchmark configuration itself (N=262144 +-1 ions, rho=0.3, kmax=16),
checked through properties that do not depend on size, in both forms.

* Verlet list: exactly the set of (owned i, owned j>i) and (owned i, ghost)
  pairs with dr*dr < cutoff_extra^2, dr*dr rounded as the reference rounds it
  ((dx*dx + dy*dy) + dz*dz, no FMA).  Candidates come from a k-d tree at a
  radius 1e-9 larger; the exact test is numpy's IEEE arithmetic.
* Newton's third law: LJ and Ewald_r act pairwise and the truncated Ewald_k
  set is symmetric under k -> -k, so the forces sum to zero; the bar is
  relative to sum |F| (1e-10 FP64, 1e-6 mixed).
* Momentum: conserved by the velocity-Verlet steps to the same bar.
"""
import numpy as np
import pytest

from caviar_amd import Engine, configs

pytestmark = pytest.mark.gpu

N, KMAX = 262144, 16


@pytest.fixture(scope="module")
def system():
    return configs.ions(n=N, rho=0.3, kmax=KMAX, seed=1)


def exact_pairs(pos, gpos, rce):
    from scipy.spatial import cKDTree
    n = len(pos)
    allp = np.concatenate([pos, gpos])
    cand = cKDTree(allp).query_pairs(rce * (1 + 1e-9), output_type="ndarray")
    i, j = np.minimum(cand[:, 0], cand[:, 1]), np.maximum(cand[:, 0], cand[:, 1])
    keep = i < n                         # ghost-ghost pairs are not listed
    i, j = i[keep], j[keep]
    d = allp[i] - allp[j]
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    inside = r2 < rce * rce
    return np.stack([i[inside], j[inside]], axis=1)


def sorted_pairs(p):
    p = np.asarray(p, dtype=np.int64).reshape(-1, 2)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_size_list_and_newton(system, precision):
    print(f"[case] full size N={N} kmax={KMAX} {precision}", flush=True)
    e = Engine(system.copy(), precision=precision)
    e.setup()
    pos = e.get("pos")
    gpos, _ = e.ghosts()
    rce = system.neighbor.cutoff * system.neighbor.cutoff_extra_coef
    got = sorted_pairs(e.pairs())
    want = sorted_pairs(exact_pairs(pos, gpos, rce))
    assert len(got) == len(want)
    assert np.array_equal(got, want)
    f = e.get("acc") * system.mass[system.types][:, None]
    tol = 1e-10 if precision == "fp64" else 1e-6
    assert np.abs(f.sum(axis=0)).max() <= tol * np.abs(f).sum(axis=0).max()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_size_momentum(system, precision):
    print(f"[case] full size momentum {precision}", flush=True)
    e = Engine(system.copy(), precision=precision)
    e.setup()
    m = system.mass[system.types][:, None]
    p0 = (m * e.get("vel")).sum(axis=0)
    scale = (m * np.abs(e.get("vel"))).sum()
    e.step(0.002, 20)
    p1 = (m * e.get("vel")).sum(axis=0)
    tol = 1e-10 if precision == "fp64" else 1e-6
    assert np.abs(p1 - p0).max() <= tol * scale
