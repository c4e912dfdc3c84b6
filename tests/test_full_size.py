"""The benchmark configuration itself (N=262144 +-1 ions, rho=0.3, kmax=16),
in both forms.

Checked against the oracle where it finishes in seconds on the host's cores,
and through properties that do not depend on size:

* ghosts: the reference's image rule (exchange_ghost_single_md_domain.cpp:
  67-215: strict tests against [lower + ghost_cutoff, upper - ghost_cutoff],
  images per owned atom in the order x, y, z, xy, xz, yz, xyz) restated in
  numpy; positions and sources bit-exact;
* Verlet list: exactly the set of (owned i, owned j>i) and (owned i, ghost)
  pairs with dr*dr < cutoff_extra^2, dr*dr rounded as the reference rounds it
  ((dx*dx + dy*dy) + dz*dz, no FMA), candidates from a k-d tree over the
  numpy ghosts at a radius 1e-9 larger; and the oracle's own list: bit-exact;
* FP64 accelerations of 1024 sampled atoms (LJ + Ewald_r + Ewald_k) and all
  32768 S(k) values against the oracle: 1e-12 relative (norm-wise);
* Newton's third law: LJ and Ewald_r act pairwise and the truncated Ewald_k
  set is symmetric under k -> -k, so the forces sum to zero; the bar is
  relative to sum |F| (1e-10 FP64, 1e-6 mixed);
* momentum: conserved by the velocity-Verlet steps to the same bar;
* the mixed form: force RMS relative error against FP64 at most 1e-5 on three
  seeded snapshots.
"""
import os

import numpy as np
import pytest

import fixtures as F
from caviar_amd import Engine, configs
from oracle import Oracle

pytestmark = pytest.mark.gpu

N, KMAX = 262144, 16


@pytest.fixture(scope="module")
def system():
    return configs.ions(n=N, rho=0.3, kmax=KMAX, seed=1)


def sorted_pairs(p):
    p = np.asarray(p, dtype=np.int64).reshape(-1, 2)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_full_size_fp64_against_oracle(system):
    """Ghosts (numpy rule and oracle), the list (k-d tree and oracle), S(k) and
    the accelerations of 1024 sampled atoms against the oracle."""
    print(f"[case] full size N={N} kmax={KMAX} fp64 vs oracle", flush=True)
    e = Engine(system.copy(), precision="fp64")
    e.setup()
    pos = e.get("pos")
    gpos, gsrc = e.ghosts()
    npos, nsrc = F.numpy_ghosts(system, pos)
    assert np.array_equal(gsrc, nsrc)
    assert np.array_equal(gpos, npos)
    rce = system.neighbor.cutoff * system.neighbor.cutoff_extra_coef
    got = sorted_pairs(e.pairs())
    want = sorted_pairs(F.exact_pairs(pos, npos, rce))
    assert len(got) == len(want)
    assert np.array_equal(got, want)
    idx = np.sort(np.random.default_rng(7).choice(N, size=1024, replace=False))
    threads = max(1, min(16, os.cpu_count() or 1))
    o = Oracle(system.copy(), threads=threads)
    print(f"[case] oracle setup over {threads} threads", flush=True)
    o.setup_sampled(idx)
    ogp, ogs = o.ghost_data()
    assert np.array_equal(ogp, gpos) and np.array_equal(ogs, gsrc)
    assert np.array_equal(sorted_pairs(o.pairs()), got)
    a = e.get("acc")[idx]
    b = o.get("acc")[idx]
    assert rel(a, b) <= 1e-12
    sk, osk = e.sk(), o.sk()
    assert len(sk) == (2 * KMAX) ** 3
    assert rel(np.c_[sk.real, sk.imag], np.c_[osk.real, osk.imag]) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_size_newton(system, precision):
    print(f"[case] full size N={N} kmax={KMAX} {precision} Newton", flush=True)
    e = Engine(system.copy(), precision=precision)
    e.setup()
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


def test_full_size_mixed_force_error():
    """Mixed vs FP64 accelerations on identical positions (three seeded
    snapshots of the workload): RMS relative error at most 1e-5."""
    for seed in (1, 2, 3):
        print(f"[case] full size mixed vs fp64 seed={seed}", flush=True)
        s = configs.ions(n=N, rho=0.3, kmax=KMAX, seed=seed)
        out = {}
        for prec in ("fp64", "mixed"):
            e = Engine(s.copy(), precision=prec)
            e.setup()
            out[prec] = e.get("acc")
            del e
        a, b = out["mixed"], out["fp64"]
        rms = np.sqrt(((a - b) ** 2).sum(1).mean() / (b ** 2).sum(1).mean())
        print(f"[case]   rms {rms:.3e}", flush=True)
        assert rms <= 1e-5
