"""Rounding of the neighbour-list cutoff test.

The reference decides list membership with dr*dr < cutoff_extra^2, where
dr*dr = (dx*dx + dy*dy) + dz*dz is rounded one operation at a time (its x86-64
build has no FMA).  A compiler that contracts the products into fused
multiply-adds rounds differently and flips pairs that lie within an ulp of the
cutoff.  The inputs here are pairs built to sit exactly there: for each,
unfused rounding and at least one fused evaluation order disagree about
membership.

CPU: the oracle's pair set equals the unfused classification of every built
pair (and the build does hold pairs that a fused evaluation would misclassify).
GPU: the HIP list equals the oracle's, bit for bit, in both precisions.
"""
from fractions import Fraction
import itertools

import numpy as np
import pytest

import fixtures as F
from caviar_amd import Engine, configs
from oracle import Oracle

CUTOFF, COEF = 3.5, 1.12347


def _round(x: Fraction) -> float:
    return float(x)  # Fraction -> float rounds to nearest, ties to even


def unfused(d):
    dx, dy, dz = (float(v) for v in d)
    return (dx * dx + dy * dy) + dz * dz


def fused_orders(d):
    """r^2 with one plain product and the other two fused, in every order."""
    out = []
    for a, b, c in itertools.permutations(range(3)):
        t = float(d[a]) * float(d[a])
        t = _round(Fraction(d[b]) * Fraction(d[b]) + Fraction(t))
        t = _round(Fraction(d[c]) * Fraction(d[c]) + Fraction(t))
        out.append(t)
    return out


def cutoff_pairs(npairs=48, seed=11):
    """Positions of `npairs` isolated pairs, each within an ulp of the list
    cutoff and on a rounding boundary, in a periodic box of side 100."""
    rce = CUTOFF * COEF
    rc2 = rce * rce
    rng = np.random.default_rng(seed)
    grid = [(x, y, z) for x in range(4) for y in range(4) for z in range(4)]
    pos, inside = [], []
    for g in grid:
        if len(inside) == npairs:
            break
        a = np.array(g, dtype=np.float64) * 22.0 + 12.0 + rng.uniform(-1, 1, 3)
        for _ in range(2000):
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
            b0 = a + rce * u
            found = False
            for k in range(-24, 25):
                b = b0.copy()
                b[2] = b0[2] + k * np.spacing(b0[2])
                d = a - b                       # the list's dr, as rounded on the device
                r2 = unfused(d)
                flips = [(f < rc2) != (r2 < rc2) for f in fused_orders(d)]
                if any(flips):
                    pos += [a, b]
                    inside.append(r2 < rc2)
                    found = True
                    break
            if found:
                break
    assert len(inside) == npairs, "not enough boundary pairs found"
    return np.array(pos), np.array(inside)


def boundary_system():
    pos, inside = cutoff_pairs()
    n = len(pos)
    s = configs.ions(n=n, rho=n / 1.0e6, kmax=2, seed=5, nl_cutoff=CUTOFF, coef=COEF,
                     with_k=False)
    s.pos = pos
    s.vel = np.zeros_like(pos)
    return s, inside


def expected_pairs(inside):
    k = np.nonzero(inside)[0]
    return np.stack([2 * k, 2 * k + 1], axis=1)


def test_oracle_follows_unfused_rounding():
    s, inside = boundary_system()
    assert inside.any() and (~inside).any()
    o = Oracle(s)
    o.setup()
    got = F.sorted_pairs(o.pairs())
    assert np.array_equal(got, F.sorted_pairs(expected_pairs(inside)))


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_list_follows_unfused_rounding(gpu, precision):
    s, inside = boundary_system()
    print(f"[case] cutoff-boundary pairs n={s.n} {precision}", flush=True)
    e = Engine(s, precision=precision)
    e.setup()
    assert np.array_equal(F.sorted_pairs(e.pairs()), F.sorted_pairs(expected_pairs(inside)))
