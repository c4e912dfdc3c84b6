"""Batched replicas (cav_batch_*): every replica of a batch against the
single-system path (cav_ctx, itself pinned to the reference) on the same input.

Bars: Verlet pair sets bit-exact per replica; FP64 accelerations and virial
within 1e-12 relative after setup (the S(k) chunking differs from the single
path, so sums are regrouped); FP64 trajectories within 1e-9 after 10 steps;
graph-launched steps bitwise equal to the same kernels launched one by one;
the mixed form within 1e-5 force RMS of FP64.
"""
import numpy as np
import pytest

import fixtures as F
from caviar_amd import Batch, Engine, configs

pytestmark = pytest.mark.gpu


def ion_replicas():
    # uneven sizes and boxes (L from 6.3 to 12.6), one shared force field
    return [configs.ions(n=300, rho=0.3, kmax=4, seed=3),
            configs.ions(n=200, rho=0.8, kmax=4, seed=4),
            configs.ions(n=640, rho=0.45, kmax=4, seed=5),
            configs.ions(n=450, rho=0.25, kmax=4, seed=6),
            configs.ions(n=1000, rho=0.6, kmax=4, seed=7)]


def rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum(1).mean() / (b ** 2).sum(1).mean()))


def single(s, precision="fp64", steps=0, dt=0.002):
    e = Engine(s.copy(), precision=precision)
    e.setup()
    if steps:
        e.step(dt, steps)
    return e


def test_batch_setup_matches_single_systems(gpu):
    systems = ion_replicas()
    print("[case] batch setup fp64", [s.n for s in systems], flush=True)
    b = Batch(systems)
    b.setup()
    for r, s in enumerate(systems):
        e = single(s)
        assert np.array_equal(F.sorted_pairs(b.pairs(r)), F.sorted_pairs(e.pairs())), r
        assert F.rel_err(b.get(r, "acc"), e.get("acc")) <= 1e-12, r
        v = e.virial()
        assert abs(b.virial(r) - v) <= 1e-12 * abs(v), r
        assert np.array_equal(b.get(r, "pos"), e.get("pos")), r


def test_batch_trajectory_matches_single_systems(gpu):
    systems = ion_replicas()
    print("[case] batch 10 steps fp64", flush=True)
    b = Batch(systems)
    b.setup()
    b.step(0.002, 10, graph=True)
    for r, s in enumerate(systems):
        e = single(s, steps=10)
        for what in ("pos", "vel", "acc"):
            assert F.rel_err(b.get(r, what), e.get(what)) <= 1e-9, (r, what)
        assert np.array_equal(F.sorted_pairs(b.pairs(r)), F.sorted_pairs(e.pairs())), r


def test_graph_replays_equal_eager_steps(gpu):
    """The captured step replayed several times, with eager work between the
    replays, against the same kernels launched one by one: bitwise equal."""
    systems = ion_replicas()[:3]
    print("[case] batch graph vs eager", flush=True)
    ga, eb = Batch(systems), Batch(systems)
    ga.setup()
    eb.setup()
    for k in range(4):
        ga.step(0.002, 1, graph=True)
        eb.step(0.002, 1, graph=False)
        for r in range(len(systems)):
            for what in ("pos", "vel", "acc"):
                assert np.array_equal(ga.get(r, what), eb.get(r, what)), (k, r, what)


def test_batch_mixed_form_force_error(gpu):
    systems = ion_replicas()
    print("[case] batch mixed", flush=True)
    b = Batch(systems, precision="mixed")
    b.setup()
    for r, s in enumerate(systems):
        assert rms(b.get(r, "acc"), single(s).get("acc")) <= 1e-5, r


def test_half_charged_replicas(gpu):
    """Replicas of the batched workload (half +-1, half neutral, 3 types) with
    kmax 6: FP64 batch against single systems, mixed batch against FP64."""
    systems = [configs.half_charged(n, rho, kmax=6, seed=20 + i)
               for i, (n, rho) in enumerate([(200, 0.8), (900, 0.2), (350, 0.55), (1300, 0.35)])]
    print("[case] half-charged batch", flush=True)
    b64 = Batch(systems)
    b64.setup()
    b32 = Batch(systems, precision="mixed")
    b32.setup()
    for r, s in enumerate(systems):
        e = single(s)
        assert F.rel_err(b64.get(r, "acc"), e.get("acc")) <= 1e-12, r
        assert rms(b32.get(r, "acc"), e.get("acc")) <= 1e-5, r


def test_autotune_keeps_results(gpu):
    systems = ion_replicas()
    print("[case] batch autotune", flush=True)
    b = Batch(systems, precision="mixed")
    b.setup()
    tiles = b.autotune()
    assert tiles["s_rows"] in (1, 2, 3) and tiles["s_stage"] in (16, 32)
    assert tiles["s_chunk"] in (256, 512, 1024)
    for r, s in enumerate(systems):
        assert rms(b.get(r, "acc"), single(s).get("acc")) <= 1e-5, r
    b.step(0.002, 5)
    for r, s in enumerate(systems):
        assert rms(b.get(r, "acc"), single(s, steps=5).get("acc")) <= 1e-4, r


@pytest.mark.parametrize("precision, tol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_batch_adjoint_matches_single_systems(gpu, precision, tol):
    """The batched force-matching adjoint of every replica against the
    single-system adjoint (itself pinned by central differences and the FP64
    CPU reference)."""
    systems = ion_replicas()[:4]
    print(f"[case] batch adjoint {precision}", flush=True)
    b = Batch(systems, precision=precision)
    b.setup()
    rng = np.random.default_rng(2)
    weights = [rng.normal(size=(s.n, 3)) for s in systems]
    got = b.adjoint(weights)
    for r, s in enumerate(systems):
        e = Engine(s.copy(), precision="fp64")
        e.setup()
        want = e.adjoint(weights[r])
        for g, w in zip(got[r], want):
            assert np.linalg.norm(g - w) <= tol * max(np.linalg.norm(w), 1e-300), (r, g, w)
