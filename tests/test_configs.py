"""Two measured workloads at their full sizes, on the GPU through the C ABI.

* Force matching, 64 frames x N=2000 (two species, eps=(1.0, 0.8),
  sigma=(1.0, 1.2), q=(+1, -1) targets; start perturbed by seeded +-10% under
  charge neutrality; kmax 6): per-frame FP64 forces at the start point against
  the oracle (1e-12), the loss (1e-12) and its gradient against the oracle's
  FP64 adjoint (1e-10), and the gradient against central differences of the
  GPU loss at h = 1e-6 (1e-6).
* Batched replicas, all 512 systems (N ~ U[200, 4000], rho ~ U[0.2, 0.8],
  half charged, kmax 6): per replica the Verlet pair set against the exact set
  (numpy ghosts by the reference's image rule, k-d tree candidates, unfused
  rounding), FP64 accelerations and virial against the oracle (1e-12),
  Newton's third law and momentum over 20 steps where they hold; 8 sampled replicas
  against single-system FP64 runs (accelerations 1e-12 after setup, 1e-9
  after 20 steps) and their batched adjoint against the single-system one at
  setup (1e-10); the mixed form with the adjoint after every 10th step within 1e-5
  force RMS of FP64.
"""
import os

import numpy as np
import pytest

import fixtures as F
from caviar_amd import Batch, Engine, configs, fitting
from oracle import Oracle

pytestmark = pytest.mark.gpu

THETA_T = np.array([1.0, 0.8, 1.0, 1.2, 1.0, -1.0])  # eps, sigma, q per species
THREADS = max(1, min(16, os.cpu_count() or 1))


def test_force_matching_64_frames_n2000(gpu):
    nf, n = 64, 2000
    print(f"[case] force matching {nf} frames x N={n}, kmax 6", flush=True)
    fs = fitting.make_frameset(nframes=nf, n=n, theta_target=THETA_T, kmax=6, seed=101)
    fm = fitting.ForceMatcher(fs)
    theta = fitting.neutral_start(THETA_T, np.array([n // 2, n - n // 2]), seed=17)
    loss, grad = fm.loss_and_grad(theta)
    assert loss > 0
    # the oracle's loss and FP64 adjoint gradient over the same frames
    em, sm = fitting.mix(theta[0:2], theta[2:4])
    scale = 2.0 / (nf * n * 3)
    tot = 0.0
    ge, gs, gq = np.zeros((2, 2)), np.zeros((2, 2)), np.zeros(2)
    for f in range(nf):
        s = fs.system.copy()
        s.lj.epsilon, s.lj.sigma = em, sm
        s.charge = theta[4:6].copy()
        s.pos = fs.frames[f]
        o = Oracle(s, threads=THREADS)
        o.setup()
        fo = o.get("acc") * s.mass[s.types][:, None]
        fg = fm.forces(theta, f)
        assert F.rel_err(fg, fo) <= 1e-12, f
        r = fo - fs.targets[f]
        tot += float(np.sum(r * r))
        a, b, c = o.adjoint(scale * r)
        ge += a
        gs += b
        gq += c
        if f % 16 == 15:
            print(f"[case]   oracle frames done: {f + 1}", flush=True)
    lo = tot / (nf * n * 3)
    e, sg = fitting.mix_vjp(theta[0:2], theta[2:4], ge, gs)
    go = np.concatenate([e, sg, gq])
    assert abs(loss - lo) <= 1e-12 * lo
    assert np.linalg.norm(grad - go) <= 1e-10 * np.linalg.norm(go), (grad, go)
    # central differences of the GPU loss
    h = 1e-6
    for k in range(6):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (fm.loss(tp) - fm.loss(tm)) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(grad[k])), (k, fd, grad[k])
    print(f"[case]   loss {loss:.6e}, |grad| {np.linalg.norm(grad):.6e}", flush=True)


@pytest.fixture(scope="module")
def replicas():
    return configs.replica_set(512, (200, 4000), (0.2, 0.8), kmax=6, seed=1)


def test_512_replicas_fp64(gpu, replicas):
    """Every replica's pair set against the exact set and its FP64
    accelerations and virial against the oracle; Newton's third law and
    momentum where they hold.  They hold for boxes of at least twice the list
    radius: in a smaller box the reference's half list reaches some partners
    through several images in one direction only, and its Ewald_r, which
    takes every listed pair, is then not pairwise (the oracle shows the same
    imbalance)."""
    print(f"[case] 512 replicas fp64, {sum(s.n for s in replicas)} atoms", flush=True)
    b = Batch(replicas)
    b.setup()
    rce = replicas[0].neighbor.cutoff * replicas[0].neighbor.cutoff_extra_coef
    p0 = []
    for r, s in enumerate(replicas):
        pos = b.get(r, "pos")
        gpos, _ = F.numpy_ghosts(s, pos)
        got = F.sorted_pairs(b.pairs(r))
        want = F.sorted_pairs(F.exact_pairs(pos, gpos, rce))
        assert np.array_equal(got, want), r
        o = Oracle(s.copy(), threads=THREADS)
        o.setup()
        a = b.get(r, "acc")
        assert F.rel_err(a, o.get("acc")) <= 1e-12, r
        m = s.mass[s.types][:, None]
        # the virial sums pair terms of both signs and can nearly cancel: its
        # bar is relative to the scale of its terms (|F| times the list radius)
        vscale = max(abs(o.virial()), float(np.abs(a * m).sum()) * rce)
        assert abs(b.virial(r) - o.virial()) <= 1e-12 * vscale, r
        if s.size.min() >= 2 * rce:
            f = a * m
            assert np.abs(f.sum(axis=0)).max() <= 1e-10 * np.abs(f).sum(axis=0).max(), r
        p0.append((m * b.get(r, "vel")).sum(axis=0))
        if r % 128 == 127:
            print(f"[case]   replicas checked against the oracle: {r + 1}", flush=True)
    print("[case]   pair sets, oracle forces and Newton checked; adjoint", flush=True)
    sample = [0, 1, 63, 123, 127, 300, 411, 511]
    first = {r: b.get(r, "acc") for r in sample}
    # batched adjoint against the single-system adjoint on the same positions
    # and list (both just set up: after steps the lists' rebuild history would
    # differ, and Ewald_r takes every listed pair)
    rng = np.random.default_rng(5)
    ws = [rng.normal(size=(s.n, 3)) for s in replicas]
    gb = b.adjoint(ws)
    for r in sample:
        e = Engine(replicas[r].copy())
        e.setup()
        for x, y in zip(gb[r], e.adjoint(ws[r])):
            assert np.linalg.norm(x - y) <= 1e-10 * max(np.linalg.norm(y), 1e-300), r
    print("[case]   20 steps", flush=True)
    b.step(0.002, 20, graph=True)
    for r, s in enumerate(replicas):
        if s.size.min() < 2 * rce:
            continue
        m = s.mass[s.types][:, None]
        v = b.get(r, "vel")
        p1 = (m * v).sum(axis=0)
        assert np.abs(p1 - p0[r]).max() <= 1e-10 * (m * np.abs(v)).sum(), r
    print("[case]   momentum checked; sampled replicas against single systems", flush=True)
    for r in sample:
        s = replicas[r]
        e = Engine(s.copy())
        e.setup()
        assert F.rel_err(first[r], e.get("acc")) <= 1e-12, r
        e.step(0.002, 20)
        for what in ("pos", "vel", "acc"):
            assert F.rel_err(b.get(r, what), e.get(what)) <= 1e-9, (r, what)
        assert np.array_equal(F.sorted_pairs(b.pairs(r)), F.sorted_pairs(e.pairs())), r


def test_512_replicas_mixed_with_adjoint(gpu, replicas):
    print("[case] 512 replicas mixed, adjoint after every 10th step", flush=True)
    bm = Batch(replicas, precision="mixed")
    bm.setup()
    bd = Batch(replicas)
    bd.setup()
    for r, s in enumerate(replicas):
        a, d = bm.get(r, "acc"), bd.get(r, "acc")
        rms = np.sqrt(((a - d) ** 2).sum(1).mean() / (d ** 2).sum(1).mean())
        assert rms <= 1e-5, (r, rms)
    rng = np.random.default_rng(6)
    for _ in range(2):
        bm.step(0.002, 10, graph=True)
        ws = [rng.normal(size=(s.n, 3)) for s in replicas]
        for g in bm.adjoint(ws):
            assert all(np.all(np.isfinite(x)) for x in g)
    rce = replicas[0].neighbor.cutoff * replicas[0].neighbor.cutoff_extra_coef
    for r, s in enumerate(replicas):
        if s.size.min() < 2 * rce:
            continue
        m = s.mass[s.types][:, None]
        f = bm.get(r, "acc") * m
        assert np.abs(f.sum(axis=0)).max() <= 1e-6 * np.abs(f).sum(axis=0).max(), r
