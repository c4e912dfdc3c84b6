"""Force-matching adjoint: oracle pinned by central finite differences (CPU),
GPU adjoint checked against the oracle and against finite differences of the
GPU loss (h = 1e-6, bar 1e-6 relative), and the torch.autograd wrapper."""
import numpy as np
import pytest

from caviar_amd import fitting
from caviar_amd.configs import random_nonoverlap
from oracle import Oracle

THETA_T = np.array([1.0, 0.8, 1.0, 1.2, 1.0, -1.0])  # eps, sigma, q per species


def small_frames(n=120, nframes=2, kmax=3, seed=4):
    base = fitting.two_species_system(n, 0.3, THETA_T, kmax, seed)
    frames = np.stack([random_nonoverlap(n, base.size[0], fitting.DMIN, seed + 17 * (f + 1))
                       for f in range(nframes)])
    return base, frames


def oracle_forces(base, pos, theta):
    s = base.copy()
    em, sm = fitting.mix(theta[0:2], theta[2:4])
    s.lj.epsilon, s.lj.sigma = em, sm
    s.charge = np.asarray(theta[4:6], dtype=np.float64)
    s.pos = pos
    o = Oracle(s)
    o.setup()
    return o, o.get("acc") * s.mass[s.types][:, None]


def oracle_loss_grad(base, frames, targets, theta):
    nf, n = frames.shape[:2]
    scale = 2.0 / (nf * n * 3)
    tot = 0.0
    ge = np.zeros((2, 2))
    gs = np.zeros((2, 2))
    gq = np.zeros(2)
    for f in range(nf):
        o, F = oracle_forces(base, frames[f], theta)
        r = F - targets[f]
        tot += float(np.sum(r * r))
        a, b, c = o.adjoint(scale * r)
        ge += a
        gs += b
        gq += c
    e, s = fitting.mix_vjp(theta[0:2], theta[2:4], ge, gs)
    return tot / (nf * n * 3), np.concatenate([e, s, gq])


def test_oracle_adjoint_matches_central_differences():
    base, frames = small_frames()
    targets = np.stack([oracle_forces(base, fr, THETA_T)[1] for fr in frames])
    theta = fitting.neutral_start(THETA_T, np.array([60, 60]), seed=3)
    loss, grad = oracle_loss_grad(base, frames, targets, theta)
    h = 1e-6
    for k in range(6):
        tp = theta.copy()
        tm = theta.copy()
        tp[k] += h
        tm[k] -= h
        lp = oracle_loss_grad(base, frames, targets, tp)[0]
        lm = oracle_loss_grad(base, frames, targets, tm)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(grad[k])), (k, fd, grad[k])


def test_mixing_vjp_against_differences():
    eps = np.array([1.1, 0.7])
    sig = np.array([0.9, 1.3])
    rng = np.random.default_rng(0)
    ge_m, gs_m = rng.normal(size=(2, 2)), rng.normal(size=(2, 2))
    ge, gs = fitting.mix_vjp(eps, sig, ge_m, gs_m)
    h = 1e-7
    for s in range(2):
        d = np.zeros(2)
        d[s] = h
        fe = (np.sum(ge_m * fitting.mix(eps + d, sig)[0]) - np.sum(ge_m * fitting.mix(eps - d, sig)[0])) / (2 * h)
        fs = (np.sum(gs_m * fitting.mix(eps, sig + d)[1]) - np.sum(gs_m * fitting.mix(eps, sig - d)[1])) / (2 * h)
        assert abs(fe - ge[s]) < 1e-8 and abs(fs - gs[s]) < 1e-8


@pytest.mark.gpu
def test_gpu_adjoint_matches_oracle(gpu):
    from caviar_amd import Engine
    base, frames = small_frames(n=300, nframes=1, kmax=4, seed=8)
    theta = fitting.neutral_start(THETA_T, np.array([150, 150]), seed=5)
    o, F = oracle_forces(base, frames[0], theta)
    w = np.random.default_rng(1).normal(size=F.shape)
    ge_o, gs_o, gq_o = o.adjoint(w)
    s = base.copy()
    s.lj.epsilon, s.lj.sigma = fitting.mix(theta[0:2], theta[2:4])
    s.charge = theta[4:6]
    s.pos = frames[0]
    e = Engine(s)
    e.setup()
    assert np.linalg.norm(e.get("acc") - F) <= 1e-12 * np.linalg.norm(F)
    ge, gs, gq = e.adjoint(w)
    for g, go in ((ge, ge_o), (gs, gs_o), (gq, gq_o)):
        assert np.linalg.norm(g - go) <= 1e-10 * max(np.linalg.norm(go), 1e-300)


@pytest.mark.gpu
def test_gpu_force_matching_gradient_vs_central_differences(gpu):
    fs = fitting.make_frameset(nframes=3, n=400, theta_target=THETA_T, kmax=4, seed=11)
    fm = fitting.ForceMatcher(fs)
    theta = fitting.neutral_start(THETA_T, np.array([200, 200]), seed=7)
    loss, grad = fm.loss_and_grad(theta)
    assert loss > 0
    h = 1e-6
    for k in range(6):
        tp = theta.copy()
        tm = theta.copy()
        tp[k] += h
        tm[k] -= h
        fd = (fm.loss(tp) - fm.loss(tm)) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-6 * max(1.0, abs(grad[k])), (k, fd, grad[k])


@pytest.mark.gpu
def test_torch_autograd_wrapper(gpu):
    import torch
    fs = fitting.make_frameset(nframes=2, n=300, theta_target=THETA_T, kmax=3, seed=21)
    fm = fitting.ForceMatcher(fs)
    theta0 = fitting.neutral_start(THETA_T, np.array([150, 150]), seed=2)
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    loss = fitting.ForceMatchingLoss.apply(th, fm)
    loss.backward()
    _, grad = fm.loss_and_grad(theta0)
    assert np.allclose(th.grad.numpy(), grad, rtol=1e-12, atol=0)
    # a few gradient steps on the species values reduce the loss
    opt = torch.optim.Adam([th], lr=2e-3)
    l0 = float(loss)
    for _ in range(5):
        opt.zero_grad()
        lo = fitting.ForceMatchingLoss.apply(th, fm)
        lo.backward()
        opt.step()
    assert float(fitting.ForceMatchingLoss.apply(th, fm)) < l0


@pytest.mark.gpu
def test_gpu_adjoint_mixed_form_close_to_fp64(gpu):
    """The adjoint in the mixed form (FP32 pair and k-space products, FP64
    sums) against the FP64 adjoint on the same frame: 1e-5 relative."""
    from caviar_amd import Engine
    base, frames = small_frames(n=400, nframes=1, kmax=5, seed=12)
    theta = fitting.neutral_start(THETA_T, np.array([200, 200]), seed=9)
    s = base.copy()
    s.lj.epsilon, s.lj.sigma = fitting.mix(theta[0:2], theta[2:4])
    s.charge = theta[4:6]
    s.pos = frames[0]
    w = np.random.default_rng(3).normal(size=(s.n, 3))
    out = {}
    for prec in ("fp64", "mixed"):
        e = Engine(s.copy(), precision=prec)
        e.setup()
        out[prec] = e.adjoint(w)
    for g64, g32 in zip(out["fp64"], out["mixed"]):
        assert np.linalg.norm(g32 - g64) <= 1e-5 * np.linalg.norm(g64)
