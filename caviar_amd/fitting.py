"""Force matching of per-species LJ and charge parameters through the GPU adjoint.

Parameters theta = (eps_1..eps_S, sigma_1..sigma_S, q_1..q_S) for S species.
The LJ type matrices follow the Lorentz-Berthelot rule of the reference's
``force_field lj`` with ``epsilon_atom``/``sigma_atom`` (lj.cpp:207-222):
eps_ab = sqrt(eps_a eps_b), sigma_ab = (sigma_a + sigma_b)/2 (the diagonal is the
species value itself).  The loss is the mean squared force error over frames,
atoms and components, accumulated in FP64:

    L = 1/(F N 3) sum_f sum_i |F_i(theta; r_f) - T_i^f|^2.

Its gradient comes from ``Engine.adjoint`` (the hand-written HIP adjoint of the
LJ, Ewald_r and Ewald_k force formulas) chained through the mixing rule here;
:class:`ForceMatchingLoss` wraps it as a ``torch.autograd.Function``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import Engine
from .configs import random_nonoverlap
from .system import EwaldKSettings, EwaldRSettings, LJSettings, NeighborSettings, System


def mix(eps: np.ndarray, sig: np.ndarray):
    """Lorentz-Berthelot type matrices from per-species values (lj.cpp:207-222)."""
    eps = np.asarray(eps, dtype=np.float64)
    sig = np.asarray(sig, dtype=np.float64)
    return np.sqrt(np.outer(eps, eps)), 0.5 * (sig[:, None] + sig[None, :])


def mix_vjp(eps: np.ndarray, sig: np.ndarray, g_eps_m: np.ndarray, g_sig_m: np.ndarray):
    """Chain dL/d(matrix entries) to dL/d(per-species values) through the mixing rule."""
    eps = np.asarray(eps, dtype=np.float64)
    ns = len(eps)
    em, _ = mix(eps, sig)
    ge = np.zeros(ns)
    gs = np.zeros(ns)
    for a in range(ns):
        for b in range(ns):
            for s in (a, b):
                ge[s] += g_eps_m[a, b] * em[a, b] / (2.0 * eps[s])
                gs[s] += g_sig_m[a, b] * 0.5
    return ge, gs


@dataclass
class FrameSet:
    system: System          # template: box, types, force-field settings
    frames: np.ndarray      # (F, N, 3) positions
    targets: np.ndarray     # (F, N, 3) target forces


# "Non-overlapping" for species of sigma 1.0 and 1.2: no pair closer than the
# mixed diameter sigma_01 = 1.1.
DMIN = 1.1


def two_species_system(n: int, rho: float, theta: np.ndarray, kmax: int, seed: int,
                       rc_lj: float = 2.5, nl_cutoff: float = 3.5, alpha: float = 0.9,
                       ke: float = 2.0) -> System:
    """Half species 0, half species 1, random non-overlapping positions."""
    box = (n / rho) ** (1.0 / 3.0)
    pos = random_nonoverlap(n, box, DMIN, seed)
    types = np.zeros(n, dtype=np.int32)
    types[n // 2:] = 1
    eps, sig, q = theta[0:2], theta[2:4], theta[4:6]
    em, sm = mix(eps, sig)
    coef = 1.12347
    return System(lower=[0, 0, 0], upper=[box] * 3, bc=[1, 1, 1],
                  ghost_cutoff=nl_cutoff * coef + 0.1, mass=[1.0, 1.0], charge=q, pos=pos,
                  vel=np.zeros((n, 3)), types=types,
                  neighbor=NeighborSettings(cutoff=nl_cutoff, cutoff_extra_coef=coef),
                  lj=LJSettings(cutoff=rc_lj, epsilon=em, sigma=sm),
                  ewald_r=EwaldRSettings(alpha=alpha, k_electrostatic=ke),
                  ewald_k=EwaldKSettings.cube(alpha, kmax, ke),
                  meta={"workload": "force_matching", "n": n, "rho": rho, "kmax": kmax})


def neutral_start(theta_target: np.ndarray, counts: np.ndarray, seed: int, spread: float = 0.1):
    """Targets times (1 + U(-spread, spread)) per parameter, then the last
    species' charge set so that sum_s counts_s q_s = 0."""
    rng = np.random.default_rng(seed)
    th = np.asarray(theta_target, dtype=np.float64) * (1.0 + rng.uniform(-spread, spread, 6))
    th[5] = -th[4] * counts[0] / counts[1]
    return th


class ForceMatcher:
    """Loss and gradient of force matching on one device-resident engine."""

    def __init__(self, frameset: FrameSet, precision: str = "fp64"):
        self.fs = frameset
        self.engine = Engine(frameset.system.copy(), precision=precision)
        self.engine.setup()
        self.nf, self.n = frameset.frames.shape[0], frameset.frames.shape[1]

    def _set_theta(self, theta):
        th = np.asarray(theta, dtype=np.float64)
        em, sm = mix(th[0:2], th[2:4])
        self.engine.set_parameters(epsilon=em, sigma=sm, charge=th[4:6])

    def forces(self, theta, frame: int) -> np.ndarray:
        self._set_theta(theta)
        e = self.engine
        e.set("pos", self.fs.frames[frame])
        e.recalculate(force_rebuild=True)
        return e.get("acc") * e.system.mass[e.system.types][:, None]

    def loss(self, theta) -> float:
        tot = 0.0
        for f in range(self.nf):
            r = self.forces(theta, f) - self.fs.targets[f]
            tot += float(np.sum(r * r))
        return tot / (self.nf * self.n * 3)

    def loss_and_grad(self, theta):
        th = np.asarray(theta, dtype=np.float64)
        scale = 2.0 / (self.nf * self.n * 3)
        tot = 0.0
        ge_m = np.zeros((2, 2))
        gs_m = np.zeros((2, 2))
        gq = np.zeros(2)
        for f in range(self.nf):
            r = self.forces(th, f) - self.fs.targets[f]
            tot += float(np.sum(r * r))
            a, b, c = self.engine.adjoint(scale * r)
            ge_m += a
            gs_m += b
            gq += c
        ge, gs = mix_vjp(th[0:2], th[2:4], ge_m, gs_m)
        return tot / (self.nf * self.n * 3), np.concatenate([ge, gs, gq])


def make_frameset(nframes: int, n: int, theta_target: np.ndarray, kmax: int = 6, rho: float = 0.3,
                  seed: int = 1, precision: str = "fp64") -> FrameSet:
    """Random non-overlapping frames with target forces from theta_target."""
    base = two_species_system(n, rho, theta_target, kmax, seed)
    frames = np.stack([random_nonoverlap(n, base.size[0], DMIN, seed + 17 * (f + 1))
                       for f in range(nframes)])
    fs = FrameSet(base, frames, np.zeros_like(frames))
    fm = ForceMatcher(fs, precision=precision)
    for f in range(nframes):
        fs.targets[f] = fm.forces(theta_target, f)
    fm.engine.close()
    return fs


try:  # the autograd wrapper needs torch; the rest of the module does not
    import torch

    class ForceMatchingLoss(torch.autograd.Function):
        """L(theta) with the GPU adjoint as its backward (theta: 6 FP64 values)."""

        @staticmethod
        def forward(ctx, theta, matcher):
            th = theta.detach().cpu().numpy().astype(np.float64)
            loss, grad = matcher.loss_and_grad(th)
            ctx.save_for_backward(torch.as_tensor(grad, dtype=theta.dtype))
            return theta.new_tensor(loss)

        @staticmethod
        def backward(ctx, grad_out):
            (grad,) = ctx.saved_tensors
            return grad_out * grad.to(grad_out.device), None
except ImportError:  # pragma: no cover
    ForceMatchingLoss = None
