"""cav_adjoint_noalloc, the force-matching adjoint whose workspace is the
library's fixed per-GPU scratch array, against cav_adjoint (which allocates
its workspace per call) through the C ABI.

Both run the same kernels with fixed-order reductions on the same inputs, so
the gradients must agree bit for bit.  Calls follow one another with the
second on a smaller system, so the second call finds the first one's data in
the scratch array and must not depend on it.  A system whose workspace is
larger than the array is refused with CAV_ERR_CAPACITY, and so is a call
after a parameter change that has not been applied yet (applying it
allocates).
"""
import numpy as np
import pytest

from caviar_amd import CaviarError, Engine, fitting, load_library
from caviar_amd.system import (EwaldKSettings, EwaldRSettings, LJSettings, NeighborSettings,
                               System)

pytestmark = pytest.mark.gpu

THETA_T = np.array([1.0, 0.8, 1.0, 1.2, 1.0, -1.0])  # eps, sigma, q per species


def frame(n, kmax, seed):
    theta = fitting.neutral_start(THETA_T, np.array([n // 2, n - n // 2]), seed=seed)
    return fitting.two_species_system(n, 0.3, theta, kmax, seed)


def both(e, w):
    a = e.adjoint(w)
    b = e.adjoint_noalloc(w)
    return a, b


def assert_same(a, b):
    for x, y in zip(a, b):
        assert np.array_equal(x, y), (x, y)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_adjoint_equals_allocating_adjoint(gpu, precision):
    print(f"[case] no-alloc adjoint {precision}: n=2000 kmax 5, then n=500 kmax 3", flush=True)
    big = Engine(frame(2000, 5, seed=31), precision=precision)
    big.setup()
    w = np.random.default_rng(1).normal(size=(2000, 3))
    a, b = both(big, w)
    assert_same(a, b)
    assert np.linalg.norm(a[2]) > 0
    # a second call in a row on the same system gives the same again
    assert_same(a, big.adjoint_noalloc(w))
    # then a smaller system: the scratch holds the larger call's data
    small = Engine(frame(500, 3, seed=32), precision=precision)
    small.setup()
    w2 = np.random.default_rng(2).normal(size=(500, 3))
    a2, b2 = both(small, w2)
    assert_same(a2, b2)
    assert not np.array_equal(a2[2], a[2])


def test_noalloc_adjoint_refuses_pending_parameters(gpu):
    print("[case] no-alloc adjoint after a parameter change", flush=True)
    e = Engine(frame(400, 3, seed=33))
    e.setup()
    w = np.random.default_rng(3).normal(size=(400, 3))
    e.set_parameters(charge=[0.9, -0.9])
    with pytest.raises(CaviarError) as err:
        e.adjoint_noalloc(w)
    assert err.value.code == 3
    e.recalculate()
    assert_same(e.adjoint(w), e.adjoint_noalloc(w))


def test_noalloc_adjoint_refuses_oversized_workspace(gpu):
    """A 600k-atom +-1 lattice needs more than the 32 MiB array; the
    allocating entry point still serves it."""
    lib = load_library()
    assert lib.cav_adjoint_noalloc_scratch_bytes() == 32 << 20
    m = 84
    n = m ** 3
    rho = 0.3
    a = (1.0 / rho) ** (1.0 / 3.0)
    g = (np.arange(m) + 0.5) * a
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    idx = np.indices((m, m, m)).reshape(3, -1).sum(axis=0)
    types = (idx % 2).astype(np.int32)
    box = m * a
    print(f"[case] no-alloc adjoint oversized: n={n}", flush=True)
    s = System(lower=[0, 0, 0], upper=[box] * 3, bc=[1, 1, 1], ghost_cutoff=3.5 * 1.12347 + 0.1,
               mass=[1.0, 1.0], charge=[1.0, -1.0], pos=pos, vel=np.zeros((n, 3)), types=types,
               neighbor=NeighborSettings(cutoff=3.5),
               lj=LJSettings(cutoff=2.5, epsilon=np.ones((2, 2)), sigma=np.ones((2, 2))),
               ewald_r=EwaldRSettings(alpha=0.9, k_electrostatic=2.0),
               ewald_k=EwaldKSettings.cube(0.9, 3, 2.0))
    e = Engine(s)
    e.setup()
    w = np.random.default_rng(4).normal(size=(n, 3))
    with pytest.raises(CaviarError) as err:
        e.adjoint_noalloc(w)
    assert err.value.code == 5
    assert "cav_adjoint" in str(err.value)
    g = e.adjoint(w)
    assert np.all(np.isfinite(g[2]))
