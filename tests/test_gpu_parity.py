"""Parity of the HIP path (through the C ABI) with the reference's results.

The golden fixtures hold the reference library's own outputs on each input
(tests/golden/make_golden.py); the oracle, itself pinned bit-for-bit to those
fixtures, supplies the same quantities for inputs generated here.

Bars (written in each assertion):
* ghost images: bit-exact positions and source ids;
* Verlet pair sets: bit-exact as sorted (i, j) lists;
* FP64 accelerations, virial, energies, S(k), point fields: 1e-12 relative
  (norm-wise), trajectories after tens of steps: 1e-9;
* mixed form: force RMS relative error <= 1e-5 against FP64.
"""
import numpy as np
import pytest

import fixtures as F
from caviar_amd import Batch, CaviarError, Engine, NaNAccelerationError, configs
from oracle import Oracle

pytestmark = pytest.mark.gpu

SETUP = [n for n in F.names() if n.endswith("_setup")]
RUNS = [n for n in F.names() if "_run" in n]


def engine_for(name, precision="fp64"):
    s, ref, meta = F.load(name)
    print(f"[case] {name} n={s.n} precision={precision}", flush=True)
    e = Engine(s, precision=precision)
    e.setup()
    if meta["run_steps"]:
        e.step(meta["dt"], meta["run_steps"])
    return e, s, ref, meta


@pytest.mark.parametrize("name", SETUP)
def test_setup_parity_fp64(gpu, name):
    e, s, ref, meta = engine_for(name)
    gp, gs = e.ghosts()
    assert np.array_equal(gp, ref["ghost_pos"])                       # bit-exact
    assert np.array_equal(s.ids[gs], ref["ghost_id"])
    assert np.array_equal(F.sorted_pairs(e.pairs()), F.sorted_pairs(ref["pairs"]))  # bit-exact set
    assert F.rel_err(e.get("acc"), ref["acc"]) <= 1e-12
    v = float(ref["virial"])
    assert abs(e.virial() - v) <= 1e-12 * max(abs(v), 1e-300) or v == e.virial()
    if "energy_r" in ref or "energy_k" in ref:
        elj, er, ek = e.energy(lj=False, r="energy_r" in ref, k="energy_k" in ref)
        if "energy_r" in ref:
            assert abs(er - ref["energy_r"]) <= 1e-12 * abs(ref["energy_r"])
        if "energy_k" in ref:
            assert abs(ek - ref["energy_k"]) <= 1e-12 * abs(ref["energy_k"])
    if "sk" in ref:
        sk = e.sk()
        assert F.rel_err(np.c_[sk.real, sk.imag], ref["sk"]) <= 1e-12
    check_thermo(e, ref, 1e-12)
    if "energy_elc" in ref:
        assert abs(e.energy_elc() - ref["energy_elc"]) <= 1e-12 * abs(ref["energy_elc"])
    shifted = e.get("pos") + np.array([0.1, -0.2, 0.3])
    if "field_elc" in ref:
        f, p = e.point_field("elc", shifted)
        assert F.rel_err(f, ref["field_elc"][:, :3]) <= 1e-12
        assert F.rel_err(p, ref["field_elc"][:, 3]) <= 1e-12
    if "field_k" in ref:
        f, p = e.point_field("k", shifted)
        assert F.rel_err(f, ref["field_k"][:, :3]) <= 1e-12
        assert F.rel_err(p, ref["field_k"][:, 3]) <= 1e-12
    if "field_r" in ref:
        f, p = e.point_field("r", shifted)
        assert F.rel_err(f, ref["field_r"][:, :3]) <= 1e-12
        assert F.rel_err(p, ref["field_r"][:, 3]) <= 1e-12


def check_thermo(e, ref, tol):
    """Kinetic energy, temperature and pressure (the last two as finalized at
    the end of the last step, or after setup) against the reference's."""
    if "temperature" not in ref:
        return
    ke, t, p = e.thermo()
    for got, key in ((ke, "kinetic_energy"), (t, "temperature"), (p, "pressure")):
        want = float(ref[key])
        assert abs(got - want) <= tol * max(abs(want), 1e-300) or got == want, key


@pytest.mark.parametrize("name", RUNS)
def test_trajectory_parity_fp64(gpu, name):
    e, s, ref, meta = engine_for(name)
    assert F.rel_err(e.get("pos"), ref["pos"]) <= 1e-9
    assert F.rel_err(e.get("vel"), ref["vel"]) <= 1e-9
    assert F.rel_err(e.get("acc"), ref["acc"]) <= 1e-9
    assert np.array_equal(F.sorted_pairs(e.pairs()), F.sorted_pairs(ref["pairs"]))
    check_thermo(e, ref, 1e-9)


def test_b2_published_known_answer(gpu):
    pub = F.published_b2()
    s = configs.b2_ewald(pub["pos"], pub["q"], kmax=25)
    e = Engine(s)
    e.setup()
    _, er, ek = e.energy(lj=False)
    assert abs(er + ek - float(pub["energy"])) <= 2e-12 * abs(float(pub["energy"]))
    assert F.rel_err(e.get("acc"), pub["forces"]) <= 1e-12


@pytest.mark.parametrize("name", ["ions600_setup", "ions2000_setup", "aniso400_setup",
                                  "lambda_dipole300_setup", "small200_setup", "lj4000_setup",
                                  "ions8000_setup", "slabz500_setup", "ljatom_sr400_setup",
                                  "ljwca_mol360_setup", "allfields500_setup", "elcz400_setup"])
def test_mixed_precision_force_error(gpu, name):
    e64, s, ref, meta = engine_for(name)
    e32, _, _, _ = engine_for(name, "mixed")
    assert np.array_equal(F.sorted_pairs(e32.pairs()), F.sorted_pairs(ref["pairs"]))
    a64, a32 = e64.get("acc"), e32.get("acc")
    rms = np.sqrt(((a32 - a64) ** 2).sum(axis=1).mean() / (a64 ** 2).sum(axis=1).mean())
    assert rms <= 1e-5                     # force RMS relative error bar of the mixed form
    v64, v32 = e64.virial(), e32.virial()
    assert abs(v32 - v64) <= 1e-5 * max(abs(v64), 1.0)


@pytest.mark.parametrize("k", [(1, 1, 1), (2, 5, 0), (8, 3, 8), (12, 12, 12), (5, 9, 16),
                               (20, 7, 20), (16, 16, 24), (31, 2, 31), (31, 31, 31)])
def test_mixed_lattice_shapes(gpu, k):
    """The split Ewald_k over lattice shapes that exercise its geometry: kz = 0
    (only the accumulator seed), K axes that are and are not multiples of 32
    (2 (nZ - 1) = 4 kz = 0 .. 124 -> 1..4 K-steps), 1..8 column tiles, record
    rows from 6 to 95, and the widest lattice (whose force-pass tiles exceed
    LDS, so the FP32 tiles take over).  Mixed forces within 1e-5 RMS
    of FP64 on the same positions; the k-space energy within 1e-6."""
    kx, ky, kz = k
    s = configs.ions(n=1200, rho=0.3, kmax=1, seed=31 + kx + ky + kz)
    s.ewald_k.kx_max, s.ewald_k.ky_max, s.ewald_k.kz_max = kx, ky, kz
    print(f"[case] mixed lattice {k}", flush=True)
    out = {}
    for prec in ("fp64", "mixed"):
        e = Engine(s.copy(), precision=prec)
        e.setup()
        out[prec] = (e.get("acc"), e.energy(lj=False)[2])
    a64, a32 = out["fp64"][0], out["mixed"][0]
    rms = np.sqrt(((a32 - a64) ** 2).sum(axis=1).mean() / (a64 ** 2).sum(axis=1).mean())
    assert rms <= 1e-5, (k, rms)
    assert abs(out["mixed"][1] - out["fp64"][1]) <= 1e-6 * abs(out["fp64"][1]), k


@pytest.mark.parametrize("k", [(2, 2, 32), (3, 1, 40), (2, 3, 70), (120, 120, 2)])
def test_wide_lattices_against_oracle(gpu, k):
    """k lattices past one force-pass K slab (|nz| > 31: 2 and 3 slabs of up
    to 32 K-steps), past 8 S-pass column tiles (2 and 3 launches of column
    groups), and with [X | Y | Z] records too long for 16-atom S-pass stages
    (kx = ky = 120: 8-atom stages).  FP64 accelerations, S(k), the k-space
    energy and point fields against the oracle (1e-12); the mixed form, which
    takes the FP32 tiles here, within 1e-5 force RMS of FP64."""
    kx, ky, kz = k
    s = configs.ions(n=300 if kx > 100 else 900, rho=0.3, kmax=1, seed=7 + kz)
    s.ewald_k.kx_max, s.ewald_k.ky_max, s.ewald_k.kz_max = kx, ky, kz
    print(f"[case] wide lattice {k}", flush=True)
    o = Oracle(s)
    o.setup()
    e = Engine(s.copy())
    e.setup()
    a64 = e.get("acc")
    assert F.rel_err(a64, o.get("acc")) <= 1e-12, k
    sk, osk = e.sk(), o.sk()
    assert F.rel_err(np.c_[sk.real, sk.imag], np.c_[osk.real, osk.imag]) <= 1e-12, k
    ek = e.energy(lj=False)[2]
    assert abs(ek - o.energy_k()) <= 1e-12 * abs(o.energy_k()), k
    pts = e.get("pos")[:50] + np.array([0.1, -0.2, 0.3])
    f, p = e.point_field("k", pts)
    fo, po = o.point_k(pts)
    assert F.rel_err(f, fo) <= 1e-12, k
    assert F.rel_err(p, po) <= 1e-12, k
    em = Engine(s.copy(), precision="mixed")
    em.setup()
    a32 = em.get("acc")
    rms = np.sqrt(((a32 - a64) ** 2).sum(axis=1).mean() / (a64 ** 2).sum(axis=1).mean())
    assert rms <= 1e-5, (k, rms)


def test_wide_lattice_batch(gpu):
    """Batched replicas with kz = 40 (two K slabs, two column groups): FP64
    against single-system runs (1e-12), the mixed form (FP32 tiles) within
    1e-5 force RMS, and the per-replica k-space field of both forms."""
    reps = []
    for r, n in enumerate((300, 900, 500)):
        s = configs.ions(n=n, rho=0.3, kmax=1, seed=40 + r)
        s.ewald_k.kx_max, s.ewald_k.ky_max, s.ewald_k.kz_max = 2, 3, 40
        reps.append(s)
    print("[case] batch of 3, lattice (2, 3, 40)", flush=True)
    bd = Batch(reps)
    bd.setup()
    bm = Batch(reps, precision="mixed")
    bm.setup()
    for r, s in enumerate(reps):
        e = Engine(s.copy())
        e.setup()
        a = bd.get(r, "acc")
        assert F.rel_err(a, e.get("acc")) <= 1e-12, r
        am = bm.get(r, "acc")
        rms = np.sqrt(((am - a) ** 2).sum(axis=1).mean() / (a ** 2).sum(axis=1).mean())
        assert rms <= 1e-5, (r, rms)
    bd.step(0.002, 5)
    bm.step(0.002, 5)
    for r, s in enumerate(reps):
        e = Engine(s.copy())
        e.setup()
        e.step(0.002, 5)
        assert F.rel_err(bd.get(r, "acc"), e.get("acc")) <= 1e-9, r


def test_fp64_is_bitwise_reproducible(gpu):
    s = configs.ions(n=3000, kmax=6, seed=9)
    runs = []
    for _ in range(2):
        e = Engine(s)
        e.setup()
        e.step(0.002, 5)
        runs.append((e.get("pos"), e.get("acc"), e.virial()))
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]


def test_against_oracle_generated_inputs(gpu):
    """Inputs made here (not fixtures): unequal kmax per axis and a box with
    the lower corner away from the origin (slab geometry has its own reference
    fixture)."""
    s = configs.ions(n=700, kmax=4, seed=77)
    shift = np.array([-3.0, 5.5, 1.25])
    s.lower = s.lower + shift
    s.upper = s.upper + shift
    s.pos = s.pos + shift
    s.ewald_k.kx_max, s.ewald_k.ky_max, s.ewald_k.kz_max = 3, 5, 4
    o = Oracle(s)
    o.setup()
    e = Engine(s)
    e.setup()
    assert np.array_equal(F.sorted_pairs(e.pairs()), F.sorted_pairs(o.pairs()))
    assert F.rel_err(e.get("acc"), o.get("acc")) <= 1e-12
    sk = e.sk()
    osk = o.sk()
    assert F.rel_err(np.c_[sk.real, sk.imag], np.c_[osk.real, osk.imag]) <= 1e-12
    _, er, ek = e.energy(lj=False)
    assert abs(ek - o.energy_k()) <= 1e-12 * abs(o.energy_k())
    assert abs(er - o.energy_r()) <= 1e-12 * abs(o.energy_r())
    # run both for 25 steps (list rebuilds included) and compare
    assert o.run(0.002, 25) == 0
    e.step(0.002, 25)
    assert F.rel_err(e.get("pos"), o.get("pos")) <= 1e-9
    assert F.rel_err(e.get("acc"), o.get("acc")) <= 1e-9


def test_lj_energy_matches_oracle(gpu):
    s = configs.lj_fcc(ncell=4, seed=5)
    o = Oracle(s)
    o.setup()
    e = Engine(s)
    e.setup()
    elj, _, _ = e.energy(lj=True, r=False, k=False)
    assert abs(elj - o.energy_lj()) <= 1e-12 * abs(o.energy_lj())


def test_lj_cell_walk_path(gpu):
    """LJ cutoff above the list cutoff: the cell-walk kernel (every step) is used."""
    s = configs.ions(n=500, kmax=3, seed=3, nl_cutoff=2.0, rc_lj=2.5)
    o = Oracle(s)
    o.setup()
    e = Engine(s)
    e.setup()
    assert F.rel_err(e.get("acc"), o.get("acc")) <= 1e-12
    assert o.run(0.002, 10) == 0
    e.step(0.002, 10)
    assert F.rel_err(e.get("acc"), o.get("acc")) <= 1e-9


def test_nan_acceleration_is_reported_and_state_kept(gpu):
    s, ref, meta = F.load("lj256_setup")
    e = Engine(s)
    e.setup()
    acc = e.get("acc")
    acc[3, 1] = np.nan
    e.set("acc", acc)
    before = e.get("pos")
    with pytest.raises(NaNAccelerationError):
        e.step(0.005, 1)
    assert np.array_equal(e.get("pos"), before)


def test_list_capacity_growth_is_transparent(gpu):
    """A compressed configuration overflows the initial row capacity; the step
    is redone with a larger list and matches the oracle."""
    s = configs.ions(n=800, kmax=3, seed=13)
    # squeeze half the atoms into one corner (still >= 0.8 apart: scaled lattice)
    g = np.stack(np.meshgrid(*[np.arange(8) * 0.82] * 3, indexing="ij"), -1).reshape(-1, 3)[:400]
    s.pos[:400] = g + 0.5
    o = Oracle(s)
    o.setup()
    e = Engine(s)
    e.setup()
    assert np.array_equal(F.sorted_pairs(e.pairs()), F.sorted_pairs(o.pairs()))
    assert F.rel_err(e.get("acc"), o.get("acc")) <= 1e-12


@pytest.mark.parametrize("name", ["ions600_setup", "lj256_setup", "lambda_dipole300_setup",
                                  "allfields500_setup", "ljwca_mol360_setup", "elcz400_setup"])
def test_per_field_calculate_acceleration(gpu, name):
    """Force_field::calculate_acceleration one field at a time (cav_forces):
    each field's acceleration against the oracle's same field, and all of them
    added together against the fused force phase (1e-12 relative).  The slab
    correction is stateful: its dipole field comes from the previous
    evaluation's dipole sum (zero at setup), so with it the fields are compared
    with one more evaluation at the same positions instead of with setup's."""
    s, ref, meta = F.load(name)
    print(f"[case] per-field {name}", flush=True)
    e = Engine(s)
    e.setup()
    fused = e.get("acc")
    o = Oracle(s)
    o.setup()
    fields = [f for f, on in (("lj", s.lj), ("lj_list", s.lj_list), ("ewald_r", s.ewald_r),
                              ("short_range", s.short_range), ("ewald_k", s.ewald_k),
                              ("elc", s.elc)) if on is not None]
    orc = {"lj": o.lib.orc_lj_acc, "ewald_r": o.lib.orc_ewald_r_acc, "ewald_k": o.lib.orc_ewald_k_acc,
           "lj_list": o.lib.orc_lj_list_acc, "short_range": o.lib.orc_short_range_acc,
           "elc": o.lib.orc_elc_acc}
    total = np.zeros_like(fused)
    for f in fields:
        e.reset_acceleration()
        e.forces(f)
        mine = e.get("acc")
        o.lib.orc_reset_acc(o.h)
        orc[f](o.h)
        want = o.get("acc")
        assert F.rel_err(mine, want) <= 1e-12, f
        total += mine
    e.reset_acceleration()
    e.forces(*fields)
    if s.elc is None:
        assert F.rel_err(e.get("acc"), fused) <= 1e-12
        assert F.rel_err(total, fused) <= 1e-12
    else:
        assert F.rel_err(e.get("acc"), total) <= 1e-12
    # the exchange + build pieces reproduce the fused path's ghosts and list
    e.exchange()
    e.neighbor_build(False)
    assert np.array_equal(F.sorted_pairs(e.pairs()), F.sorted_pairs(ref["pairs"]))


def test_autotune_keeps_results(gpu):
    """The tile autotuner changes only the Ewald_k tiling: FP64 forces stay
    within 1e-12 of the oracle, mixed within 1e-5 of FP64."""
    s = configs.ions(n=1500, kmax=5, seed=9)
    print("[case] autotune", flush=True)
    o = Oracle(s)
    o.setup()
    want = o.get("acc")
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-5)):
        e = Engine(s, precision=prec)
        e.setup()
        tiles = e.autotune()
        assert tiles["s_chunk"] in (256, 512, 1024, 2048, 4096)
        a = e.get("acc")
        err = F.rel_err(a, want) if prec == "fp64" else np.sqrt(((a - want) ** 2).sum(1).mean() / (want ** 2).sum(1).mean())
        assert err <= tol, (prec, tiles, err)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("dipole", [False, True])
def test_concurrent_ewald_k_same_trajectory(gpu, precision, dipole):
    """Ewald_k on its own stream, concurrent with the ghost exchange, list
    build and pair kernel, gives the in-order trajectory bit for bit (the side
    accelerations are added to the short-range ones once, then the
    epsilon_dipole term, as in order), over steps that rebuild the list."""
    s = configs.ions(n=3000, kmax=6, seed=4)
    if dipole:
        s.ewald_k.epsilon_dipole = 1.5
    print(f"[case] concurrent Ewald_k {precision} dipole={dipole}", flush=True)
    out = []
    for overlap in (True, False):
        e = Engine(s.copy(), precision=precision)
        e.set_overlap(overlap)
        e.setup()
        e.step(0.002, 25)
        out.append([e.get(k) for k in ("pos", "vel", "acc")])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_pinned_tiles_keep_results(gpu):
    """Every pinned Ewald_k tile shape (row tiles 1..4, chunks down to one
    32-atom table block and past n, force tiles 1/2 x 8/16 waves) gives the
    oracle's forces: FP64 within 1e-12, mixed within 1e-5; an out-of-range
    shape is refused."""
    s = configs.ions(n=1500, kmax=5, seed=9)
    print("[case] pinned tiles", flush=True)
    o = Oracle(s)
    o.setup()
    want = o.get("acc")
    e = Engine(s, precision="mixed")
    e.setup()
    for rt, ch, nt, fw in ((1, 32, 1, 8), (2, 96, 2, 8), (3, 512, 1, 16), (4, 4096, 2, 16)):
        tiles = e.set_tiles(rt, ch, nt, fw)
        assert (tiles["s_rows"], tiles["s_chunk"], tiles["f_tiles"], tiles["f_waves"]) == (rt, ch, nt, fw)
        a = e.get("acc")
        err = np.sqrt(((a - want) ** 2).sum(1).mean() / (want ** 2).sum(1).mean())
        assert err <= 1e-5, (rt, ch, nt, fw, err)
    with pytest.raises(CaviarError):
        e.set_tiles(5, 1024, 1, 16)
    with pytest.raises(CaviarError):
        e.set_tiles(3, 1000, 1, 16)
    with pytest.raises(CaviarError):
        e.set_tiles(3, 8192, 1, 16)
    e = Engine(s, precision="fp64")
    e.setup()
    e.set_tiles(3, 256, 1, 8)
    assert F.rel_err(e.get("acc"), want) <= 1e-12


@pytest.mark.parametrize("name", ["ions2000_setup", "aniso400_setup", "lj4000_setup",
                                  "lambda_dipole300_setup"])
def test_pair_kernels_agree(gpu, name):
    """The cell-staged pair kernel (neighbours read from the cell's LDS tile)
    and the row-gather kernel: both within 1e-12 of the reference in FP64 and
    within 1e-5 force RMS in the mixed form, over a 20-step trajectory with
    energy queries (which re-bin the cells between list builds) interleaved."""
    s, ref, meta = F.load(name)
    print(f"[case] pair kernels {name}", flush=True)
    o = Oracle(s)
    o.setup()
    want0 = o.get("acc")
    assert o.run(0.002, 20) == 0
    want20 = o.get("acc")
    for prec in ("fp64", "mixed"):
        got = {}
        for kind in ("rows", "cells"):
            e = Engine(s.copy(), precision=prec)
            e.set_pair_kernel(kind)
            e.setup()
            lam = s.ewald_r is not None and s.ewald_r.lambda_ is not None
            expect = "rows" if (kind == "rows" or lam) else "cells"
            assert e.pair_kernel() == expect, (kind, e.pair_kernel())
            a0 = e.get("acc")
            for _ in range(4):
                e.step(0.002, 5)
                e.energy(lj=s.lj is not None, r=s.ewald_r is not None, k=False)
            got[kind] = (a0, e.get("acc"))
        for kind, (a0, a20) in got.items():
            if prec == "fp64":
                assert F.rel_err(a0, ref["acc"]) <= 1e-12, kind
                assert F.rel_err(a0, want0) <= 1e-12, kind
                assert F.rel_err(a20, want20) <= 1e-9, kind
            else:
                rms = np.sqrt(((a0 - want0) ** 2).sum(1).mean() / (want0 ** 2).sum(1).mean())
                assert rms <= 1e-5, kind
        if prec == "fp64":
            assert F.rel_err(got["cells"][0], got["rows"][0]) <= 1e-13


def test_pair_kernel_switch_mid_run(gpu):
    """Switching to the cell-staged kernel after setup: it takes over once a
    list build has written the staging positions (the row kernel runs until
    then); the 20-step trajectory matches the oracle's."""
    s, ref, meta = F.load("ions2000_setup")
    print("[case] pair kernel switch mid-run", flush=True)
    o = Oracle(s)
    o.setup()
    assert o.run(0.002, 20) == 0
    e = Engine(s.copy())
    e.set_pair_kernel("rows")
    e.setup()
    e.step(0.002, 5)
    e.set_pair_kernel("cells")
    e.step(0.002, 15)
    assert e.pair_kernel() == "cells"
    assert F.rel_err(e.get("acc"), o.get("acc")) <= 1e-9
    assert F.rel_err(e.get("pos"), o.get("pos")) <= 1e-9


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_slot_rows_and_index_rows(gpu, prec):
    """The default pair mode's list builds write the entries as bin-order slots
    only; the atom-index rows are filled from them when something reads them.
    Two engines on the same system, one in the default mode and one gathering
    by atom index throughout, step 12 times (every step rebuilds the list).
    Then every reader of the index rows sees the same rows and gives the same
    numbers: the pair list, the list statistics, the LJ + Ewald_r energy
    (cell walk) and the adjoint; the accelerations of the run are bit-identical,
    and so are the row kernel's accelerations on the filled rows after a
    switch to CAV_PAIRS_ROWS without a rebuild."""
    s = configs.ions(n=3000, rho=0.3, kmax=6, seed=5)
    print(f"[case] slot rows vs index rows {prec}", flush=True)
    a = Engine(s.copy(), precision=prec)
    b = Engine(s.copy(), precision=prec)
    b.set_pair_kernel("rows")
    for e in (a, b):
        e.setup()
        e.step(0.002, 12)
    assert np.array_equal(a.get("acc"), b.get("acc"))
    assert np.array_equal(a.get("pos"), b.get("pos"))
    assert np.array_equal(F.sorted_pairs(a.pairs()), F.sorted_pairs(b.pairs()))
    assert a.list_stats() == b.list_stats()
    ea, eb = a.energy(k=False), b.energy(k=False)
    assert ea[0] == eb[0] and ea[1] == eb[1]
    w = np.random.default_rng(3).standard_normal((len(s.pos), 3))
    ga, gb = a.adjoint(w), b.adjoint(w)
    for x, y in zip(ga, gb):
        assert np.array_equal(x, y)
    # the last build wrote slots only: the row kernel reads the filled rows
    a.set_pair_kernel("rows")
    a.reset_acceleration()
    a.forces("lj", "ewald_r")
    b.reset_acceleration()
    b.forces("lj", "ewald_r")
    assert np.array_equal(a.get("acc"), b.get("acc"))


def test_charges_changed_after_setup(gpu):
    """New type charges set after setup reach the owned atoms and the current
    ghosts before the next force, energy and point-field call (cav_forces on
    the existing ghosts and list, no new exchange), as a freshly set-up engine
    with those charges computes them."""
    s = configs.ions(n=900, kmax=4, seed=21)
    print("[case] charges changed after setup", flush=True)
    e = Engine(s.copy())
    e.setup()
    newq = np.array([0.7, -0.7])
    e.set_parameters(charge=newq)
    e.reset_acceleration()
    e.reset_virial()
    e.forces("ewald_r", "ewald_k")
    t = s.copy()
    t.charge = newq
    f = Engine(t)
    f.setup()
    f.reset_acceleration()
    f.reset_virial()
    f.forces("ewald_r", "ewald_k")
    assert F.rel_err(e.get("acc"), f.get("acc")) <= 1e-12
    assert abs(e.virial() - f.virial()) <= 1e-12 * abs(f.virial())
    _, er0, ek0 = e.energy(lj=False)
    _, er1, ek1 = f.energy(lj=False)
    assert abs(er0 - er1) <= 1e-12 * abs(er1)
    assert abs(ek0 - ek1) <= 1e-12 * abs(ek1)
    pts = np.random.default_rng(3).random((16, 3)) * s.size + s.lower
    for which in ("r", "k"):
        a, pa = e.point_field(which, pts)
        b, pb = f.point_field(which, pts)
        assert F.rel_err(a, b) <= 1e-12 and F.rel_err(pa, pb) <= 1e-12
