"""The C++ drop-in layer: the reference-named objects, the CASL reader and the
caviar_amd_run driver (caviar_amd/host/).

CPU tests read the fixtures' own CASL scripts (the scripts the reference ran
to produce them, tests/golden/make_golden.py) with --parse-only and check every
setting the driver collected against the fixture's recorded settings, plus the
reader's handling of variables, expressions, comments, quoting and errors.
GPU tests run those scripts end to end through caviar_amd_run and compare its
dump with the reference's outputs held in the fixtures, to the same bars as
the C-ABI parity tests: ghosts and Verlet pair sets bit-exact, FP64
accelerations / virial / energies / S(k) / fields within 1e-12 relative,
trajectories after tens of steps within 1e-9.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import fixtures as F

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402  (CASL / XYZ writers and the dump parser)

ROOT = os.path.dirname(HERE)
DRIVER = os.path.join(ROOT, "caviar_amd", "caviar_amd_run")


def needs_driver():
    if not os.path.exists(DRIVER):
        pytest.fail("caviar_amd_run is not built (python -c 'import __graft_entry__ as g; g.build()')")


def write_case(tmp_path, name):
    s, ref, meta = F.load(name)
    make_golden.write_xyz(str(tmp_path / "in.xyz"), s)
    script = tmp_path / "in.casl"
    script.write_text(make_golden.casl_for(s, "in.xyz", meta["run_steps"], meta["dt"]))
    return s, ref, meta, str(script)


def molecule_args(tmp_path, s, mode):
    """With molecule indices (set by the reference harness from a file before
    its run), the driver gets the same file and runs after setting them."""
    if s.mol is None:
        return [], mode
    f = tmp_path / "mol.txt"
    np.savetxt(f, s.mol, fmt="%d")
    return ["--molecules", str(f)], ("mrun" if mode == "run" else mode)


def run_driver(*args, cwd=None):
    return subprocess.run([DRIVER, *args], capture_output=True, text=True, cwd=cwd, timeout=600)


@pytest.mark.parametrize("name", F.names())
def test_parse_only_collects_the_fixture_settings(tmp_path, name):
    needs_driver()
    s, ref, meta, script = write_case(tmp_path, name)
    r = run_driver("--parse-only", script)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert d["lower"] == list(map(float, s.lower)) and d["upper"] == list(map(float, s.upper))
    assert d["bc"] == list(map(int, s.bc))
    assert d["n"] == s.n and d["ghost_cutoff"] == s.ghost_cutoff
    assert d["mass"] == list(map(float, s.mass)) and d["charge"] == list(map(float, s.charge))
    ps = vs = 0.0
    ts = 0
    for i in range(s.n):
        p, v = s.pos[i], s.vel[i]
        ps += float(p[0]) + 2.0 * float(p[1]) + 3.0 * float(p[2])
        vs += float(v[0]) + 2.0 * float(v[1]) + 3.0 * float(v[2])
        ts += int(s.types[i]) * (i % 7 + 1)
    assert d["pos_checksum"] == ps and d["vel_checksum"] == vs and d["type_checksum"] == ts
    nl = s.neighbor
    assert d["neighbor"] == {"cutoff": nl.cutoff, "cutoff_extra_coef": nl.cutoff_extra_coef,
                             "from_cells": bool(nl.from_cells), "build_cells": bool(nl.build_cells),
                             "all_atom_test": bool(nl.all_atom_test), "rebuild_test": bool(nl.rebuild_test)}
    if s.lj is not None:
        assert d["lj"]["cutoff"] == s.lj.cutoff
        assert np.array_equal(np.array(d["lj"]["epsilon"]), np.asarray(s.lj.epsilon))
        assert np.array_equal(np.array(d["lj"]["sigma"]), np.asarray(s.lj.sigma))
    else:
        assert "lj" not in d
    if s.lj_list is not None:
        g, w = d["lj_list"], s.lj_list
        assert g["cutoff"] == w.cutoff and g["wca"] == w.wca
        assert g["ignore_intra_molecule"] == w.ignore_intra_molecule and g["jump_tol"] == w.jump_tol
        nt = s.ntypes

        def dense(m, fill):
            out = np.full((nt, nt), fill)
            for i, row in enumerate(m):
                out[i, :len(row)] = row
            return out
        for key, fill in (("epsilon", 0.0), ("sigma", 0.0), ("cutoff_list", 0.0), ("lambda_s", 1.0),
                          ("lambda_e", 1.0)):
            want = getattr(w, key)
            if want is None:
                assert g[key] == [], key
            else:
                assert np.array_equal(dense(g[key], fill), np.asarray(want)), key
        for key in ("epsilon_atom", "sigma_atom"):
            want = getattr(w, key)
            assert (g[key] == []) if want is None else np.array_equal(g[key], want), key
    if s.short_range is not None:
        g, w = d["short_range"], s.short_range
        assert (g["cutoff"], g["k_electrostatic"], g["beta"]) == (w.cutoff, w.k_electrostatic, w.beta)
        assert g["external_field"] == list(map(float, w.external_field))
        if w.lambda_ is None:
            assert g["lambda_"] is None
        else:
            assert np.array_equal(np.array(g["lambda_"]), np.asarray(w.lambda_))
    if s.ewald_r is not None:
        e = d["ewald_r"]
        assert e["alpha"] == s.ewald_r.alpha and e["k_electrostatic"] == s.ewald_r.k_electrostatic
        if s.ewald_r.lambda_ is None:
            assert e["lambda_"] is None
        else:
            assert np.array_equal(np.array(e["lambda_"]), np.asarray(s.ewald_r.lambda_))
    if s.ewald_k is not None:
        e, k = d["ewald_k"], s.ewald_k
        assert (e["alpha"], e["k_electrostatic"]) == (k.alpha, k.k_electrostatic)
        assert (e["kx_max"], e["ky_max"], e["kz_max"]) == (k.kx_max, k.ky_max, k.kz_max)
        assert e["epsilon_dipole"] == k.epsilon_dipole
    if s.elc is not None:
        assert d["elc"] == vars(s.elc)
    assert d["dt"] == meta["dt"] and d["final_step"] == meta["run_steps"]
    if s.random_velocity is not None:
        # add_random_velocity ran in the script: the driver's velocities are the
        # reference's bit for bit (checksum above, formed from s.vel = those)
        assert d["vel_checksum"] == vs


EXAMPLE = """
#===== variables (the reference's examples/e2-periodic style)
real cutoff_max 3.5
real half = cutoff_max / 2 - 0.25   # an expression
int  nk 3

domain box dom
dom xmin -6 xmax +6 ymin -6
dom ymax 6 zmin -6.0 zmax 6e0
dom boundary_condition 1 1 1
dom generate

atom_data basic adata
adata ghost_cutoff cutoff_max
adata set_domain dom
adata add_type_mass 0 1.0
adata add_type_charge 0 1
adata add_type_mass 1 2
adata add_type_charge 1 -1
adata add_xyz_data_file file_name 'atoms.xyz' read_velocity

neighborlist basic neigh
neigh set_atom_data adata
neigh set_domain dom
neigh cutoff half cutoff_extra_coef 1.1
neigh make_verlet_list_from_cell_list build_cell_list dt 0.01

force_field lj_cell_list f_lj
f_lj cutoff 1.5 epsilon 0 0 1.0 epsilon 1 1 0.5
f_lj sigma 0 0 1.0 sigma 1 1 1.2
f_lj set_neighborlist neigh set_atom_data adata

force_field electrostatic_ewald_k f_ek
f_ek set_atom_data adata set_domain dom
f_ek k_max nk alpha 0.9 k_electrostatic 2
f_ek kz_max 4

md_simulator basic sim
sim set_atom_data adata
sim integrator_type velocity_verlet
sim add_force_field f_lj
sim add_force_field f_ek
sim add_neighborlist neigh
sim initial_step 0 final_step 5 dt 0.002
sim run
exit
this line is never read
"""


def test_reader_variables_expressions_and_ragged_matrices(tmp_path):
    needs_driver()
    (tmp_path / "atoms.xyz").write_text("3\ncomment line\n0 0 0 0 0.1 0 0\n1 1.5 0 0 0 -0.1 0\n"
                                        "0 -1.5 +2 1e-1 0 0 .5\n")
    (tmp_path / "ex.casl").write_text(EXAMPLE)
    r = run_driver("--parse-only", str(tmp_path / "ex.casl"))
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert d["lower"] == [-6.0] * 3 and d["upper"] == [6.0] * 3 and d["bc"] == [1, 1, 1]
    assert d["n"] == 3 and d["ghost_cutoff"] == 3.5
    assert d["mass"] == [1.0, 2.0] and d["charge"] == [1.0, -1.0]
    assert d["pos_checksum"] == (0.0) + (1.5) + (-1.5 + 2.0 * 2.0 + 3.0 * 0.1)
    assert d["vel_checksum"] == (0.1) + (2.0 * -0.1) + (3.0 * 0.5)
    assert d["neighbor"]["cutoff"] == 3.5 / 2 - 0.25 and d["neighbor"]["cutoff_extra_coef"] == 1.1
    assert d["neighbor"]["from_cells"] and d["neighbor"]["build_cells"]
    # only the entries written are set; rows grow on demand with 0 (row 1 has [0, 0.5])
    assert d["lj"]["epsilon"] == [[1.0], [0.0, 0.5]]
    assert d["lj"]["sigma"] == [[1.0], [0.0, 1.2]]
    assert (d["ewald_k"]["kx_max"], d["ewald_k"]["ky_max"], d["ewald_k"]["kz_max"]) == (3, 3, 4)
    assert d["final_step"] == 5 and d["dt"] == 0.002


@pytest.mark.parametrize("script, message", [
    ("foo bar\n", "Invalid command or object name: foo"),
    ("force_field lj_gpu f\n", "Undefined 'force_field::lj_gpu'"),
    ("domain box d\nd xmin -1 wrong 2\n", "Undefined variable or command: wrong"),
    ("neighborlist basic n\nn cutoff -1\n", "cutoff have to be positive"),
    ("neighborlist basic n\nn cutoff undefined_name\n", "Invalid object NAME"),
    ("int k 2.5\n", "non-integral"),
    ("domain box d\ndomain box d\n", "NAME is already used: d"),
    ("force_field electrostatic_ewald_k e\ne k_max 1.5\n", "Expected an int number"),
    ("force_field electrostatic_ewald_k e\ne slab_normal_axis w\n", "expected 'x' or 'y' or 'z'"),
    ("md_simulator basic s\ns integrator_type leap_frog\n", "only velocity_verlet"),
])
def test_reader_reports_errors_like_the_reference(tmp_path, script, message):
    needs_driver()
    (tmp_path / "bad.casl").write_text(script)
    r = run_driver("--parse-only", str(tmp_path / "bad.casl"))
    assert r.returncode == 1
    assert message in r.stderr, r.stderr


@pytest.mark.parametrize("kind, cells", [("basic", False), ("cell_list", True), ("verlet_list", False)])
def test_neighborlist_mode_aliases(tmp_path, kind, cells):
    needs_driver()
    s, ref, meta, script = write_case(tmp_path, "ions600_brute_setup")
    text = open(script).read().replace("neighborlist basic neigh", f"neighborlist {kind} neigh")
    (tmp_path / "alias.casl").write_text(text)
    r = run_driver("--parse-only", str(tmp_path / "alias.casl"))
    assert r.returncode == 0, r.stderr
    nl = json.loads(r.stdout)["neighbor"]
    assert nl["build_cells"] == (cells or s.neighbor.build_cells)
    assert nl["from_cells"] == (cells or s.neighbor.from_cells)


def test_missing_settings_fail_at_setup(tmp_path):
    """Ewald_k without alpha (no default in the reference) and a force field
    without its neighbour list are refused when the run sets up."""
    needs_driver()
    s, ref, meta, script = write_case(tmp_path, "small200_run10")
    text = open(script).read()
    bad = text.replace("f_ek alpha", "# f_ek alpha")
    (tmp_path / "noalpha.casl").write_text(bad)
    r = run_driver("--parse-only", str(tmp_path / "noalpha.casl"))
    assert r.returncode == 1 and "alpha is not set" in r.stderr, r.stderr


# ----------------------------------------------------------------- GPU ------

def check_dump(out, s, ref, meta, exact_tol=1e-12):
    got = make_golden.parse(out)
    assert got["N"] == s.n
    run = meta["run_steps"] > 0
    tol = 1e-9 if run else exact_tol
    for key in ("pos", "vel", "acc"):
        assert F.rel_err(got[key], ref[key]) <= tol, key
    assert np.array_equal(F.sorted_pairs(got["pairs"]), F.sorted_pairs(ref["pairs"]))
    if not run:
        assert np.array_equal(got["ghost_pos"], ref["ghost_pos"])
        assert np.array_equal(got["ghost_id"], ref["ghost_id"])
        v = float(ref["virial"])
        assert abs(got["virial"] - v) <= 1e-12 * max(abs(v), 1e-300) or v == got["virial"]
    for key in ("energy_r", "energy_k", "energy_k_self", "energy_elc", "kinetic_energy", "temperature",
                "pressure"):
        if key in ref:
            assert abs(got[key] - ref[key]) <= 1e-12 * abs(ref[key]), key
    for key in ("sk", "field_r", "field_k", "field_elc"):
        if key in ref:
            assert F.rel_err(got[key], ref[key]) <= 1e-12, key


@pytest.mark.gpu
@pytest.mark.parametrize("name", F.names())
def test_driver_reproduces_the_reference_run(gpu, tmp_path, name):
    needs_driver()
    s, ref, meta, script = write_case(tmp_path, name)
    print(f"[case] caviar_amd_run {name} mode={meta['mode']}", flush=True)
    out = str(tmp_path / "out.txt")
    extra, mode = molecule_args(tmp_path, s, meta["mode"])
    r = run_driver(*extra, script, mode, out, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    check_dump(out, s, ref, meta)
    check_writer_files(tmp_path, ref, meta)


def check_writer_files(tmp_path, ref, meta):
    """Files the script's writer wrote, against the reference's: the same
    lines and tokens, numbers (printed with 6 significant digits) within one
    unit in their last printed digit."""
    for key in ref:
        if not key.startswith("file_"):
            continue
        name = key[5:].rsplit("_", 1)
        path = tmp_path / (name[0] + "." + name[1])
        assert path.exists(), path
        got = path.read_text().split("\n")
        want = str(ref[key]).split("\n")
        assert len(got) == len(want), key
        for g, w in zip(got, want):
            gt, wt = g.split(), w.split()
            assert len(gt) == len(wt), (key, g, w)
            for a, b in zip(gt, wt):
                if a == b:
                    continue
                x, y = float(a), float(b)
                assert abs(x - y) <= 1e-5 * max(abs(y), 1e-300) + 1e-300, (key, g, w)


@pytest.mark.gpu
def test_driver_mixed_precision_force_error(gpu, tmp_path):
    needs_driver()
    s, ref, meta, script = write_case(tmp_path, "ions2000_setup")
    out = str(tmp_path / "out.txt")
    r = run_driver("--precision", "mixed", script, "setup", out)
    assert r.returncode == 0, r.stderr
    acc = make_golden.parse(out)["acc"]
    err = np.sqrt(((acc - ref["acc"]) ** 2).sum(1).mean() / (ref["acc"] ** 2).sum(1).mean())
    assert err <= 1e-5
