"""caviar_amd: MI355X-native short-range + Ewald force path of CAVIAR.

The compute path is the HIP library ``caviar_amd/libcaviar_amd.so`` (gfx950),
reached through its C ABI (``include/caviar_amd.h``) with ctypes.  There is no
CPU fallback: importing works without a GPU, but creating an :class:`Engine`
fails loudly when the library or a device is missing.

:class:`Engine` is one device-resident single-domain system.  The C++ classes
with the reference's own names and CASL commands (Domain, Atom_data,
Neighborlist, force_field::Lj_cell_list / Electrostatic_ewald_r /
Electrostatic_ewald_k, Md_simulator) and the ``caviar_amd_run`` script driver
are in ``caviar_amd/host`` on the same C ABI.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .system import (BerendsenSettings, ELCSettings, EwaldKSettings,  # noqa: F401
                     EwaldRSettings, LJListSettings, LJSettings, NeighborSettings, NVESettings,
                     ShortRangeSettings, System, ThermoSettings, lj_list_args, short_range_args)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcaviar_amd.so")

CAV_FP64 = 0
CAV_MIXED = 1
FORCE_FIELDS = {"lj": 1, "ewald_r": 2, "ewald_k": 4, "lj_list": 8, "short_range": 16, "elc": 32}
_ARRAYS = {"pos": 0, "vel": 1, "acc": 2, "acc_old": 3}
_ERRORS = {1: "invalid argument", 2: "HIP runtime error", 3: "call out of order",
           4: "NaN acceleration", 5: "capacity"}

_lib = None


class CaviarError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"caviar_amd error {code} ({_ERRORS.get(code, '?')}): {msg}")
        self.code = code


class NaNAccelerationError(CaviarError):
    pass


def load_library():
    """Load libcaviar_amd.so (built in-tree by ``caviar_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m caviar_amd.build` "
                           "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    I = ctypes.c_int
    D = ctypes.c_double
    L = ctypes.c_long
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "cav_create": [I, ctypes.POINTER(P)],
        "cav_destroy": [P],
        "cav_set_precision": [P, I],
        "cav_set_box": [P, dp, dp, ip],
        "cav_set_ghost_cutoff": [P, D],
        "cav_set_types": [P, I, dp, dp],
        "cav_set_neighbor": [P, D, D, I, I, I, I],
        "cav_set_lj": [P, D, dp, dp],
        "cav_set_ewald_r": [P, D, D, dp],
        "cav_set_lj_list": [P, D, dp, dp, dp, dp, dp, I, dp, dp, I, I, D],
        "cav_set_short_range": [P, D, D, D, dp, dp],
        "cav_set_elc": [P, I, I, I, D, D],
        "cav_set_thermo": [P, D, I, I, I],
        "cav_set_berendsen": [P, D, D, D, I, D, D],
        "cav_set_nve": [P, D, D, D, D, D, I],
        "cav_set_step_index": [P, L],
        "cav_thermo": [P, dp, dp, dp],
        "cav_energy_elc": [P, dp],
        "cav_set_ewald_k": [P, D, D, I, I, I, I, D, I, D],
        "cav_upload_atoms": [P, I, ip, ip, ip, dp, dp],
        "cav_download": [P, I, dp],
        "cav_upload": [P, I, dp],
        "cav_num_atoms": [P, ip],
        "cav_setup": [P],
        "cav_step": [P, D, I],
        "cav_recalculate": [P, I],
        "cav_neighbor_build": [P, I, ip],
        "cav_num_ghosts": [P, ip],
        "cav_get_ghosts": [P, dp, ip],
        "cav_num_pairs": [P, ctypes.POINTER(L)],
        "cav_get_pairs": [P, ip, L, ctypes.POINTER(L)],
        "cav_virial": [P, dp],
        "cav_num_kvectors": [P, ip],
        "cav_get_sk": [P, dp],
        "cav_energy": [P, dp, dp, dp],
        "cav_point_field": [P, I, I, dp, dp, dp],
        "cav_list_stats": [P, dp, dp],
        "cav_kernel_time": [P, ctypes.c_char_p, dp, ip],
        "cav_enable_timing": [P, I],
        "cav_adjoint": [P, dp, dp, dp, dp],
        "cav_adjoint_noalloc": [P, dp, dp, dp, dp],
        "cav_exchange": [P, ip],
        "cav_batch_create": [I, I, ctypes.POINTER(P)],
        "cav_batch_destroy": [P],
        "cav_batch_set_types": [P, I, dp, dp],
        "cav_batch_set_neighbor": [P, D, D, D],
        "cav_batch_set_lj": [P, D, dp, dp],
        "cav_batch_set_ewald_r": [P, D, D],
        "cav_batch_set_ewald_k": [P, D, D, I, I, I],
        "cav_batch_add_replica": [P, dp, dp, I, ip, dp, dp, ip],
        "cav_batch_num_replicas": [P, ip],
        "cav_batch_num_atoms": [P, I, ip],
        "cav_batch_setup": [P],
        "cav_batch_step": [P, D, I, I],
        "cav_batch_download": [P, I, I, dp],
        "cav_batch_virial": [P, I, dp],
        "cav_batch_get_pairs": [P, I, ip, L, ctypes.POINTER(L)],
        "cav_batch_autotune": [P],
        "cav_batch_tiles": [P, ip, ip, ip, ip],
        "cav_batch_adjoint": [P, dp, dp, dp, dp],
        "cav_autotune": [P],
        "cav_tiles": [P, ip, ip, ip, ip, ip],
        "cav_set_tiles": [P, I, I, I, I],
        "cav_set_overlap": [P, I],
        "cav_set_pair_kernel": [P, I],
        "cav_pair_kernel": [P, ip],
        "cav_reset_acceleration": [P],
        "cav_reset_virial": [P],
        "cav_forces": [P, I],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = I
    lib.cav_last_error.argtypes = [P]
    lib.cav_last_error.restype = ctypes.c_char_p
    lib.cav_adjoint_noalloc_scratch_bytes.argtypes = []
    lib.cav_adjoint_noalloc_scratch_bytes.restype = ctypes.c_size_t
    lib.cav_batch_last_error.argtypes = [P]
    lib.cav_batch_last_error.restype = ctypes.c_char_p
    _lib = lib
    return lib


def exported_symbols() -> list:
    """Names of the C-ABI entry points declared in include/caviar_amd.h."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "caviar_amd.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(cav_[a-z_]+)\s*\(", txt)))


def _dp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class Engine:
    """One single-domain CAVIAR system resident on one GPU.

    Mirrors the reference's Md_simulator + Atom_data + Neighborlist + force-field
    objects for the LJ / Ewald path (md_simulator.cpp:361-617).
    """

    def __init__(self, system: System, precision: str = "fp64", device: int = 0,
                 timing: bool = False):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.cav_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise CaviarError(rc, "cav_create failed (no usable HIP device?)")
        self.h = h
        self.system = system
        self.n = system.n
        self._check(self.lib.cav_set_precision(self.h, CAV_FP64 if precision == "fp64" else CAV_MIXED))
        self.precision = precision
        s = system
        self._check(self.lib.cav_set_box(self.h, _dp(s.lower.copy()), _dp(s.upper.copy()), _ip(s.bc.copy())))
        self._check(self.lib.cav_set_ghost_cutoff(self.h, float(s.ghost_cutoff)))
        self._check(self.lib.cav_set_types(self.h, s.ntypes, _dp(s.mass.copy()), _dp(s.charge.copy())))
        nt = s.ntypes
        if s.neighbor is not None:
            nl = s.neighbor
            self._check(self.lib.cav_set_neighbor(self.h, nl.cutoff, nl.cutoff_extra_coef,
                                                  int(nl.from_cells), int(nl.build_cells),
                                                  int(nl.all_atom_test), int(nl.rebuild_test)))
        if s.lj is not None:
            eps = np.ascontiguousarray(s.lj.epsilon, dtype=np.float64).reshape(nt * nt)
            sig = np.ascontiguousarray(s.lj.sigma, dtype=np.float64).reshape(nt * nt)
            self._check(self.lib.cav_set_lj(self.h, s.lj.cutoff, _dp(eps), _dp(sig)))
        if s.lj_list is not None:
            a = lj_list_args(s.lj_list, nt)
            self._check(self.lib.cav_set_lj_list(self.h, a[0], *[_dp(x) for x in a[1:6]], a[6],
                                                 _dp(a[7]), _dp(a[8]), a[9], a[10], a[11]))
        if s.short_range is not None:
            a = short_range_args(s.short_range, nt)
            self._check(self.lib.cav_set_short_range(self.h, a[0], a[1], a[2], _dp(a[3]), _dp(a[4])))
        if s.ewald_r is not None:
            lam = s.ewald_r.lambda_
            lam = None if lam is None else np.ascontiguousarray(lam, dtype=np.float64).reshape(nt * nt)
            self._check(self.lib.cav_set_ewald_r(self.h, s.ewald_r.alpha, s.ewald_r.k_electrostatic,
                                                 _dp(lam)))
        if s.ewald_k is not None:
            e = s.ewald_k
            self._check(self.lib.cav_set_ewald_k(
                self.h, e.alpha, e.k_electrostatic, e.kx_max, e.ky_max, e.kz_max,
                int(e.epsilon_dipole is not None), float(e.epsilon_dipole or 0.0),
                -1 if e.slab_normal_axis is None else int(e.slab_normal_axis), float(e.slab_lz)))
        if s.elc is not None:
            c = s.elc
            self._check(self.lib.cav_set_elc(self.h, int(c.slab_normal_axis), int(c.kx_max),
                                             int(c.ky_max), float(c.lz), float(c.k_electrostatic)))
        if s.thermo is not None:
            t = s.thermo
            self._check(self.lib.cav_set_thermo(self.h, float(t.k_b), int(t.n_r_df),
                                                int(t.temperature_process), int(t.pressure_process)))
        if s.berendsen is not None:
            b = s.berendsen
            self._check(self.lib.cav_set_berendsen(self.h, b.temperature, b.coupling, b.dt, int(b.step),
                                                   b.lambda_min, b.lambda_max))
        if s.nve is not None:
            v = s.nve
            self._check(self.lib.cav_set_nve(self.h, v.energy_per_dof, v.energy_tot, v.kbt, v.kb,
                                             v.temperature, int(v.step)))
        ids = np.ascontiguousarray(s.ids, dtype=np.int32)
        typ = np.ascontiguousarray(s.types, dtype=np.int32)
        mol = None if s.mol is None else np.ascontiguousarray(s.mol, dtype=np.int32)
        pos = np.ascontiguousarray(s.pos, dtype=np.float64)
        vel = np.ascontiguousarray(s.vel, dtype=np.float64)
        self._check(self.lib.cav_upload_atoms(self.h, s.n, _ip(ids), _ip(typ), _ip(mol), _dp(pos), _dp(vel)))
        if timing:
            self._check(self.lib.cav_enable_timing(self.h, 1))

    def _check(self, rc: int):
        if rc != 0:
            msg = self.lib.cav_last_error(self.h).decode()
            if rc == 4:
                raise NaNAccelerationError(rc, msg)
            raise CaviarError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.cav_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- MD path
    def setup(self):
        self._check(self.lib.cav_setup(self.h))

    def step(self, dt: float, nsteps: int = 1):
        self._check(self.lib.cav_step(self.h, float(dt), int(nsteps)))

    def recalculate(self, force_rebuild: bool = False):
        self._check(self.lib.cav_recalculate(self.h, int(force_rebuild)))

    def neighbor_build(self, update: bool) -> bool:
        r = ctypes.c_int()
        self._check(self.lib.cav_neighbor_build(self.h, int(update), ctypes.byref(r)))
        return bool(r.value)

    # ---- tuning
    def autotune(self) -> dict:
        """Pick the Ewald_k tile shapes by timing the candidates on this system."""
        self._check(self.lib.cav_autotune(self.h))
        return self.tiles()

    def tiles(self) -> dict:
        v = [ctypes.c_int() for _ in range(5)]
        self._check(self.lib.cav_tiles(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("s_rows", "s_stage", "s_chunk", "f_tiles", "f_waves"), (x.value for x in v)))

    def set_tiles(self, s_rows: int, s_chunk: int, f_tiles: int = 1, f_waves: int = 16) -> dict:
        """Pin the Ewald_k tiles instead of tuning them."""
        self._check(self.lib.cav_set_tiles(self.h, int(s_rows), int(s_chunk), int(f_tiles),
                                           int(f_waves)))
        return self.tiles()

    def set_overlap(self, on: bool):
        """Ewald_k of a step concurrent with the short-range passes (default on)."""
        self._check(self.lib.cav_set_overlap(self.h, int(bool(on))))

    PAIR_KERNELS = {"auto": 0, "rows": 1, "cells": 2}

    def set_pair_kernel(self, kind: str):
        """'auto' (cell-staged where it applies), 'rows' or 'cells'."""
        self._check(self.lib.cav_set_pair_kernel(self.h, self.PAIR_KERNELS[kind]))

    def pair_kernel(self) -> str:
        v = ctypes.c_int()
        self._check(self.lib.cav_pair_kernel(self.h, ctypes.byref(v)))
        return {1: "rows", 2: "cells"}[v.value]

    # ---- the pieces of a step, one object at a time
    def exchange(self) -> bool:
        """Wrap owned atoms and rebuild the ghosts; True if any atom wrapped."""
        w = ctypes.c_int()
        self._check(self.lib.cav_exchange(self.h, ctypes.byref(w)))
        return bool(w.value)

    def reset_acceleration(self) -> None:
        self._check(self.lib.cav_reset_acceleration(self.h))

    def reset_virial(self) -> None:
        self._check(self.lib.cav_reset_virial(self.h))

    def forces(self, *fields: str) -> None:
        """Add the accelerations (and virial) of the named force fields
        ("lj", "ewald_r", "ewald_k"), as each Force_field::calculate_acceleration does."""
        mask = 0
        for f in fields:
            mask |= FORCE_FIELDS[f]
        self._check(self.lib.cav_forces(self.h, mask))

    # ---- state
    def get(self, what: str) -> np.ndarray:
        out = np.empty((self.n, 3), dtype=np.float64)
        self._check(self.lib.cav_download(self.h, _ARRAYS[what], _dp(out)))
        return out

    def set(self, what: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.n, 3)
        self._check(self.lib.cav_upload(self.h, _ARRAYS[what], _dp(a)))

    def ghosts(self):
        g = ctypes.c_int()
        self._check(self.lib.cav_num_ghosts(self.h, ctypes.byref(g)))
        gp = np.empty((g.value, 3))
        gs = np.empty(g.value, dtype=np.int32)
        if g.value:
            self._check(self.lib.cav_get_ghosts(self.h, _dp(gp), _ip(gs)))
        return gp, gs

    def pairs(self) -> np.ndarray:
        npairs = ctypes.c_long()
        self._check(self.lib.cav_num_pairs(self.h, ctypes.byref(npairs)))
        out = np.empty((max(npairs.value, 1), 2), dtype=np.int32)
        got = ctypes.c_long()
        self._check(self.lib.cav_get_pairs(self.h, _ip(out), npairs.value, ctypes.byref(got)))
        return out[:got.value]

    def virial(self) -> float:
        v = ctypes.c_double()
        self._check(self.lib.cav_virial(self.h, ctypes.byref(v)))
        return v.value

    def sk(self) -> np.ndarray:
        nk = ctypes.c_int()
        self._check(self.lib.cav_num_kvectors(self.h, ctypes.byref(nk)))
        out = np.empty((nk.value, 2))
        self._check(self.lib.cav_get_sk(self.h, _dp(out)))
        return out[:, 0] + 1j * out[:, 1]

    def energy(self, lj=True, r=True, k=True):
        vals = [ctypes.c_double() if f else None for f in (lj, r, k)]
        args = [ctypes.byref(v) if v is not None else None for v in vals]
        self._check(self.lib.cav_energy(self.h, *args))
        return tuple(v.value if v is not None else None for v in vals)

    def point_field(self, which: str, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        f = np.empty_like(pts)
        p = np.empty(len(pts))
        self._check(self.lib.cav_point_field(self.h, {"r": 0, "k": 1, "elc": 2}[which], len(pts),
                                             _dp(pts), _dp(f), _dp(p)))
        return f, p

    def thermo(self):
        """(kinetic energy now, temperature and pressure as last finalized)."""
        v = [ctypes.c_double() for _ in range(3)]
        self._check(self.lib.cav_thermo(self.h, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def set_step_index(self, step: int):
        self._check(self.lib.cav_set_step_index(self.h, int(step)))

    def energy_elc(self) -> float:
        """Electrostatic_ewald_slab_correction::energy of the last force evaluation."""
        v = ctypes.c_double()
        self._check(self.lib.cav_energy_elc(self.h, ctypes.byref(v)))
        return v.value

    def set_parameters(self, epsilon=None, sigma=None, charge=None, mass=None):
        """Change force-field parameter values after setup (same structure);
        they take effect at the next recalculate()/step()."""
        s = self.system
        nt = s.ntypes
        if charge is not None or mass is not None:
            if charge is not None:
                s.charge = np.asarray(charge, dtype=np.float64).reshape(nt)
            if mass is not None:
                s.mass = np.asarray(mass, dtype=np.float64).reshape(nt)
            self._check(self.lib.cav_set_types(self.h, nt, _dp(s.mass.copy()), _dp(s.charge.copy())))
        if epsilon is not None or sigma is not None:
            if epsilon is not None:
                s.lj.epsilon = np.asarray(epsilon, dtype=np.float64).reshape(nt, nt)
            if sigma is not None:
                s.lj.sigma = np.asarray(sigma, dtype=np.float64).reshape(nt, nt)
            eps = np.ascontiguousarray(s.lj.epsilon, dtype=np.float64).reshape(nt * nt)
            sig = np.ascontiguousarray(s.lj.sigma, dtype=np.float64).reshape(nt * nt)
            self._check(self.lib.cav_set_lj(self.h, s.lj.cutoff, _dp(eps), _dp(sig)))

    def adjoint(self, weights):
        """g = sum_i w_i . dF_i/dtheta for the LJ epsilon/sigma matrices and the
        type charges at the current positions and parameters."""
        nt = self.system.ntypes
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(self.n, 3)
        ge = np.zeros(nt * nt)
        gs = np.zeros(nt * nt)
        gq = np.zeros(nt)
        self._check(self.lib.cav_adjoint(self.h, _dp(w), _dp(ge), _dp(gs), _dp(gq)))
        return ge.reshape(nt, nt), gs.reshape(nt, nt), gq

    def adjoint_noalloc(self, weights):
        """adjoint() through cav_adjoint_noalloc: the same result, with the
        workspace in the library's fixed per-GPU scratch array (no device
        allocation in the call)."""
        nt = self.system.ntypes
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(self.n, 3)
        ge = np.zeros(nt * nt)
        gs = np.zeros(nt * nt)
        gq = np.zeros(nt)
        self._check(self.lib.cav_adjoint_noalloc(self.h, _dp(w), _dp(ge), _dp(gs), _dp(gq)))
        return ge.reshape(nt, nt), gs.reshape(nt, nt), gq

    def list_stats(self):
        a = ctypes.c_double()
        b = ctypes.c_double()
        self._check(self.lib.cav_list_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def kernel_time(self, name: str):
        ms = ctypes.c_double()
        nl = ctypes.c_int()
        self._check(self.lib.cav_kernel_time(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(nl)))
        return ms.value, nl.value

    def enable_timing(self, on: bool = True):
        self._check(self.lib.cav_enable_timing(self.h, int(on)))


class Batch:
    """Many independent systems of one force field stepped together on one GPU
    (``cav_batch_*``): each replica has its own atoms and periodic box; every
    kernel covers all replicas and one step of all of them is one graph launch.

    The systems must share types, cutoffs and force-field settings (Verlet list
    from cells, Lj_cell_list + Ewald_r + Ewald_k without lambda / dipole / slab).
    """

    def __init__(self, systems, precision: str = "fp64", device: int = 0):
        self.lib = load_library()
        if not systems:
            raise ValueError("no systems")
        s0 = systems[0]
        for s in systems[1:]:
            if (s.ntypes != s0.ntypes or not np.array_equal(s.mass, s0.mass)
                    or not np.array_equal(s.charge, s0.charge) or s.ghost_cutoff != s0.ghost_cutoff
                    or s.neighbor != s0.neighbor):
                raise ValueError("batched systems must share types and neighbour settings")
        h = ctypes.c_void_p()
        rc = self.lib.cav_batch_create(int(device), CAV_FP64 if precision == "fp64" else CAV_MIXED,
                                       ctypes.byref(h))
        if rc != 0:
            raise CaviarError(rc, "cav_batch_create failed (no usable HIP device?)")
        self.h = h
        self.systems = systems
        self.precision = precision
        self._check(self.lib.cav_batch_set_types(h, s0.ntypes, _dp(s0.mass.copy()), _dp(s0.charge.copy())))
        nl = s0.neighbor
        if nl is None or not (nl.from_cells and nl.build_cells) or nl.all_atom_test or nl.rebuild_test:
            raise ValueError("batched replicas use the Verlet list built from cells")
        self._check(self.lib.cav_batch_set_neighbor(h, float(s0.ghost_cutoff), float(nl.cutoff),
                                                    float(nl.cutoff_extra_coef)))
        nt = s0.ntypes
        if s0.lj is not None:
            e = np.ascontiguousarray(s0.lj.epsilon, dtype=np.float64).reshape(nt * nt)
            g = np.ascontiguousarray(s0.lj.sigma, dtype=np.float64).reshape(nt * nt)
            self._check(self.lib.cav_batch_set_lj(h, float(s0.lj.cutoff), _dp(e), _dp(g)))
        if s0.ewald_r is not None:
            if s0.ewald_r.lambda_ is not None:
                raise ValueError("lambda is not supported for batched replicas")
            self._check(self.lib.cav_batch_set_ewald_r(h, s0.ewald_r.alpha, s0.ewald_r.k_electrostatic))
        if s0.ewald_k is not None:
            k = s0.ewald_k
            if k.epsilon_dipole is not None or k.slab_normal_axis is not None:
                raise ValueError("dipole / slab Ewald_k is not supported for batched replicas")
            self._check(self.lib.cav_batch_set_ewald_k(h, k.alpha, k.k_electrostatic, k.kx_max,
                                                       k.ky_max, k.kz_max))
        for s in systems:
            if list(s.bc) != [1, 1, 1]:
                raise ValueError("batched replicas are periodic in every direction")
            typ = np.ascontiguousarray(s.types, dtype=np.int32)
            pos = np.ascontiguousarray(s.pos, dtype=np.float64)
            vel = np.ascontiguousarray(s.vel, dtype=np.float64)
            self._check(self.lib.cav_batch_add_replica(h, _dp(s.lower.copy()), _dp(s.upper.copy()),
                                                       s.n, _ip(typ), _dp(pos), _dp(vel), None))

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.cav_batch_last_error(self.h).decode() if getattr(self, "h", None) else ""
            if rc == 4:
                raise NaNAccelerationError(rc, msg)
            raise CaviarError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            self.lib.cav_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setup(self):
        self._check(self.lib.cav_batch_setup(self.h))

    def step(self, dt: float, nsteps: int = 1, graph: bool = True):
        self._check(self.lib.cav_batch_step(self.h, float(dt), int(nsteps), int(bool(graph))))

    def get(self, replica: int, what: str) -> np.ndarray:
        out = np.empty((self.systems[replica].n, 3), dtype=np.float64)
        self._check(self.lib.cav_batch_download(self.h, int(replica), _ARRAYS[what], _dp(out)))
        return out

    def virial(self, replica: int) -> float:
        v = ctypes.c_double()
        self._check(self.lib.cav_batch_virial(self.h, int(replica), ctypes.byref(v)))
        return v.value

    def pairs(self, replica: int) -> np.ndarray:
        n = ctypes.c_long()
        self._check(self.lib.cav_batch_get_pairs(self.h, int(replica), None, 0, ctypes.byref(n)))
        out = np.empty((n.value, 2), dtype=np.int32)
        self._check(self.lib.cav_batch_get_pairs(self.h, int(replica), _ip(out), n.value, ctypes.byref(n)))
        return out

    def autotune(self) -> dict:
        self._check(self.lib.cav_batch_autotune(self.h))
        return self.tiles()

    def adjoint(self, weights):
        """Force-matching adjoint of every replica: ``weights`` is a list of
        (n_r, 3) arrays; returns a list of (g_eps, g_sigma, g_charge) per replica."""
        nt = self.systems[0].ntypes
        B = len(self.systems)
        w = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1, 3)
                                                 for x in weights]))
        ge = np.zeros(B * nt * nt)
        gs = np.zeros(B * nt * nt)
        gq = np.zeros(B * nt)
        self._check(self.lib.cav_batch_adjoint(self.h, _dp(w), _dp(ge), _dp(gs), _dp(gq)))
        return [(ge[r * nt * nt:(r + 1) * nt * nt].reshape(nt, nt),
                 gs[r * nt * nt:(r + 1) * nt * nt].reshape(nt, nt), gq[r * nt:(r + 1) * nt])
                for r in range(B)]

    def tiles(self) -> dict:
        v = [ctypes.c_int() for _ in range(4)]
        self._check(self.lib.cav_batch_tiles(self.h, *[ctypes.byref(x) for x in v]))
        return dict(zip(("s_rows", "s_stage", "s_chunk", "f_tiles"), (x.value for x in v)))
