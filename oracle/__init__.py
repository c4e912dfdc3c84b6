"""ctypes front-end of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module, and only as the checker: the product path (``caviar_amd``)
never imports, links or calls it.  The C restatement it wraps is
``oracle/caviar_oracle.c``; see that file for the reference citations.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build() -> None:
    """Compile oracle/liboracle.so (gcc, plain C99)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    sig = {
        "orc_create": (_p, []),
        "orc_destroy": (None, [_p]),
        "orc_set_threads": (None, [_i]),
        "orc_set_box": (None, [_p, _dp, _dp, _ip]),
        "orc_set_ghost_cutoff": (None, [_p, _d]),
        "orc_set_types": (None, [_p, _i, _dp, _dp]),
        "orc_set_atoms": (None, [_p, _i, _ip, _ip, _ip, _dp, _dp]),
        "orc_set_neighbor": (None, [_p, _d, _d, _i, _i, _i, _i]),
        "orc_set_lj": (None, [_p, _d, _dp, _dp]),
        "orc_set_ewald_r": (None, [_p, _d, _d, _dp]),
        "orc_set_lj_list": (_i, [_p, _d, _dp, _dp, _dp, _dp, _dp, _i, _dp, _dp, _i, _i, _d]),
        "orc_set_short_range": (None, [_p, _d, _d, _d, _dp, _dp]),
        "orc_lj_list_acc": (None, [_p]),
        "orc_set_elc": (None, [_p, _i, _i, _i, _d, _d]),
        "orc_set_thermo": (None, [_p, _d, _i, _i, _i]),
        "orc_set_step_index": (None, [_p, ctypes.c_long]),
        "orc_kinetic_energy": (_d, [_p]),
        "orc_temperature": (_d, [_p]),
        "orc_pressure": (_d, [_p]),
        "orc_set_berendsen": (None, [_p, _d, _d, _d, _i, _d, _d]),
        "orc_set_nve": (None, [_p, _d, _d, _d, _d, _d, _i]),
        "orc_elc_acc": (None, [_p]),
        "orc_energy_elc": (_d, [_p]),
        "orc_point_elc": (None, [_p, _i, _dp, _dp, _dp]),
        "orc_short_range_acc": (None, [_p]),
        "orc_set_ewald_k": (None, [_p, _d, _d, _i, _i, _i, _i, _d, _i, _d]),
        "orc_setup": (_i, [_p]),
        "orc_step": (_i, [_p, _d]),
        "orc_forces": (None, [_p]),
        "orc_forces_sampled": (None, [_p, _i, _ip]),
        "orc_setup_sampled": (_i, [_p, _i, _ip]),
        "orc_wrap": (_i, [_p]),
        "orc_ghosts": (None, [_p]),
        "orc_build": (_i, [_p, _i]),
        "orc_reset_acc": (None, [_p]),
        "orc_lj_acc": (None, [_p]),
        "orc_ewald_r_acc": (None, [_p]),
        "orc_ewald_k_acc": (None, [_p]),
        "orc_ewald_k_sample": (None, [_p, _i, _i, _i, _i, _dp, _dp]),
        "orc_n": (_i, [_p]),
        "orc_get": (None, [_p, _i, _dp]),
        "orc_set": (None, [_p, _i, _dp]),
        "orc_num_ghosts": (_i, [_p]),
        "orc_get_ghosts": (None, [_p, _dp, _ip]),
        "orc_num_pairs": (ctypes.c_long, [_p]),
        "orc_get_pairs": (None, [_p, _ip]),
        "orc_get_nbins": (None, [_p, _ip]),
        "orc_virial": (_d, [_p]),
        "orc_reset_virial": (None, [_p]),
        "orc_num_kvectors": (_i, [_p]),
        "orc_get_kvectors": (None, [_p, _dp, _dp]),
        "orc_get_sk": (None, [_p, _dp]),
        "orc_energy_r": (_d, [_p]),
        "orc_energy_k": (_d, [_p]),
        "orc_self_energy": (_d, [_p]),
        "orc_energy_lj": (_d, [_p]),
        "orc_point_r": (None, [_p, _i, _dp, _dp, _dp]),
        "orc_point_k": (None, [_p, _i, _dp, _dp, _dp]),
        "orc_fm_adjoint": (None, [_p, _dp, _dp, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else None


class Oracle:
    """One CAVIAR single-domain system on the CPU (FP64 restatement).

    ``cfg`` is a :class:`caviar_amd.System` description (box, types, atoms and
    force-field settings); the same object configures the GPU path, so tests
    compare the two on identical inputs.
    """

    def __init__(self, cfg, threads: int = 1):
        self.lib = _load()
        self.lib.orc_set_threads(int(threads))
        self.h = self.lib.orc_create()
        self.cfg = cfg
        self._keep = []
        L = self.lib
        lo = np.ascontiguousarray(cfg.lower, dtype=np.float64)
        hi = np.ascontiguousarray(cfg.upper, dtype=np.float64)
        bc = np.ascontiguousarray(cfg.bc, dtype=np.int32)
        L.orc_set_box(self.h, _dptr(lo), _dptr(hi), _iptr(bc))
        L.orc_set_ghost_cutoff(self.h, float(cfg.ghost_cutoff))
        mass = np.ascontiguousarray(cfg.mass, dtype=np.float64)
        charge = np.ascontiguousarray(cfg.charge, dtype=np.float64)
        L.orc_set_types(self.h, len(mass), _dptr(mass), _dptr(charge))
        n = cfg.n
        ids = np.ascontiguousarray(cfg.ids, dtype=np.int32)
        typ = np.ascontiguousarray(cfg.types, dtype=np.int32)
        mol = None if cfg.mol is None else np.ascontiguousarray(cfg.mol, dtype=np.int32)
        pos = np.ascontiguousarray(cfg.pos, dtype=np.float64)
        vel = np.ascontiguousarray(cfg.vel, dtype=np.float64)
        L.orc_set_atoms(self.h, n, _iptr(ids), _iptr(typ), _iptr(mol), _dptr(pos), _dptr(vel))
        nl = cfg.neighbor
        if nl is not None:
            L.orc_set_neighbor(self.h, nl.cutoff, nl.cutoff_extra_coef, int(nl.from_cells),
                               int(nl.build_cells), int(nl.all_atom_test), int(nl.rebuild_test))
        nt = len(mass)
        if cfg.lj is not None:
            eps = np.ascontiguousarray(cfg.lj.epsilon, dtype=np.float64).reshape(nt * nt)
            sig = np.ascontiguousarray(cfg.lj.sigma, dtype=np.float64).reshape(nt * nt)
            L.orc_set_lj(self.h, cfg.lj.cutoff, _dptr(eps), _dptr(sig))
        if cfg.lj_list is not None:
            from caviar_amd.system import lj_list_args
            a = lj_list_args(cfg.lj_list, nt)
            self._keep.append(a)
            if L.orc_set_lj_list(self.h, a[0], *[_dptr(x) for x in a[1:6]], a[6], _dptr(a[7]),
                                 _dptr(a[8]), a[9], a[10], a[11]):
                raise ValueError("force_field lj: give epsilon/sigma by array or by atom, not both")
        if cfg.short_range is not None:
            from caviar_amd.system import short_range_args
            a = short_range_args(cfg.short_range, nt)
            self._keep.append(a)
            L.orc_set_short_range(self.h, a[0], a[1], a[2], _dptr(a[3]), _dptr(a[4]))
        if cfg.ewald_r is not None:
            lam = cfg.ewald_r.lambda_
            lam = None if lam is None else np.ascontiguousarray(lam, dtype=np.float64).reshape(nt * nt)
            L.orc_set_ewald_r(self.h, cfg.ewald_r.alpha, cfg.ewald_r.k_electrostatic, _dptr(lam))
        if cfg.thermo is not None:
            t = cfg.thermo
            L.orc_set_thermo(self.h, float(t.k_b), int(t.n_r_df), int(t.temperature_process),
                             int(t.pressure_process))
        if cfg.berendsen is not None:
            b = cfg.berendsen
            L.orc_set_berendsen(self.h, b.temperature, b.coupling, b.dt, int(b.step), b.lambda_min,
                                b.lambda_max)
        if cfg.nve is not None:
            v = cfg.nve
            L.orc_set_nve(self.h, v.energy_per_dof, v.energy_tot, v.kbt, v.kb, v.temperature, int(v.step))
        if cfg.elc is not None:
            c = cfg.elc
            L.orc_set_elc(self.h, int(c.slab_normal_axis), int(c.kx_max), int(c.ky_max), float(c.lz),
                          float(c.k_electrostatic))
        if cfg.ewald_k is not None:
            e = cfg.ewald_k
            L.orc_set_ewald_k(self.h, e.alpha, e.k_electrostatic, e.kx_max, e.ky_max, e.kz_max,
                              int(e.epsilon_dipole is not None),
                              float(e.epsilon_dipole or 0.0),
                              -1 if e.slab_normal_axis is None else int(e.slab_normal_axis),
                              float(e.slab_lz))

    def __del__(self):
        try:
            self.lib.orc_destroy(self.h)
        except Exception:
            pass

    # driver
    def setup(self):
        return self.lib.orc_setup(self.h)

    def setup_sampled(self, idx):
        """setup() with the Ewald_k force pass over the atoms ``idx`` only
        (after the full S(k)); other atoms lack their k-space term."""
        self._idx = np.ascontiguousarray(idx, dtype=np.int32)
        return self.lib.orc_setup_sampled(self.h, len(self._idx), _iptr(self._idx))

    def step(self, dt: float):
        return self.lib.orc_step(self.h, float(dt))

    def run(self, dt: float, nsteps: int):
        for _ in range(nsteps):
            rc = self.step(dt)
            if rc:
                return rc
        return 0

    def forces(self):
        self.lib.orc_forces(self.h)

    def wrap(self):
        return self.lib.orc_wrap(self.h)

    def ghosts(self):
        self.lib.orc_ghosts(self.h)

    def build(self, update: bool):
        return self.lib.orc_build(self.h, int(update))

    # state
    def get(self, what: str) -> np.ndarray:
        idx = {"pos": 0, "vel": 1, "acc": 2, "acc_old": 3}[what]
        out = np.empty((self.lib.orc_n(self.h), 3), dtype=np.float64)
        self.lib.orc_get(self.h, idx, _dptr(out))
        return out

    def set(self, what: str, arr) -> None:
        idx = {"pos": 0, "vel": 1, "acc": 2, "acc_old": 3}[what]
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self.lib.orc_set(self.h, idx, _dptr(a))

    def ghost_data(self):
        g = self.lib.orc_num_ghosts(self.h)
        gp = np.empty((g, 3), dtype=np.float64)
        gs = np.empty(g, dtype=np.int32)
        self.lib.orc_get_ghosts(self.h, _dptr(gp), _iptr(gs))
        return gp, gs

    def pairs(self) -> np.ndarray:
        npairs = self.lib.orc_num_pairs(self.h)
        out = np.empty((npairs, 2), dtype=np.int32)
        if npairs:
            self.lib.orc_get_pairs(self.h, _iptr(out))
        return out

    def nbins(self):
        nb = np.zeros(3, dtype=np.int32)
        self.lib.orc_get_nbins(self.h, _iptr(nb))
        return tuple(int(x) for x in nb)

    def virial(self) -> float:
        return self.lib.orc_virial(self.h)

    def reset_virial(self):
        self.lib.orc_reset_virial(self.h)

    def kvectors(self):
        nk = self.lib.orc_num_kvectors(self.h)
        kv = np.empty((nk, 3))
        fk = np.empty(nk)
        self.lib.orc_get_kvectors(self.h, _dptr(kv), _dptr(fk))
        return kv, fk

    def sk(self) -> np.ndarray:
        nk = self.lib.orc_num_kvectors(self.h)
        out = np.empty((nk, 2))
        self.lib.orc_get_sk(self.h, _dptr(out))
        return out[:, 0] + 1j * out[:, 1]

    def energy_r(self) -> float:
        return self.lib.orc_energy_r(self.h)

    def energy_k(self) -> float:
        return self.lib.orc_energy_k(self.h)

    def self_energy(self) -> float:
        return self.lib.orc_self_energy(self.h)

    def energy_lj(self) -> float:
        return self.lib.orc_energy_lj(self.h)

    def point_r(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        f = np.empty_like(pts)
        p = np.empty(len(pts))
        self.lib.orc_point_r(self.h, len(pts), _dptr(pts), _dptr(f), _dptr(p))
        return f, p

    def point_k(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        f = np.empty_like(pts)
        p = np.empty(len(pts))
        self.lib.orc_point_k(self.h, len(pts), _dptr(pts), _dptr(f), _dptr(p))
        return f, p

    def thermo(self):
        """(kinetic energy now, temperature and pressure as last finalized)."""
        return (self.lib.orc_kinetic_energy(self.h), self.lib.orc_temperature(self.h),
                self.lib.orc_pressure(self.h))

    def energy_elc(self) -> float:
        return self.lib.orc_energy_elc(self.h)

    def point_elc(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        f = np.empty_like(pts)
        p = np.empty(len(pts))
        self.lib.orc_point_elc(self.h, len(pts), _dptr(pts), _dptr(f), _dptr(p))
        return f, p

    def adjoint(self, weights):
        nt = len(self.cfg.mass)
        w = np.ascontiguousarray(weights, dtype=np.float64).reshape(-1, 3)
        ge = np.zeros(nt * nt)
        gs = np.zeros(nt * nt)
        gq = np.zeros(nt)
        self.lib.orc_fm_adjoint(self.h, _dptr(w), _dptr(ge), _dptr(gs), _dptr(gq))
        return ge.reshape(nt, nt), gs.reshape(nt, nt), gq

    def ewald_k_sample(self, k0, k1, i0, i1):
        ts = ctypes.c_double()
        tf = ctypes.c_double()
        self.lib.orc_ewald_k_sample(self.h, k0, k1, i0, i1, ctypes.byref(ts), ctypes.byref(tf))
        return ts.value, tf.value
