"""Host-side description of one CAVIAR single-domain system.

These dataclasses mirror the settings the reference's CASL objects hold for
the short-range + Ewald path, with the same names and defaults:

* ``domain box`` (domain.cpp:206-266): lower/upper bounds, boundary_condition;
* ``atom_data basic`` (atom_data.cpp:78-306): ghost_cutoff, per-type mass and
  charge, owned atoms (id, type, molecule index, position, velocity);
* ``neighborlist basic`` (neighborlist.cpp:193-251, neighborlist.h:134);
* ``force_field lj_cell_list`` (lj_cell_list.cpp:33-76);
* ``force_field lj`` (lj.cpp:42-147) and ``force_field
  electrostatic_short_range`` (electrostatic_short_range.cpp:39-99);
* ``force_field electrostatic_ewald_r`` (electrostatic_ewald_r.cpp:31-84);
* ``force_field electrostatic_ewald_k`` (electrostatic_ewald_k.cpp:31-139);
* ``force_field electrostatic_ewald_slab_correction``
  (electrostatic_ewald_slab_correction.cpp:44-126).

A :class:`System` is consumed by :class:`caviar_amd.Engine` (GPU path).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class NeighborSettings:
    """``neighborlist basic`` (neighborlist.cpp:193-251)."""
    cutoff: float
    cutoff_extra_coef: float = 1.12347          # neighborlist.h:134
    from_cells: bool = True                      # make_verlet_list_from_cell_list
    build_cells: bool = True                     # build_cell_list
    all_atom_test: bool = False
    rebuild_test: bool = False

    @property
    def cutoff_extra(self) -> float:
        return self.cutoff * self.cutoff_extra_coef


@dataclass
class LJSettings:
    """``force_field lj_cell_list`` (lj_cell_list.cpp:33-76).

    ``epsilon[i][j]`` / ``sigma[i][j]`` are the full type matrices (the
    reference does not symmetrise them, preprocessors.h:568-578)."""
    cutoff: float
    epsilon: np.ndarray
    sigma: np.ndarray


@dataclass
class LJListSettings:
    """``force_field lj``, the Verlet-list LJ (lj.cpp:42-147): type matrices
    ``epsilon``/``sigma`` (nt x nt, unset entries 0) or per-type
    ``epsilon_atom``/``sigma_atom`` (Lorentz-Berthelot off the diagonal,
    lj.cpp:207-222); per-pair ``cutoff_list`` or ``wca``; ``lambda_s`` /
    ``lambda_e`` scale sigma / epsilon (unset entries 1);
    ``ignore_intra_molecule``; ``jump_tol`` (None: off)."""
    cutoff: float
    epsilon: Optional[np.ndarray] = None
    sigma: Optional[np.ndarray] = None
    epsilon_atom: Optional[np.ndarray] = None
    sigma_atom: Optional[np.ndarray] = None
    cutoff_list: Optional[np.ndarray] = None
    wca: bool = False
    lambda_s: Optional[np.ndarray] = None
    lambda_e: Optional[np.ndarray] = None
    ignore_intra_molecule: bool = False
    jump_tol: Optional[float] = None


@dataclass
class ShortRangeSettings:
    """``force_field electrostatic_short_range`` (electrostatic_short_range.cpp:
    39-99): shifted Coulomb within ``cutoff`` over the Verlet list, plus a
    uniform ``external_field``; ``lambda_`` per type pair for same-molecule
    pairs."""
    cutoff: float
    k_electrostatic: float = 1.0
    beta: float = 1.0
    external_field: tuple = (0.0, 0.0, 0.0)
    lambda_: Optional[np.ndarray] = None


@dataclass
class ELCSettings:
    """``force_field electrostatic_ewald_slab_correction``, the ELC term of
    Ewald sums in slab geometry (electrostatic_ewald_slab_correction.cpp:
    44-126): the slab normal axis (0, 1, 2), the in-plane k counts, the slab
    thickness ``lz`` (at least the box height along the normal)."""
    slab_normal_axis: int = 2
    kx_max: int = 1
    ky_max: int = 1
    lz: float = 1.0
    k_electrostatic: float = 1.0


@dataclass
class ThermoSettings:
    """``atom_data basic`` k_b, n_r_df (-1: from the boundary conditions,
    atom_data.cpp:322-350), temperature_process (on by default,
    atom_data.h:735) and pressure_process (atom_data.cpp:121-151)."""
    k_b: float = -1.0
    n_r_df: int = -1
    temperature_process: bool = True
    pressure_process: bool = False


@dataclass
class BerendsenSettings:
    """``constraint berendsen`` (constraint/berendsen.cpp:33-146)."""
    temperature: float
    coupling: float
    dt: float
    step: int = 1
    lambda_min: float = 0.8
    lambda_max: float = 1.25


@dataclass
class NVESettings:
    """``constraint nve`` (constraint/nve.cpp:33-178); one of energy_per_dof,
    energy_tot, kbt or kb + temperature (unset values <= 0)."""
    energy_per_dof: float = -1.0
    energy_tot: float = -1.0
    kbt: float = -1.0
    kb: float = -1.0
    temperature: float = -1.0
    step: int = 1


@dataclass
class EwaldRSettings:
    """``force_field electrostatic_ewald_r`` (electrostatic_ewald_r.cpp:31-84)."""
    alpha: float = 1.0
    k_electrostatic: float = 1.0
    lambda_: Optional[np.ndarray] = None         # per type pair, same-molecule pairs only


@dataclass
class EwaldKSettings:
    """``force_field electrostatic_ewald_k`` (electrostatic_ewald_k.cpp:31-139)."""
    alpha: float
    k_electrostatic: float = 1.0
    kx_max: int = 1
    ky_max: int = 1
    kz_max: int = 1
    epsilon_dipole: Optional[float] = None       # None: no dipole term
    slab_normal_axis: Optional[int] = None       # None: no slab geometry
    slab_lz: float = 1.0

    @staticmethod
    def cube(alpha: float, k_max: int, k_electrostatic: float = 1.0) -> "EwaldKSettings":
        return EwaldKSettings(alpha=alpha, k_electrostatic=k_electrostatic,
                              kx_max=k_max, ky_max=k_max, kz_max=k_max)


@dataclass
class System:
    lower: np.ndarray
    upper: np.ndarray
    bc: np.ndarray
    ghost_cutoff: float
    mass: np.ndarray
    charge: np.ndarray
    pos: np.ndarray
    vel: np.ndarray
    types: np.ndarray
    ids: Optional[np.ndarray] = None
    mol: Optional[np.ndarray] = None
    neighbor: Optional[NeighborSettings] = None
    lj: Optional[LJSettings] = None
    lj_list: Optional[LJListSettings] = None
    short_range: Optional[ShortRangeSettings] = None
    ewald_r: Optional[EwaldRSettings] = None
    ewald_k: Optional[EwaldKSettings] = None
    elc: Optional[ELCSettings] = None
    thermo: Optional[ThermoSettings] = None
    berendsen: Optional[BerendsenSettings] = None
    nve: Optional[NVESettings] = None
    # atom_data add_random_velocity SEED AMP (atom_data.cpp:280-288), applied
    # by the CASL layer; ``vel`` then holds the velocities it produced
    random_velocity: Optional[tuple] = None
    meta: dict = field(default_factory=dict)

    def __post_init__(self):
        self.lower = np.asarray(self.lower, dtype=np.float64).reshape(3)
        self.upper = np.asarray(self.upper, dtype=np.float64).reshape(3)
        self.bc = np.asarray(self.bc, dtype=np.int32).reshape(3)
        self.mass = np.asarray(self.mass, dtype=np.float64).reshape(-1)
        self.charge = np.asarray(self.charge, dtype=np.float64).reshape(-1)
        self.pos = np.ascontiguousarray(self.pos, dtype=np.float64).reshape(-1, 3)
        self.vel = np.ascontiguousarray(self.vel, dtype=np.float64).reshape(-1, 3)
        self.types = np.ascontiguousarray(self.types, dtype=np.int32).reshape(-1)
        if self.ids is None:
            self.ids = np.arange(len(self.pos), dtype=np.int32)
        self.ids = np.ascontiguousarray(self.ids, dtype=np.int32)
        if self.mol is not None:
            self.mol = np.ascontiguousarray(self.mol, dtype=np.int32)
        if len(self.charge) != len(self.mass):
            raise ValueError("mass and charge must cover the same types")
        if self.types.size and (self.types.min() < 0 or self.types.max() >= len(self.mass)):
            raise ValueError("atom type outside the declared types")

    @property
    def n(self) -> int:
        return len(self.pos)

    @property
    def ntypes(self) -> int:
        return len(self.mass)

    @property
    def size(self) -> np.ndarray:
        return self.upper - self.lower

    def copy(self) -> "System":
        import copy
        return copy.deepcopy(self)


def lj_list_args(lj: "LJListSettings", nt: int) -> tuple:
    """The ``force_field lj`` settings as the C-ABI argument tuple (after the
    handle): cutoff, eps, sig, eps_atom, sig_atom, cutoff_list, wca, lambda_s,
    lambda_e, ignore_intra_molecule, jump_fix, jump_tol; arrays as contiguous
    FP64 or None."""
    def mat(a):
        return None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(nt * nt)

    def vec(a):
        return None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(nt)
    return (float(lj.cutoff), mat(lj.epsilon), mat(lj.sigma), vec(lj.epsilon_atom),
            vec(lj.sigma_atom), mat(lj.cutoff_list), int(bool(lj.wca)), mat(lj.lambda_s),
            mat(lj.lambda_e), int(bool(lj.ignore_intra_molecule)), int(lj.jump_tol is not None),
            float(lj.jump_tol or 0.0))


def short_range_args(sr: "ShortRangeSettings", nt: int) -> tuple:
    """The ``force_field electrostatic_short_range`` settings as the C-ABI
    argument tuple (after the handle)."""
    lam = None if sr.lambda_ is None else np.ascontiguousarray(sr.lambda_, dtype=np.float64).reshape(nt * nt)
    ext = np.ascontiguousarray(sr.external_field, dtype=np.float64).reshape(3)
    return float(sr.cutoff), float(sr.k_electrostatic), float(sr.beta), ext, lam
