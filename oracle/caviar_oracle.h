/* CPU oracle for the CAVIAR short-range + Ewald force path (TEST INFRASTRUCTURE).
 *
 * A plain-C, FP64 restatement of the reference algorithm, used only as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * The product path (caviar_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference tree).  Pins (what checks each function against an
 * independent answer) are listed in tests/test_oracle.py.
 *
 * Layout follows the reference: AoS xyz doubles, owned atoms 0..n-1, ghost
 * atoms addressed as n+g in neighbour lists (neighborlist.cpp:287,470).
 */
#ifndef CAVIAR_ORACLE_H
#define CAVIAR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sys orc_sys;

orc_sys *orc_create(void);
void orc_destroy(orc_sys *s);
void orc_set_threads(int nthreads);

/* Domain box + boundary flags (domain.cpp:206-266,470-477). */
void orc_set_box(orc_sys *s, const double lo[3], const double hi[3], const int bc[3]);
/* Atom_data::ghost_cutoff (atom_data.cpp:87-91). */
void orc_set_ghost_cutoff(orc_sys *s, double gc);
/* add_type_mass / add_type_charge (atom_data.cpp:206-245); mass 0 -> mass_inv 0. */
void orc_set_types(orc_sys *s, int nt, const double *mass, const double *charge);
/* Owned atoms (atom_data.cpp:487-521).  mol may be NULL (-> -1 like add_atom). */
void orc_set_atoms(orc_sys *s, int n, const int *id, const int *type, const int *mol,
                   const double *pos, const double *vel);
/* neighborlist basic: cutoff, cutoff_extra_coef, make_verlet_list_from_cell_list
 * (from_cells), build_cell_list, all_atom_test, rebuild_test (neighborlist.cpp:193-251). */
void orc_set_neighbor(orc_sys *s, double cutoff, double coef, int from_cells, int build_cells,
                      int all_atom_test, int rebuild_test);
/* force_field lj_cell_list (lj_cell_list.cpp:33-76); eps/sigma are nt*nt row-major [ti][tj]. */
void orc_set_lj(orc_sys *s, double rc, const double *eps, const double *sigma);
/* force_field lj (lj.cpp:42-147 read, 162-339 verify_settings): eps/sig nt*nt
 * (by array) or eps_atom/sig_atom nt (by atom, Lorentz-Berthelot), cutoff_list
 * nt*nt or NULL, wca, lambda_s/lambda_e nt*nt or NULL, ignore_intra_molecule,
 * jump_fix + jump_tol.  Returns 1 if both or neither input forms are given. */
int orc_set_lj_list(orc_sys *s, double cutoff, const double *eps, const double *sig,
                    const double *eps_atom, const double *sig_atom, const double *cutoff_list,
                    int wca, const double *lambda_s, const double *lambda_e,
                    int ignore_intra_molecule, int jump_fix, double jump_tol);
/* force_field electrostatic_short_range (electrostatic_short_range.cpp:39-99):
 * cutoff, k_electrostatic, beta, external_field (3 or NULL), lambda nt*nt or NULL. */
void orc_set_short_range(orc_sys *s, double cutoff, double ke, double beta, const double ext[3],
                         const double *lambda);
/* force_field electrostatic_ewald_slab_correction (electrostatic_ewald_slab_correction.cpp:
 * 44-126): slab_normal_axis 0/1/2, kx_max, ky_max, lz, k_electrostatic. */
void orc_set_elc(orc_sys *s, int axis, int kx, int ky, double lz, double ke);
/* force_field electrostatic_ewald_r (electrostatic_ewald_r.cpp:31-110); lambda nt*nt or NULL. */
void orc_set_ewald_r(orc_sys *s, double alpha, double ke, const double *lambda);
/* force_field electrostatic_ewald_k (electrostatic_ewald_k.cpp:31-139).
 * dipole!=0 enables epsilon_dipole; slab_axis -1 = no slab geometry. */
void orc_set_ewald_k(orc_sys *s, double alpha, double ke, int kx, int ky, int kz, int dipole,
                     double eps_dipole, int slab_axis, double slab_lz);

/* Md_simulator::setup (md_simulator.cpp:504-549): neighbour init, wrap, ghosts,
 * build(true), reset acceleration, all force fields. */
int orc_setup(orc_sys *s);
/* One velocity-Verlet step (md_simulator.cpp:555-617).  Returns 0, or 1 when an
 * acceleration is NaN before the drift (md_simulator.cpp:584-590; state untouched). */
int orc_step(orc_sys *s, double dt);
/* reset_owned_acceleration + every enabled force field at the current list. */
void orc_forces(orc_sys *s);
/* Individual stages, exposed for the parity tests. */
int orc_wrap(orc_sys *s);            /* exchange_owned_single_md_domain */
void orc_ghosts(orc_sys *s);         /* exchange_ghost_single_md_domain */
int orc_build(orc_sys *s, int update); /* Neighborlist::build */
void orc_reset_acc(orc_sys *s);
void orc_lj_acc(orc_sys *s);
void orc_ewald_r_acc(orc_sys *s);
void orc_lj_list_acc(orc_sys *s);
void orc_short_range_acc(orc_sys *s);
void orc_elc_acc(orc_sys *s);
void orc_ewald_k_acc(orc_sys *s);
/* Bounded-sample timing of Ewald_k for the CPU baseline: S(k) over all atoms
 * for k in [k0,k1), and the force pass for atoms [i0,i1) over all k. */
void orc_ewald_k_sample(orc_sys *s, int k0, int k1, int i0, int i1, double *t_s, double *t_f);

/* orc_forces / orc_setup with the Ewald_k force pass restricted to the m
 * atoms idx[0..m) (after the full structure factor), for comparisons at sizes
 * where the whole O(N K) force pass would take too long on the CPU; the other
 * atoms' accelerations lack their k-space term.  Same arithmetic per atom. */
void orc_forces_sampled(orc_sys *s, int m, const int *idx);
int orc_setup_sampled(orc_sys *s, int m, const int *idx);

/* atom_data k_b, n_r_df (-1: from the boundary conditions), temperature_process,
 * pressure_process (atom_data.cpp:121-151; kinetic_energy.cpp:162-292). */
void orc_set_thermo(orc_sys *s, double k_b, int n_r_df, int temperature_process, int pressure_process);
/* Md_simulator::current_step of the next orc_step (thermostats act when it is
 * a multiple of their step). */
void orc_set_step_index(orc_sys *s, long step);
double orc_kinetic_energy(const orc_sys *s);   /* 0 for n_r_df 4..6 unless I is singular */
double orc_temperature(const orc_sys *s);      /* last finalized (0 until then) */
double orc_pressure(const orc_sys *s);
/* constraint berendsen (berendsen.cpp:33-146) and nve (nve.cpp:33-178). */
void orc_set_berendsen(orc_sys *s, double t0, double coupling, double dt, int step, double lmin,
                       double lmax);
void orc_set_nve(orc_sys *s, double e_dof, double e_tot, double kbt, double kb, double t, int step);

/* Getters. what: 0 pos, 1 vel, 2 acc, 3 acc_old (3n doubles). */
int orc_n(const orc_sys *s);
void orc_get(const orc_sys *s, int what, double *out);
void orc_set(orc_sys *s, int what, const double *in);
int orc_num_ghosts(const orc_sys *s);
void orc_get_ghosts(const orc_sys *s, double *gpos, int *gsrc);
long orc_num_pairs(const orc_sys *s);
void orc_get_pairs(const orc_sys *s, int *ij);          /* reference order, 2*np ints */
void orc_get_nbins(const orc_sys *s, int nb[3]);
double orc_virial(const orc_sys *s);
void orc_reset_virial(orc_sys *s);
int orc_num_kvectors(const orc_sys *s);
void orc_get_kvectors(const orc_sys *s, double *kvec, double *fk);
void orc_get_sk(const orc_sys *s, double *sk);          /* 2*nk (re,im) */

/* Energies and point queries (off the MD path). */
double orc_energy_r(orc_sys *s);      /* Electrostatic_ewald_r::energy, P7 cell walk */
double orc_energy_k(orc_sys *s);      /* Electrostatic_ewald_k::energy (k + self [+ dipole]) */
double orc_self_energy(orc_sys *s);
double orc_energy_lj(orc_sys *s);     /* shifted-force LJ energy (no reference counterpart) */
double orc_energy_elc(orc_sys *s);    /* Electrostatic_ewald_slab_correction::energy */
void orc_point_elc(orc_sys *s, int np, const double *pts, double *field, double *pot);
void orc_point_r(orc_sys *s, int np, const double *pts, double *field, double *pot);
void orc_point_k(orc_sys *s, int np, const double *pts, double *field, double *pot);

/* Force-matching adjoint reference: g = sum_i w_i . dF_i/dtheta for the LJ
 * epsilon/sigma matrices (nt*nt) and the type charges (nt), F = mass * acc.
 * No reference counterpart. */
void orc_fm_adjoint(orc_sys *s, const double *w, double *g_eps, double *g_sig, double *g_q);

#ifdef __cplusplus
}
#endif
#endif
