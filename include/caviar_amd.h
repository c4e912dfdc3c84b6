/* caviar_amd: MI355X (gfx950) short-range + Ewald force path of CAVIAR,
 * exposed as a plain C ABI (no torch types, plain pointers and sizes).
 *
 * One context = one single-domain system resident on one GPU.  Host buffers
 * passed in and out use the reference's layout: AoS xyz doubles, owned atoms
 * 0..n-1, ghost atoms n+g.  Every function returns CAV_OK or an error code
 * and never exits; cav_last_error() gives the message (the reference prints
 * and calls exit(1) instead, interpreter/error.cpp:27-30,175-202).
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference tree, src/caviar/objects/ unless given in full).
 */
#ifndef CAVIAR_AMD_H
#define CAVIAR_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cav_ctx cav_ctx;

enum cav_status {
  CAV_OK = 0,
  CAV_ERR_ARG = 1,      /* invalid argument or setting */
  CAV_ERR_HIP = 2,      /* HIP runtime failure */
  CAV_ERR_STATE = 3,    /* call out of order (e.g. step before setup) */
  CAV_ERR_NAN = 4,      /* NaN acceleration before a drift (md_simulator.cpp:584-590) */
  CAV_ERR_CAPACITY = 5  /* a device buffer could not be grown */
};

enum cav_precision {
  CAV_FP64 = 0,  /* plain form: every term in FP64 */
  CAV_MIXED = 1  /* FP32 pair/erfc/LJ math and FP32 k-space products; FP64 S(k)
                    accumulation across atom blocks, energy/virial, integration */
};

enum cav_array { CAV_POS = 0, CAV_VEL = 1, CAV_ACC = 2, CAV_ACC_OLD = 3 };

/* Force fields of one context, as bits of the cav_forces mask. */
enum cav_force_field {
  CAV_FF_LJ = 1,           /* force_field lj_cell_list */
  CAV_FF_EWALD_R = 2,      /* force_field electrostatic_ewald_r */
  CAV_FF_EWALD_K = 4,      /* force_field electrostatic_ewald_k */
  CAV_FF_LJ_LIST = 8,      /* force_field lj (Verlet list) */
  CAV_FF_SHORT_RANGE = 16, /* force_field electrostatic_short_range */
  CAV_FF_ELC = 32          /* force_field electrostatic_ewald_slab_correction */
};

/* ---- lifetime ---------------------------------------------------------- */
/* Replaces the objects the interpreter creates (object_creator/objects_types.cpp:25-51). */
int cav_create(int device, cav_ctx **out);
int cav_destroy(cav_ctx *ctx);
const char *cav_last_error(const cav_ctx *ctx);
int cav_set_precision(cav_ctx *ctx, int precision);

/* ---- settings (the CASL object commands) ------------------------------- */
/* domain box: xmin..zmax, boundary_condition (domain.cpp:206-266). */
int cav_set_box(cav_ctx *ctx, const double lower[3], const double upper[3], const int bc[3]);
/* atom_data basic: ghost_cutoff (atom_data.cpp:87-91). */
int cav_set_ghost_cutoff(cav_ctx *ctx, double ghost_cutoff);
/* atom_data basic: add_type_mass / add_type_charge (atom_data.cpp:206-245). */
int cav_set_types(cav_ctx *ctx, int ntypes, const double *mass, const double *charge);
/* neighborlist basic: cutoff, cutoff_extra_coef, make_verlet_list_from_cell_list,
 * build_cell_list, all_atom_test, rebuild_test (neighborlist.cpp:193-251). */
int cav_set_neighbor(cav_ctx *ctx, double cutoff, double cutoff_extra_coef, int from_cells,
                     int build_cells, int all_atom_test, int rebuild_test);
/* force_field lj_cell_list: cutoff, epsilon i j, sigma i j (lj_cell_list.cpp:33-76);
 * eps/sigma are ntypes*ntypes row-major [ti][tj]. */
int cav_set_lj(cav_ctx *ctx, double cutoff, const double *epsilon, const double *sigma);
/* force_field lj (Verlet-list LJ): cutoff, epsilon i j / sigma i j (by array:
 * eps and sig, ntypes*ntypes, unset entries 0) or epsilon_atom i / sigma_atom
 * i (by atom: eps_atom and sig_atom, ntypes, Lorentz-Berthelot off the
 * diagonal), cutoff_list i j (ntypes*ntypes, unset 0) or wca (cutoff
 * 2^(1/6) sigma_ij), lambda_s / lambda_e i j (ntypes*ntypes, unset 1; scale
 * sigma / epsilon), ignore_intra_molecule, jump_tol (jump_fix != 0).  NULL
 * for what is not given.  Resolved as Lj::read / verify_settings resolve them
 * (lj.cpp:42-147,162-339); forces as Lj::calculate_acceleration
 * (lj.cpp:341-483). */
int cav_set_lj_list(cav_ctx *ctx, double cutoff, const double *eps, const double *sig,
                    const double *eps_atom, const double *sig_atom, const double *cutoff_list,
                    int wca, const double *lambda_s, const double *lambda_e,
                    int ignore_intra_molecule, int jump_fix, double jump_tol);
/* force_field electrostatic_short_range: cutoff, k_electrostatic, beta,
 * external_field (3 doubles or NULL), lambda i j (ntypes*ntypes or NULL)
 * (electrostatic_short_range.cpp:39-99); forces as calculate_acceleration
 * (electrostatic_short_range.cpp:133-258). */
int cav_set_short_range(cav_ctx *ctx, double cutoff, double k_electrostatic, double beta,
                        const double external_field[3], const double *lambda);
/* force_field electrostatic_ewald_slab_correction (ELC): slab_normal_axis
 * (0 x, 1 y, 2 z), kx_max, ky_max, lz (slab thickness, at least the box
 * height along the normal), k_electrostatic
 * (electrostatic_ewald_slab_correction.cpp:44-126); forces as its
 * calculate_acceleration (electrostatic_ewald_slab_correction.cpp:134-181,
 * electrostatic_ewald_slab_correction/initialize.cpp, field.cpp). */
int cav_set_elc(cav_ctx *ctx, int slab_normal_axis, int kx_max, int ky_max, double lz,
                double k_electrostatic);
/* force_field electrostatic_ewald_r: alpha, k_electrostatic, lambda i j
 * (electrostatic_ewald_r.cpp:31-110); lambda may be NULL. */
int cav_set_ewald_r(cav_ctx *ctx, double alpha, double k_electrostatic, const double *lambda);
/* force_field electrostatic_ewald_k: alpha, k_electrostatic, kx/ky/kz_max,
 * epsilon_dipole (dipole != 0), slab_normal_axis (-1 = none), slab_lz
 * (electrostatic_ewald_k.cpp:31-139). */
int cav_set_ewald_k(cav_ctx *ctx, double alpha, double k_electrostatic, int kx_max, int ky_max,
                    int kz_max, int dipole, double epsilon_dipole, int slab_normal_axis,
                    double slab_lz);

/* atom_data basic: k_b, n_r_df (-1: from the boundary conditions,
 * atom_data.cpp:322-350), temperature_process, pressure_process
 * (atom_data.cpp:121-151).  With temperature_process the temperature is
 * finalized after setup and at the end of every step
 * (kinetic_energy.cpp:162-202), with pressure_process the pressure at the end
 * of every step (kinetic_energy.cpp:210-292).  The kinetic energy removes the
 * centre-of-mass velocity for n_r_df 1..3; for 4..6 it is 0 unless the
 * inertia tensor about the mass centre is exactly singular, as the reference's
 * test of matrix_inverse_3d's return code makes it (kinetic_energy.cpp:92-105). */
int cav_set_thermo(cav_ctx *ctx, double k_b, int n_r_df, int temperature_process,
                   int pressure_process);
/* constraint berendsen: temperature, coupling, dt, step, lambda_min, lambda_max
 * (constraint/berendsen.cpp:33-146; defaults 1, 0.8, 1.25), applied at the end
 * of each step whose index is a multiple of step. */
int cav_set_berendsen(cav_ctx *ctx, double temperature, double coupling, double dt, int step,
                      double lambda_min, double lambda_max);
/* constraint nve: energy_per_dof, energy_tot, kbt, kb, temperature (unset:
 * <= 0), step (constraint/nve.cpp:33-178). */
int cav_set_nve(cav_ctx *ctx, double energy_per_dof, double energy_tot, double kbt, double kb,
                double temperature, int step);
/* Md_simulator::current_step of the next step (thermostats act at multiples
 * of their step); cav_step advances it by one per step. */
int cav_set_step_index(cav_ctx *ctx, long step);
/* Atom_data::kinetic_energy now (kinetic_energy.cpp:140-154), and temperature()
 * and pressure() as last finalized (0 before; NULL outputs skipped). */
int cav_thermo(cav_ctx *ctx, double *kinetic_energy, double *temperature, double *pressure);

/* ---- atoms --------------------------------------------------------------- */
/* Atom_data::add_atom for n atoms at once (atom_data.cpp:487-521); mol may be
 * NULL (molecule index -1, as add_atom sets).  Host AoS input. */
int cav_upload_atoms(cav_ctx *ctx, int n, const int *id, const int *type, const int *mol,
                     const double *pos, const double *vel);
/* Read / overwrite one owned array (atom_struct_owned.position, velocity,
 * acceleration, acceleration_old), host AoS 3n doubles. */
int cav_download(cav_ctx *ctx, int which, double *out);
int cav_upload(cav_ctx *ctx, int which, const double *in);
/* Number of owned atoms. */
int cav_num_atoms(const cav_ctx *ctx, int *n);

/* ---- the MD path ----------------------------------------------------------- */
/* Md_simulator::setup (md_simulator.cpp:504-549): neighbour init, wrap, ghosts,
 * forced list build, reset acceleration, all force fields. */
int cav_setup(cav_ctx *ctx);
/* nsteps velocity-Verlet steps, Md_simulator::step -> integrate_velocity_verlet
 * (md_simulator.cpp:382-617).  On CAV_ERR_NAN the state is that before the
 * offending step. */
int cav_step(cav_ctx *ctx, double dt, int nsteps);
/* re_calculate_acc without the drift (md_simulator.cpp:361-374): wrap, ghosts,
 * Neighborlist::build(update), reset acceleration, every force field. */
int cav_recalculate(cav_ctx *ctx, int force_rebuild);
/* Neighborlist::build(update) alone (neighborlist.cpp:66-88) on the current
 * ghosts; *rebuilt (may be NULL) receives whether the Verlet list was rebuilt. */
int cav_neighbor_build(cav_ctx *ctx, int update, int *rebuilt);

/* ---- the pieces of a step, for callers that drive the objects one by one -- */
/* Atom_data::exchange_owned + exchange_ghost for one domain
 * (exchange_owned_single_md_domain.cpp:31-59, exchange_ghost_single_md_domain.cpp:
 * 26-218): wrap owned positions, rebuild the ghosts; *wrapped (may be NULL)
 * receives whether any atom was wrapped (which forces the next list rebuild). */
int cav_exchange(cav_ctx *ctx, int *wrapped);
/* Atom_data::reset_owned_acceleration (atom_data.cpp:308-316). */
int cav_reset_acceleration(cav_ctx *ctx);
/* Atom_data::reset_virial (atom_data.h:250; called per step, md_simulator.cpp:408). */
int cav_reset_virial(cav_ctx *ctx);
/* Force_field::calculate_acceleration of the fields in mask (lj_cell_list.cpp:
 * 85-175, electrostatic_ewald_r.cpp:115-317, electrostatic_ewald_k.cpp:161-294):
 * ADDS their accelerations and virial on the current ghosts and Verlet list,
 * as each reference force field adds into atom_struct_owned.acceleration. */
int cav_forces(cav_ctx *ctx, int mask);

/* ---- inspection (parity) ------------------------------------------------- */
/* atom_struct_ghost after exchange_ghost (exchange_ghost_single_md_domain.cpp):
 * positions (3g doubles) and the owned index each ghost images (g ints). */
int cav_num_ghosts(cav_ctx *ctx, int *g);
int cav_get_ghosts(cav_ctx *ctx, double *gpos, int *gsrc);
/* Neighborlist::neighlist as (i,j) pairs with the reference's membership
 * (owned j>i, every ghost j as n+g; neighborlist.cpp:257-307,515-573).
 * Order within an atom is not the reference's; compare sorted. */
int cav_num_pairs(cav_ctx *ctx, long *npairs);
int cav_get_pairs(cav_ctx *ctx, int *ij, long capacity, long *npairs);
/* Atom_data::virialForce accumulated by the last force evaluation (atom_data.h:250). */
int cav_virial(cav_ctx *ctx, double *virial);
/* Electrostatic_ewald_k::potential_k_coef_cmplx, i.e. S(k) in the reference's
 * k order (electrostatic_ewald_k/initialize.cpp:117-146); 2*nk doubles. */
int cav_num_kvectors(cav_ctx *ctx, int *nk);
int cav_get_sk(cav_ctx *ctx, double *sk);

/* ---- energies and point queries (off the MD path) ------------------------ */
/* Electrostatic_ewald_r::energy (electrostatic_ewald_r/energy.cpp:30-51),
 * Electrostatic_ewald_k::energy (electrostatic_ewald_k/energy.cpp:109-116) and
 * the shifted-force LJ energy matching Lj_cell_list's force (no reference
 * counterpart; the reference's Lj_cell_list has no energy). NULL outputs skipped. */
int cav_energy(cav_ctx *ctx, double *e_lj, double *e_r, double *e_k);
/* Electrostatic_ewald_slab_correction::energy (electrostatic_ewald_slab_correction/
 * energy.cpp:32-146) from the last force evaluation. */
int cav_energy_elc(cav_ctx *ctx, double *e);
/* Batched Force_field::field(const Vector&) / potential(const Vector&):
 * which 0 = Electrostatic_ewald_r (electrostatic_ewald_r/field.cpp:31-96,
 * potential.cpp:30-86), 1 = Electrostatic_ewald_k (electrostatic_ewald_k/
 * field.cpp:80-132, potential.cpp:31-129), 2 = Electrostatic_ewald_slab_
 * correction (electrostatic_ewald_slab_correction/field.cpp:29-136,
 * potential.cpp:30-79; state of the last force evaluation).  np points (AoS
 * 3np doubles). */
int cav_point_field(cav_ctx *ctx, int which, int np, const double *pts, double *field,
                    double *potential);

/* ---- differentiation (no counterpart in the reference) ------------------ */
/* Force-matching adjoint: given residual weights w (3n doubles, w_i = dL/dF_i
 * with F_i = mass_i * acceleration_i at the current positions and parameters),
 * returns g = sum_i w_i . dF_i/dtheta for theta = lj_cell_list epsilon[a][b]
 * and sigma[a][b] (nt*nt each, the force shift at rc included) and the type
 * charges (nt) through Electrostatic_ewald_r and Electrostatic_ewald_k.  The
 * gradient of the reference's force formulas (lj_cell_list.cpp:143-155,
 * electrostatic_ewald_r.cpp:183-211, electrostatic_ewald_k.cpp:197-226).
 * Parameter values may be changed with the cav_set_* calls after setup; they
 * take effect at the next call that evaluates forces, energies, point fields
 * or this adjoint (owned and ghost charges alike).  NULL outputs skipped. */
int cav_adjoint(cav_ctx *ctx, const double *weights, double *g_epsilon, double *g_sigma,
                double *g_charge);

/* ---- the adjoint without allocation -------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight fitting loops, harnesses that
 * forbid hipMalloc): cav_adjoint's kernels and results, with the call's
 * workspace (residual weights, partial sums, weighted structure factors,
 * per-atom k-space field) carved from one fixed device array per GPU of
 * cav_adjoint_noalloc_scratch_bytes() bytes (32 MiB, held until the process
 * ends) instead of per-call allocations; the call allocates nothing.
 * CAV_ERR_CAPACITY when this system's workspace exceeds the array (about
 * 400k atoms with two types at kmax 16: use cav_adjoint), CAV_ERR_STATE when
 * parameter values were changed after the last force evaluation (applying
 * them allocates: call cav_recalculate or cav_forces first).  Every call on
 * one GPU shares the array and leaves its contents behind: calls that
 * overlap on one GPU (from several host threads) race. */
int cav_adjoint_noalloc(cav_ctx *ctx, const double *weights, double *g_epsilon, double *g_sigma,
                        double *g_charge);
size_t cav_adjoint_noalloc_scratch_bytes(void);

/* ---- batched replicas (no counterpart in the reference) ------------------ */
/* Many independent systems of one force field, each with its own atoms and
 * periodic box, stepped together: every kernel covers all replicas and one
 * velocity-Verlet step of all of them is one launch of a captured hipGraph.
 * Each replica follows the single-system semantics of this header (the
 * reference's Md_simulator step on its own Atom_data / Domain): its own
 * ghosts, rebuild decisions and Verlet list.  Scope: periodic boxes, Verlet
 * list from cells, Lj_cell_list + Electrostatic_ewald_r + Electrostatic_ewald_k
 * (no lambda, dipole or slab). */
typedef struct cav_batch cav_batch;
int cav_batch_create(int device, int precision, cav_batch **out);
int cav_batch_destroy(cav_batch *b);
const char *cav_batch_last_error(const cav_batch *b);
/* Settings shared by every replica (meaning as cav_set_types / cav_set_neighbor
 * with make_verlet_list_from_cell_list + build_cell_list / cav_set_lj /
 * cav_set_ewald_r without lambda / cav_set_ewald_k without dipole or slab). */
int cav_batch_set_types(cav_batch *b, int ntypes, const double *mass, const double *charge);
int cav_batch_set_neighbor(cav_batch *b, double ghost_cutoff, double cutoff, double cutoff_extra_coef);
int cav_batch_set_lj(cav_batch *b, double cutoff, const double *epsilon, const double *sigma);
int cav_batch_set_ewald_r(cav_batch *b, double alpha, double k_electrostatic);
int cav_batch_set_ewald_k(cav_batch *b, double alpha, double k_electrostatic, int kx_max, int ky_max,
                          int kz_max);
/* One replica: its box, n atoms (types, AoS positions, AoS velocities or NULL);
 * ids are 0..n-1.  *index (may be NULL) receives its replica number. */
int cav_batch_add_replica(cav_batch *b, const double lower[3], const double upper[3], int n,
                          const int *type, const double *pos, const double *vel, int *index);
int cav_batch_num_replicas(const cav_batch *b, int *n);
int cav_batch_num_atoms(const cav_batch *b, int replica, int *n);
/* Md_simulator::setup for every replica; then nsteps steps of all replicas,
 * each step one graph launch (use_graph != 0) or the same kernels launched
 * one by one.  Capacities grow transparently; CAV_ERR_NAN restores the state
 * before the offending step. */
int cav_batch_setup(cav_batch *b);
int cav_batch_step(cav_batch *b, double dt, int nsteps, int use_graph);
/* Per replica: owned arrays (which = CAV_POS/VEL/ACC/ACC_OLD, AoS 3n doubles),
 * virial, Verlet pairs with the reference's half-list membership. */
int cav_batch_download(cav_batch *b, int replica, int which, double *out);
int cav_batch_virial(cav_batch *b, int replica, double *virial);
int cav_batch_get_pairs(cav_batch *b, int replica, int *ij, long capacity, long *npairs);
/* Autotuner: times the batched Ewald_k passes for each candidate tile shape on
 * this batch and keeps the fastest; cav_batch_tiles reports the choice (S-pass
 * row tiles per wave, atoms per LDS stage (32 in the split form: one table
 * block), atoms per chunk; force-pass atom tiles per wave). */
int cav_batch_autotune(cav_batch *b);
/* cav_adjoint for every replica at its current positions: weights holds the
 * replicas' residual weights one after another (3 n_r doubles each, replica
 * order); outputs are [replica][nt*nt], [replica][nt*nt], [replica][nt]. */
int cav_batch_adjoint(cav_batch *b, const double *weights, double *g_epsilon, double *g_sigma,
                      double *g_charge);
int cav_batch_tiles(const cav_batch *b, int *s_rows, int *s_stage, int *s_chunk, int *f_tiles);

/* ---- tuning (no counterpart in the reference) ------------------------------ */
/* Autotuner: times the Ewald_k passes for each candidate tile shape on this
 * context's own system, as cav_step runs them (S-pass row tiles per wave, atoms
 * per chunk; split form: force-pass atom tiles per wave and waves per
 * workgroup; the split S pass stages 32 atoms, one table block) and keeps the
 * fastest; the forces of the current state are recomputed afterwards.
 * cav_tiles reports the tiles in use (0 = the built-in choice). */
int cav_autotune(cav_ctx *ctx);
int cav_tiles(const cav_ctx *ctx, int *s_rows, int *s_stage, int *s_chunk, int *f_tiles,
              int *f_waves);
/* Pins the Ewald_k tiles instead of tuning them (reproducible timings): S-pass
 * row tiles per wave 1..4, atoms per chunk (a multiple of 32, 32..4096: the
 * mixed form accumulates a chunk in FP32); split form:
 * force-pass atom tiles per wave 1..2 and waves per workgroup 8 or 16.  After
 * cav_setup; the forces of the current state are recomputed.  FP64 uses the
 * chunk only. */
int cav_set_tiles(cav_ctx *ctx, int s_rows, int s_chunk, int f_tiles, int f_waves);
/* cav_step runs Ewald_k on a second HIP stream, concurrently with the ghost
 * exchange, neighbour build and pair kernel (on = 1, the default); 0 runs
 * every pass in order on one stream (per-kernel timings without contention).
 * The accelerations are the same either way. */
int cav_set_overlap(cav_ctx *ctx, int on);
/* Short-range pair kernel.  CAV_PAIRS_AUTO (default): the row kernel (8 lanes
 * per atom) reads each Verlet row as the bin-order slots the list build
 * recorded and gathers the neighbours from bin-ordered copies of the atoms,
 * where the neighbours of cell-ordered atoms lie close together; the atom-index
 * rows (cav_num_pairs, the other kernels) are filled from the slots on demand.
 * CAV_PAIRS_ROWS gathers the same rows by atom index; CAV_PAIRS_CELLS stages
 * each cell's 27-cell stencil in LDS and reads the neighbours from there
 * (where it applies: list built from cells, no molecule lambda; otherwise
 * rows).  AUTO and ROWS give the same forces bit for bit, CELLS to rounding.
 * A new mode takes effect at the next list build. */
enum cav_pair_kernel { CAV_PAIRS_AUTO = 0, CAV_PAIRS_ROWS = 1, CAV_PAIRS_CELLS = 2 };
int cav_set_pair_kernel(cav_ctx *ctx, int kind);
int cav_pair_kernel(const cav_ctx *ctx, int *kind_in_use);

/* ---- measurement ----------------------------------------------------------- */
/* Statistics of the current Verlet list: list entries per atom with the
 * reference's half-list membership, and those inside the LJ cutoff. */
int cav_list_stats(cav_ctx *ctx, double *pairs_per_atom, double *lj_pairs_per_atom);
/* Average device time (ms) per launch of each kernel family over the last
 * cav_step call, measured with HIP events on the context's stream.
 * names: "pair", "ewald_k_sk", "ewald_k_force", "ewald_k_tables", "neighbor", "integrate". */
int cav_kernel_time(cav_ctx *ctx, const char *name, double *ms_per_launch, int *launches);
int cav_enable_timing(cav_ctx *ctx, int on);

#ifdef __cplusplus
}
#endif
#endif
