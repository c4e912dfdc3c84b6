// Internal state and helpers of the caviar_amd HIP library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/caviar_amd.h"

#define CAV_PI 3.14159265358979323846
#define CAV_2PI 6.28318530717958647692     // macro_constants.h:22
#define CAV_4PI 12.56637061435917295384    // macro_constants.h:24
#define CAV_PIS_INV 0.56418958354775628695 // macro_constants.h:28

namespace cav {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// Device status words, read back by the host after every step.
enum Flag : int {
  F_WRAPPED = 0,   // an owned atom was shifted by one box length this step
  F_REBUILD = 1,   // the Verlet list is rebuilt this step
  F_NAN = 2,       // a NaN acceleration was seen before the drift
  F_OVERFLOW = 3,  // a Verlet row exceeded the list capacity
  F_NGHOST = 4,    // ghost count of the current step
  F_NGHOST_OLD = 5,// ghost count at the last list build
  F_MAXNB = 6,     // longest overflowing Verlet row at the last build (set with F_OVERFLOW)
  F_FORCE = 7,     // host-requested rebuild
  F_GOVER = 8,     // batched replicas: ghosts exceeded the replica's ghost capacity
  F_MAXST = 9,     // largest cell stencil (candidate atoms) the cell-staged pair kernel met
  F_NL_STALE = 10, // the last list build wrote bin-order slots only: index rows are filled on demand
  F_COUNT = 16
};

// IEEE double operations rounded one at a time, never contracted into an FMA
// (hipcc contracts by default): where a decision or a trajectory must follow
// the reference's rounding exactly.
__device__ __forceinline__ double xadd(double a, double b)
{
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ double xsub(double a, double b)
{
#pragma clang fp contract(off)
  return a - b;
}
__device__ __forceinline__ double xmul(double a, double b)
{
#pragma clang fp contract(off)
  return a * b;
}

constexpr int kBlock = 256;
#ifndef CAV_KLPA
#define CAV_KLPA 8
#endif
constexpr int kLPA = CAV_KLPA;  // lanes per atom in the row-gather pair kernel

struct Timer {
  hipEvent_t a = nullptr, b = nullptr;
  double total_ms = 0;
  int launches = 0;
  bool pending = false;
};

struct EwaldKPlan {
  // k lattice: nx in [0,kx] (half space), ny in [-ky,ky], nz in [-kz,kz]
  int kx = 0, ky = 0, kz = 0, nX = 0, nY = 0, nZ = 0;
  int R = 0;      // rows (nx,ny)
  int Rpad = 0;   // rows rounded up to 16
  int KZ = 0;     // 2*nZ rounded up to a multiple of 4 (real-embedded nz axis)
  int nslab = 1;  // K slabs of the FP32/FP64 force pass (at most 32 K-steps each)
  int KSS = 0;    // K-steps per slab (the last slab may hold fewer)
  int KSP = 0;    // K-steps of a slab padded to whole, odd 16-byte vectors: fragment stride
  int nk_ref = 0; // (2kx)(2ky)(2kz), the reference's used count
  double lx = 1, ly = 1, lz = 1, lxyz_inv = 0, dip_coef = 0;
  std::vector<double> W;        // R*nZ: multiplicity(k) * f(k) over the half space
  std::vector<int> ref_r, ref_c, ref_conj;  // reference k order -> (row, col, conjugate?)
};

// One replica of a batch as the device kernels see it (see cav_batch.hip).
// Slot arrays hold each replica's owned atoms then its ghosts from `base`;
// owned-indexed arrays (velocities, accelerations, Verlet rows, FP64 phase
// tables) hold its owned atoms from `obase`; the split form's blocked phase
// tables hold them from `tbase`, a multiple of 32.
struct RepDev {
  double lo[3], hi[3], L[3], il[3];
  double c0, kxs, kys, kzs;  // k-space force constants of this box
  int base, n, gcap, obase, tbase;
  int cbase, nbins, nb[3];
  int pbase;                 // first per-block partial slot
};

}  // namespace cav

struct cav_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int precision = CAV_FP64;

  // domain
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, L[3] = {0, 0, 0};
  int bc[3] = {0, 0, 0};
  bool box_set = false;
  double ghost_cutoff = 0;

  // types
  int nt = 0;
  std::vector<double> mass, minv, charge;

  // neighbour list
  bool nl_on = false;
  double nl_cutoff = 0, nl_coef = 1.12347, rce = 0, rce2 = 0, thr_sq = 0, cinv = 0;
  int from_cells = 0, build_cells = 0, all_atom_test = 0, rebuild_test = 0;
  int nb[3] = {0, 0, 0}, nbins = 0;
  std::vector<int> st_off, st;
  // the list build's grid: the reference grid plus a layer of cells past
  // every face (nb + 2 per axis) with plain 3x3x3 stencils
  int xnbins = 0;

  // force fields
  bool lj_on = false;
  double lj_rc = 0;
  std::vector<double> eps, sig;
  bool lj_cellwalk = false;  // LJ walks the cells every step (cutoff beyond the list guarantee)
  bool er_on = false;
  double er_alpha = 1, er_ke = 1;
  bool lambda_set = false;
  std::vector<double> lambda;
  // force_field lj (Verlet list): type matrices resolved as Lj::verify_settings
  // resolves them, per-pair cutoffs (cutoff_list or wca), flags
  bool ljl_on = false, ljl_cut_on = false;
  double ljl_rc = 0, ljl_jump_tol = 0;
  int ljl_ignore_intra = 0, ljl_jump = 0;
  std::vector<double> ljl_eps, ljl_sig, ljl_cut;
  // Electrostatic_short_range
  bool sr_on = false, sr_lambda_set = false;
  double sr_rc = 0, sr_ke = 1, sr_beta = 1, sr_ext[3] = {0, 0, 0};
  std::vector<double> sr_lambda;
  // temperature / pressure (atom_data k_b, n_r_df, temperature_process,
  // pressure_process) and the velocity-rescaling constraints of a step
  double k_b = -1.0;
  int n_r_df = -1;
  bool temperature_process = false, pressure_process = false, thermo_on = false;
  bool ber_on = false, nve_on = false;
  int ber_step = 1, nve_step = 1;
  double ber_t0 = 0, ber_coupling = 1, ber_dt = 1, ber_lmin = 0.8, ber_lmax = 1.25;
  double nve_e_dof = -1, nve_e_tot = -1, nve_kbt = -1, nve_kb = -1, nve_t = -1;
  long step_index = 0;          // Md_simulator::current_step of the next step
  double *d_thermo = nullptr;   // scalars and per-block partials
  double *d_mass = nullptr;     // per-type mass
  // Electrostatic_ewald_slab_correction (ELC)
  bool elc_on = false;
  int elc_axis = 2, elc_kx = 1, elc_ky = 1;
  double elc_lz = 1, elc_ke = 1, elc_sum_e_coef = 0, elc_dip_coef = 0;
  double *d_elc_k = nullptr;    // k grids and coefficients
  double *d_elc_chi = nullptr;  // chi sums of the last evaluation
  double *d_elc_dip = nullptr;  // previous dipole sum, dipole field; saved copy
  bool ek_on = false;
  double ek_alpha = 0, ek_ke = 1, eps_dipole = 1, slab_lz = 1;
  int kx = 1, ky = 1, kz = 1, dipole = 0, slab_axis = -1;
  cav::EwaldKPlan ek;

  // atoms
  int n = 0, gcap = 0;
  std::vector<int> h_id, h_type, h_mol;

  // device arrays: positions/type/id/mol/charge over owned + ghost capacity
  double *x = nullptr, *y = nullptr, *z = nullptr;
  double *xo = nullptr, *yo = nullptr, *zo = nullptr;  // positions at the last list build
  double *vx = nullptr, *vy = nullptr, *vz = nullptr;
  double *ax = nullptr, *ay = nullptr, *az = nullptr;
  double *aox = nullptr, *aoy = nullptr, *aoz = nullptr;
  double *snap = nullptr;  // 9n doubles: pos/vel/acc at the start of a step
  int *type = nullptr, *id = nullptr, *mol = nullptr;
  double *q = nullptr;
  double4 *xq = nullptr;   // packed (x, y, z, q) for the pair gathers
  int2 *tid = nullptr;     // packed (type, id)
  int *gsrc = nullptr;
  int *flags = nullptr;
  int *h_flags = nullptr;  // pinned
  int *blk = nullptr;      // scan block sums
  int nblk = 0;

  // cells
  int *bin_of = nullptr, *bin_cnt = nullptr, *bin_start = nullptr, *bin_items = nullptr;
  int *d_st_off = nullptr, *d_st = nullptr;
  int *d_xst_off = nullptr, *d_xst = nullptr;  // stencils of the list build's grid
  int *scan_tmp = nullptr;
  int *order = nullptr;    // owned atoms in cell order (compaction of bin_items)
  double4 *cpos = nullptr; // (x, y, z, index) in bin order, for the list build
  int *binpos = nullptr;   // bin-order position of every slot at the last list build
  int *lstart = nullptr, *litems = nullptr;  // cells of the last list build
  double4 *xqs = nullptr;  // (x, y, z, q) in bin order, refreshed every step
  int2 *tids = nullptr;    // (type, id) in bin order
  unsigned short *nll = nullptr;  // Verlet entries as positions in the row's cell staging order
  int *nlb = nullptr;             // Verlet entries as bin-order slots of the list build
  int *ord_blk = nullptr;  // scan block sums for the compaction

  // list (row-major: nl[i*nl_cap + k], nn[i] entries in row i)
  int *nl = nullptr, *nn = nullptr;
  int nl_cap = 0;

  // per-type tables on the device
  double *d_minv = nullptr, *d_charge = nullptr, *d_eps = nullptr, *d_sig = nullptr,
         *d_lambda = nullptr;
  double *d_raw_eps = nullptr, *d_raw_sig = nullptr;  // epsilon/sigma matrices as given
  double *d_ljl = nullptr;        // force_field lj: [c12 | c6 | cshift | rc2] per type pair
  double *d_sr_lambda = nullptr;  // Electrostatic_short_range lambda per type pair
  bool params_dirty = false;  // parameter values changed after setup

  // reductions: per-block partials and results
  double *part = nullptr;  // per-block partial sums (virial, energies)
  int npart = 0;
  // Ewald_k of a step on a second stream, concurrent with the short-range
  // passes: its own accelerations and partial sums, added after the join
  bool overlap = true;
  hipStream_t kstream = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  double *ekx = nullptr, *eky = nullptr, *ekz = nullptr, *part_k = nullptr;
  // the side stream's Ewald_k leaves the dipole term to the main stream, which
  // adds it after the k-space sum (the in-order rounding: (a_sr + a_k) + a_dip)
  bool defer_dipole = false;
  double *scal = nullptr;  // [0] virial, [1] e_lj, [2] e_r, [3] e_k, [4..] scratch
  double *h_scal = nullptr;

  // Ewald_k device buffers
  void *tab = nullptr;       // phase tables (T): X[nX][n], Y[nY][n], Z[nZ][n] complex
  size_t tab_bytes = 0;
  int tab_cap = 0;           // atoms the tables hold
  double *sk_part = nullptr; // [chunks][Rpad][2nZ]
  int sk_chunks = 0;
  double *sk = nullptr;      // [Rpad][2nZ] FP64
  double *sk_sum = nullptr;  // [Rpad][SC] chunk partials summed (fixed order)
  void *bmat = nullptr;      // [KZ][4*Rpad] (T) force-pass operand
  // mixed form: split-FP16 Z records and force-pass operand (hi + lo halves)
  void *ztab = nullptr;      // [n][nZ] x 16 B: (Zr,-Zi) hi, (Zr,-Zi) lo, (Zi,Zr) hi, (Zi,Zr) lo
  void *bmat_h = nullptr;    // [Rpad/16][4][NKH][64] x 32 B: 8 hi halves, 8 lo halves
  int nkh = 0;               // 32-deep K-steps along the force-pass reduction axis
  double sA = 1, sB = 1;     // power-of-two operand scales of the split form
  // Ewald_k tiles (cav_autotune): S-pass row tiles per wave, atoms per LDS
  // stage (0 = largest that fits), atoms per chunk (0 = default), force-pass
  // atom tiles per wave of the split form
  int tune_rt = 3, tune_stage = 0, tune_chunk = 0, tune_fnt = 2, tune_fw = 8;
  int pair_kind = 0;  // cav_pair_kernel
  // staging positions (nll) valid for the current list / written by the last
  // conditional build launched
  bool nll_built = false, nll_pending = false;
  // bin-order slots (nlb) valid for the current list / written by the last
  // conditional build launched
  bool nlb_built = false, nlb_pending = false;
  bool tuned = false;
  double *d_W = nullptr;     // [Rpad][nZ]
  double *kacc = nullptr;    // per-atom k-space (fx, fy, fz, phi) scratch for point queries

  bool setup_done = false;
  bool list_valid = false;
  bool cells_current = false;

  bool timing = false;
  cav::Timer t_pair, t_sk, t_force, t_tab, t_nb, t_int, t_list;
};

// Many independent single-domain systems (replicas) of one force field, each
// with its own atoms and periodic box, stepped together: every kernel covers
// all replicas through the replica table, and one step is one graph launch.
struct cav_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int precision = CAV_FP64;

  // shared settings
  int nt = 0;
  std::vector<double> mass, minv, charge;
  double ghost_cutoff = 0;
  double nl_cutoff = 0, nl_coef = 1.12347, rce = 0, rce2 = 0, thr_sq = 0, cinv = 0;
  bool nl_on = false, lj_on = false, er_on = false, ek_on = false;
  double lj_rc = 0;
  std::vector<double> eps, sig;
  double er_alpha = 1, er_ke = 1;
  double ek_alpha = 0, ek_ke = 1;
  int kx = 1, ky = 1, kz = 1;

  // replicas as given (host)
  struct HostRep {
    double lo[3], hi[3];
    int n = 0, gcap = 0;
    std::vector<int> type;
    std::vector<double> pos, vel;  // AoS
  };
  std::vector<HostRep> hreps;
  std::vector<cav::RepDev> reps;
  int B = 0;
  long nslots = 0, nown = 0, ntab = 0;
  int ncells = 0, maxn = 0, pair_blocks = 0, scan_blocks = 0;
  std::vector<int> h_cell_rep, h_st_off, h_st;

  // device
  cav::RepDev *d_rep = nullptr;
  int *cell_rep = nullptr, *st_off = nullptr, *st = nullptr;
  double *x = nullptr, *y = nullptr, *z = nullptr;     // slots
  double *xo = nullptr, *yo = nullptr, *zo = nullptr;  // slots at the last list build
  int *type = nullptr;                                 // slots
  int *id = nullptr;                                   // slots: local owned index (ghost: its source)
  double *q = nullptr;                                 // slots
  double4 *xq = nullptr;
  int2 *tid = nullptr;
  double *vx = nullptr, *vy = nullptr, *vz = nullptr;  // owned
  double *ax = nullptr, *ay = nullptr, *az = nullptr;
  double *aox = nullptr, *aoy = nullptr, *aoz = nullptr;
  int *type_o = nullptr;                               // owned
  double *q_o = nullptr;
  double *snap = nullptr;  // [9][nown]: positions, velocities, accelerations
  int *bin_of = nullptr, *bin_cnt = nullptr, *bin_start = nullptr, *bin_items = nullptr;
  double4 *cpos = nullptr; // slots: (x, y, z, slot) in bin order
  int *gblk = nullptr;     // [B][scan_blocks] ghost-count block sums
  int *nl = nullptr, *nn = nullptr;
  // the pair kernel's bin-order gathers: rows as replica-local bin slots
  // (beside nl), bin-ordered copies of (x, y, z, q) and (type, id), slot ->
  // local atom, local atom -> slot, owned atoms in cell order, and per cell
  // the owned count and its offset within the replica
  int *nlb = nullptr;
  double4 *xqs = nullptr;
  int2 *tids = nullptr;
  int *litems = nullptr, *binpos = nullptr, *order = nullptr, *ocnt = nullptr, *ostart = nullptr;
  int cap = 0;
  int *flags = nullptr, *h_flags = nullptr;  // [B][F_COUNT]
  double *d_minv = nullptr, *d_lj = nullptr;
  double *d_raw_eps = nullptr, *d_raw_sig = nullptr, *d_charge = nullptr;  // for the adjoint
  double *part = nullptr;  // [B][pair_blocks] virial partials
  double *vir = nullptr;   // [B]

  // Ewald_k
  cav::EwaldKPlan ek;      // lattice shared by the replicas (W of replica 0)
  std::vector<double> W;   // [B][R*nZ]
  double *d_W = nullptr;
  void *tab = nullptr, *ztab = nullptr, *bmat = nullptr, *bmat_h = nullptr;
  double *sk_part = nullptr, *sk = nullptr, *epart = nullptr;
  // S-pass items in table positions (blocked in the split form) and in owned
  // positions (FP32/FP64 tables)
  int4 *s_items = nullptr, *s_items_o = nullptr, *f_items = nullptr;
  int2 *rchunks = nullptr;
  int n_sitems = 0, n_fitems = 0, n_chunks = 0, nkh = 0;
  long bm_stride = 0, bmh_stride = 0;  // bytes (FP operand) / halves (split operand) per replica
  double sA = 1, sB = 1;

  // tiles chosen by the autotuner (cav_batch_autotune)
  int s_rt = 3, s_stage = 32, s_chunk = 1024, f_nt = 2;
  bool tuned = false;

  // one step as a graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  double graph_dt = 0;
  bool setup_done = false;
};

// helpers shared by the translation units
namespace cav {
int fail(cav_ctx *c, int code, const std::string &msg);
int hip_check(cav_ctx *c, hipError_t e, const char *what);
void timer_begin(cav_ctx *c, Timer &t);
int tuning_phase(cav_ctx *c);
int refresh_params(cav_ctx *c);  // parameter values set after setup -> device
void timer_end(cav_ctx *c, Timer &t);

// launchers (defined in the .hip files)
int launch_ghosts(cav_ctx *c);
int launch_rebuild_check(cav_ctx *c);
// for_list: the binning of a list build, which also keeps the bin order the
// list refers to (litems, binpos)
int launch_bins(cav_ctx *c, bool conditional, bool for_list = false);
int launch_list(cav_ctx *c, bool conditional);
int launch_save_old(cav_ctx *c, bool conditional);
bool cell_staged(const cav_ctx *c);
bool sorted_rows(const cav_ctx *c);
// Whether list builds bin on the grid with the ghost layer: the cell walks
// (LJ beyond the list cutoff) and the cell-staged pair kernel read the cells
// of the reference grid.
inline bool list_on_ghost_grid(const cav_ctx *c)
{
  return !c->lj_cellwalk && c->pair_kind != CAV_PAIRS_CELLS && !c->all_atom_test;
}
int launch_nl_fill(cav_ctx *c);
int launch_pairs(cav_ctx *c, bool lj, bool er, bool accumulate);
int launch_pack(cav_ctx *c);
int launch_pairs_list(cav_ctx *c, bool ljl, bool sr, bool accumulate, bool pack);  // packed (x, y, z, q), (type, id) of owned + ghost slots
int launch_lj_cellwalk(cav_ctx *c);
int launch_zero_acc(cav_ctx *c);
int launch_ewald_k(cav_ctx *c);
int launch_dipole(cav_ctx *c);  // epsilon_dipole term into the accelerations
int thermo_alloc(cav_ctx *c);     // thermo scalars and partials (setup)
int thermo_temperature(cav_ctx *c);  // finalize_temperature (with temperature_process)
int thermo_end_of_step(cav_ctx *c, long ts);  // temperature, thermostats, pressure
int thermo_read(cav_ctx *c, double *ke, double *temp, double *press);
int elc_plan(cav_ctx *c);        // ELC k grids and buffers (setup)
int launch_elc(cav_ctx *c);      // ELC chi sums + forces into the accelerations
int elc_rewind(cav_ctx *c);       // ELC dipole state back before a re-evaluation
int energy_elc(cav_ctx *c, double *e);
int point_field_elc(cav_ctx *c, int np, const double *pts, double *field, double *pot);
int forces_all(cav_ctx *c);
int launch_drift(cav_ctx *c, double dt);
int launch_kick(cav_ctx *c, double dt);
int launch_begin_step(cav_ctx *c);
int ewald_k_plan(cav_ctx *c);
const char *plan_k_lattice(EwaldKPlan &p, int kx, int ky, int kz, int precision, double alpha,
                           double lx, double ly, double lz);
void split_scales(double qmax, double qsum, double wmax, double *sA, double *sB);
int ewald_k_alloc(cav_ctx *c);
int ewald_k_autotune(cav_ctx *c);
int launch_energy_r(cav_ctx *c, double *e);
int launch_energy_lj(cav_ctx *c, double *e);
int launch_energy_k(cav_ctx *c, double *e);
int point_field_r(cav_ctx *c, int np, const double *pts, double *field, double *pot);
int point_field_k(cav_ctx *c, int np, const double *pts, double *field, double *pot);
int reduce_partials(cav_ctx *c, int count, int slot, bool add = false);
int bfail(cav_batch *b, int code, const std::string &msg);
int bhip_check(cav_batch *b, hipError_t e, const char *what);
int batch_neighbor(cav_batch *b, bool forced);   // ghosts, rebuild test, cells, list
int batch_pairs(cav_batch *b);                   // pack + LJ/Ewald_r + virial
int batch_ewald_k_plan(cav_batch *b);
int batch_ewald_k_alloc(cav_batch *b);
int batch_ewald_k(cav_batch *b);
int batch_ewald_k_field(cav_batch *b, double *out4);
int batch_weighted_sk(cav_batch *b, const double *u, double *out);

inline int grid_for(long n, int block = kBlock) { return (int)((n + block - 1) / block); }

}  // namespace cav

#define CAV_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) return cav::hip_check(c, e_, #call);               \
  } while (0)

#define BAT_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) return cav::bhip_check(b, e_, #call);              \
  } while (0)

#define CAV_TRY(call)                                                        \
  do {                                                                       \
    int r_ = (call);                                                         \
    if (r_ != CAV_OK) return r_;                                             \
  } while (0)
