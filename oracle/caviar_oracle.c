/* CPU oracle for the CAVIAR short-range + Ewald force path (TEST INFRASTRUCTURE).
 *
 * Plain C99, FP64, compiled with -ffp-contract=off so that every expression is
 * rounded in the order the reference's C++ expressions are (x86-64 g++ -O3
 * without -march emits no FMA).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, and only as the checker.
 *
 * Citations are reference file:line (paths below src/caviar/objects/ unless
 * given in full).
 */
#include "caviar_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <time.h>

#define ORC_4PI 12.56637061435917295384    /* macro_constants.h:24 */
#define ORC_2PI 6.28318530717958647692     /* macro_constants.h:22 */
#define ORC_PIS_INV 0.56418958354775628695 /* macro_constants.h:28 */

struct orc_sys {
  double lo[3], hi[3], L[3];
  int bc[3];
  double ghost_cutoff;
  int nt;
  double *mass, *minv, *charge;
  int n;
  int *id, *type, *mol;
  double *pos, *vel, *acc, *acc_old;
  int g, gcap;
  double *gpos;
  int *gsrc;
  /* neighbour list */
  double nl_cutoff, nl_coef, rce, thr_sq, cinv;
  int from_cells, build_cells, all_atom_test, rebuild_test, nl_on;
  int nb[3], nbins;
  int *st_off, *st;            /* de-duplicated 27-stencil per bin, linear indices */
  int *bin_off, *bin_items;    /* cell list CSR, insertion order */
  long *nl_off;
  int *nl_items;
  long nl_cap;
  double *pos_old;
  int n_old;
  double *gpos_old;
  int g_old;
  /* Lj_cell_list */
  int lj_on;
  double lj_rc;
  double *eps, *sig;
  /* force_field lj (Verlet-list LJ): resolved type matrices */
  int ljl_on, ljl_cut_on, ljl_ignore_intra, ljl_jump;
  double ljl_rc, ljl_jump_tol;
  double *ljl_eps, *ljl_sig, *ljl_cut;
  /* Electrostatic_short_range */
  int sr_on, sr_lambda_set;
  double sr_rc, sr_ke, sr_beta, sr_ext[3];
  double *sr_lambda;
  /* Electrostatic_ewald_slab_correction (ELC) */
  int elc_on, elc_init, elc_axis, elc_kx, elc_ky;
  double elc_ke, elc_lz, elc_lx, elc_ly, elc_hz, elc_sum_e_coef, elc_dip_coef;
  double elc_dip_sum[3], elc_dip_field[3];
  double *elc_kxv, *elc_kyv, *elc_kpv, *elc_kxc, *elc_kyc, *elc_kpc;
  double *elc_chi_x, *elc_chi_y, *elc_chi_p; /* [kx][2][2], [ky][2][2], [kx*ky][2][2][2] */
  /* Electrostatic_ewald_r */
  int er_on, lambda_set;
  double er_alpha, er_ke;
  double *lambda;
  /* Electrostatic_ewald_k */
  int ek_on, ek_init, kx, ky, kz, dipole, slab_axis, nk;
  double ek_alpha, ek_ke, eps_dipole, slab_lz, lxyz_inv, dip_coef;
  double dip_sum[3], dip_field[3];
  double *kvec, *fk, *sk;
  double virial;
  /* temperature / pressure (atom_data.cpp:121-151, kinetic_energy.cpp) */
  double k_b;
  int n_r_df, temperature_process, pressure_process;
  double temperature, pressure;
  long step_index; /* Md_simulator::current_step of the next orc_step */
  /* constraint berendsen (constraint/berendsen.cpp) and nve (constraint/nve.cpp) */
  int ber_on, ber_step;
  double ber_t0, ber_coupling, ber_dt, ber_lmin, ber_lmax;
  int nve_on, nve_step;
  double nve_e_dof, nve_e_tot, nve_kbt, nve_kb, nve_t;
};

static int g_threads = 1;
void orc_set_threads(int nthreads) { g_threads = nthreads > 0 ? nthreads : 1; }

orc_sys *orc_create(void)
{
  orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
  s->nl_coef = 1.12347; /* neighborlist.h:134 */
  s->slab_axis = -1;
  s->k_b = -1.0;   /* atom_data.cpp:54 */
  s->n_r_df = -1;  /* from the boundary conditions (atom_data.cpp:322-350) */
  return s;
}

void orc_destroy(orc_sys *s)
{
  if (!s)
    return;
  void *ptrs[] = {s->mass, s->minv, s->charge, s->id, s->type, s->mol, s->pos, s->vel,
                  s->acc, s->acc_old, s->gpos, s->gsrc, s->st_off, s->st, s->bin_off,
                  s->bin_items, s->nl_off, s->nl_items, s->pos_old, s->gpos_old, s->eps,
                  s->sig, s->lambda, s->kvec, s->fk, s->sk, s->ljl_eps, s->ljl_sig,
                  s->ljl_cut, s->sr_lambda, s->elc_kxv, s->elc_kyv, s->elc_kpv, s->elc_kxc,
                  s->elc_kyc, s->elc_kpc, s->elc_chi_x, s->elc_chi_y, s->elc_chi_p};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i)
    free(ptrs[i]);
  free(s);
}

static void *dup_arr(const void *src, size_t bytes)
{
  void *p = malloc(bytes ? bytes : 1);
  if (src && bytes)
    memcpy(p, src, bytes);
  else
    memset(p, 0, bytes ? bytes : 1);
  return p;
}

void orc_set_box(orc_sys *s, const double lo[3], const double hi[3], const int bc[3])
{
  for (int d = 0; d < 3; ++d) {
    s->lo[d] = lo[d];
    s->hi[d] = hi[d];
    s->L[d] = hi[d] - lo[d]; /* domain.cpp:470-477 size_global = upper - lower */
    s->bc[d] = bc[d];
  }
}

void orc_set_ghost_cutoff(orc_sys *s, double gc) { s->ghost_cutoff = gc; }

void orc_set_types(orc_sys *s, int nt, const double *mass, const double *charge)
{
  free(s->mass);
  free(s->minv);
  free(s->charge);
  s->nt = nt;
  s->mass = (double *)dup_arr(mass, sizeof(double) * nt);
  s->charge = (double *)dup_arr(charge, sizeof(double) * nt);
  s->minv = (double *)malloc(sizeof(double) * (nt ? nt : 1));
  for (int t = 0; t < nt; ++t) /* atom_data.cpp:236-239 */
    s->minv[t] = (mass[t] == 0.0) ? 0.0 : 1.0 / mass[t];
}

void orc_set_atoms(orc_sys *s, int n, const int *id, const int *type, const int *mol,
                   const double *pos, const double *vel)
{
  free(s->id);
  free(s->type);
  free(s->mol);
  free(s->pos);
  free(s->vel);
  free(s->acc);
  free(s->acc_old);
  s->n = n;
  s->id = (int *)dup_arr(id, sizeof(int) * n);
  s->type = (int *)dup_arr(type, sizeof(int) * n);
  s->mol = (int *)malloc(sizeof(int) * (n ? n : 1));
  for (int i = 0; i < n; ++i)
    s->mol[i] = mol ? mol[i] : -1; /* atom_data.cpp:509 */
  s->pos = (double *)dup_arr(pos, sizeof(double) * 3 * n);
  s->vel = (double *)dup_arr(vel, sizeof(double) * 3 * n);
  s->acc = (double *)dup_arr(NULL, sizeof(double) * 3 * n);
  s->acc_old = (double *)dup_arr(NULL, sizeof(double) * 3 * n);
}

void orc_set_neighbor(orc_sys *s, double cutoff, double coef, int from_cells, int build_cells,
                      int all_atom_test, int rebuild_test)
{
  s->nl_on = 1;
  s->nl_cutoff = cutoff;
  s->nl_coef = coef;
  s->from_cells = from_cells;
  /* make_verlet_list_from_cell_list implies build_cell_list (neighborlist.cpp:216-220) */
  s->build_cells = build_cells || from_cells;
  s->all_atom_test = all_atom_test;
  s->rebuild_test = rebuild_test;
}

void orc_set_lj(orc_sys *s, double rc, const double *eps, const double *sigma)
{
  s->lj_on = 1;
  s->lj_rc = rc;
  free(s->eps);
  free(s->sig);
  s->eps = (double *)dup_arr(eps, sizeof(double) * s->nt * s->nt);
  s->sig = (double *)dup_arr(sigma, sizeof(double) * s->nt * s->nt);
}

/* force_field lj: Lj::read + Lj::verify_settings (lj.cpp:42-147,162-339) for
 * nt types: matrices by array (eps/sig, nt*nt, unset entries 0) or by atom
 * (eps_atom/sig_atom, nt, Lorentz-Berthelot off the diagonal, lj.cpp:207-222);
 * cutoff_list (nt*nt, unset 0) or wca (2^(1/6) sigma_ij, before the lambda_s
 * scaling, lj.cpp:278-294); lambda_s scales sigma, lambda_e epsilon (nt*nt,
 * unset 1; lj.cpp:310-338).  Returns 1 when both input forms are given
 * (lj.cpp:167-170) or neither. */
int orc_set_lj_list(orc_sys *s, double cutoff, const double *eps, const double *sig,
                    const double *eps_atom, const double *sig_atom, const double *cutoff_list,
                    int wca, const double *lambda_s, const double *lambda_e,
                    int ignore_intra_molecule, int jump_fix, double jump_tol)
{
  const int nt = s->nt, nt2 = nt * nt;
  const int by_array = eps || sig, by_atom = eps_atom || sig_atom;
  if ((by_array && by_atom) || (!by_array && !by_atom))
    return 1;
  free(s->ljl_eps);
  free(s->ljl_sig);
  free(s->ljl_cut);
  s->ljl_eps = (double *)calloc(nt2 ? nt2 : 1, sizeof(double));
  s->ljl_sig = (double *)calloc(nt2 ? nt2 : 1, sizeof(double));
  s->ljl_cut = (double *)calloc(nt2 ? nt2 : 1, sizeof(double));
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) {
      const int p = i * nt + j;
      if (by_atom) {
        const double ea = eps_atom ? eps_atom[i] : 0, eb = eps_atom ? eps_atom[j] : 0;
        const double sa = sig_atom ? sig_atom[i] : 0, sb = sig_atom ? sig_atom[j] : 0;
        if (i == j) {
          s->ljl_sig[p] = sa;
          s->ljl_eps[p] = ea;
        } else {
          s->ljl_sig[p] = 0.5 * (sa + sb);
          s->ljl_eps[p] = sqrt(ea * eb);
        }
      } else {
        s->ljl_eps[p] = eps ? eps[p] : 0;
        s->ljl_sig[p] = sig ? sig[p] : 0;
      }
    }
  s->ljl_cut_on = wca || cutoff_list != NULL;
  if (wca) {
    const double cut_coef = pow(2.0, 1.0 / 6.0);
    for (int p = 0; p < nt2; ++p)
      s->ljl_cut[p] = cut_coef * s->ljl_sig[p];
  } else if (cutoff_list) {
    memcpy(s->ljl_cut, cutoff_list, sizeof(double) * nt2);
  }
  for (int p = 0; p < nt2; ++p) {
    if (lambda_s)
      s->ljl_sig[p] *= lambda_s[p];
    if (lambda_e)
      s->ljl_eps[p] *= lambda_e[p];
  }
  s->ljl_rc = cutoff;
  s->ljl_ignore_intra = ignore_intra_molecule;
  s->ljl_jump = jump_fix;
  s->ljl_jump_tol = jump_tol;
  s->ljl_on = 1;
  return 0;
}

/* Electrostatic_short_range::read (electrostatic_short_range.cpp:39-99):
 * cutoff, k_electrostatic, beta, external_field, lambda (nt*nt, unset 1). */
void orc_set_short_range(orc_sys *s, double cutoff, double ke, double beta, const double ext[3],
                         const double *lambda)
{
  const int nt2 = s->nt * s->nt;
  s->sr_rc = cutoff;
  s->sr_ke = ke;
  s->sr_beta = beta;
  for (int d = 0; d < 3; ++d)
    s->sr_ext[d] = ext ? ext[d] : 0.0;
  free(s->sr_lambda);
  s->sr_lambda = (double *)dup_arr(lambda, sizeof(double) * (nt2 ? nt2 : 1));
  s->sr_lambda_set = lambda != NULL;
  s->sr_on = 1;
}

/* Electrostatic_ewald_slab_correction::read (electrostatic_ewald_slab_correction.cpp:
 * 44-126): slab_normal_axis (default z), kx_max / ky_max, lz, k_electrostatic. */
void orc_set_elc(orc_sys *s, int axis, int kx, int ky, double lz, double ke)
{
  s->elc_axis = axis;
  s->elc_kx = kx;
  s->elc_ky = ky;
  s->elc_lz = lz;
  s->elc_ke = ke;
  s->elc_init = 0;
  for (int d = 0; d < 3; ++d)
    s->elc_dip_sum[d] = s->elc_dip_field[d] = 0.0; /* Vector<double>() is zero */
  s->elc_on = 1;
}

void orc_set_ewald_r(orc_sys *s, double alpha, double ke, const double *lambda)
{
  s->er_on = 1;
  s->er_alpha = alpha;
  s->er_ke = ke;
  free(s->lambda);
  s->lambda = NULL;
  s->lambda_set = lambda != NULL;
  if (lambda)
    s->lambda = (double *)dup_arr(lambda, sizeof(double) * s->nt * s->nt);
}

void orc_set_ewald_k(orc_sys *s, double alpha, double ke, int kx, int ky, int kz, int dipole,
                     double eps_dipole, int slab_axis, double slab_lz)
{
  s->ek_on = 1;
  s->ek_init = 0;
  s->ek_alpha = alpha;
  s->ek_ke = ke;
  s->kx = kx;
  s->ky = ky;
  s->kz = kz;
  s->dipole = dipole;
  s->eps_dipole = eps_dipole;
  s->slab_axis = slab_axis;
  s->slab_lz = slab_lz;
}

/* ---------------------------------------------------------------- domain -- */

/* exchange_owned_single_md_domain.cpp:31-59 with Domain::fix_position
 * (domain.cpp:418-468): one +-L shift per periodic axis, strict comparisons;
 * returns whether any atom was shifted. */
int orc_wrap(orc_sys *s)
{
  int wrapped = 0;
  for (int i = 0; i < s->n; ++i) {
    double *p = s->pos + 3 * i;
    for (int d = 0; d < 3; ++d) {
      if (s->bc[d] != 1)
        continue;
      if (p[d] < s->lo[d]) {
        wrapped = 1;
        p[d] += s->L[d];
      } else if (p[d] > s->hi[d]) {
        wrapped = 1;
        p[d] -= s->L[d];
      }
    }
  }
  return wrapped;
}

static void push_ghost(orc_sys *s, int src, double x, double y, double z)
{
  if (s->g == s->gcap) {
    s->gcap = s->gcap ? 2 * s->gcap : 1024;
    s->gpos = (double *)realloc(s->gpos, sizeof(double) * 3 * s->gcap);
    s->gsrc = (int *)realloc(s->gsrc, sizeof(int) * s->gcap);
  }
  s->gpos[3 * s->g + 0] = x;
  s->gpos[3 * s->g + 1] = y;
  s->gpos[3 * s->g + 2] = z;
  s->gsrc[s->g] = src;
  s->g++;
}

/* exchange_ghost_single_md_domain.cpp:26-215: for owned i ascending, images
 * in the order x, y, z, xy, xz, yz, xyz; image = pos - s*L with s=-1 below
 * lower+ghost_cutoff and s=+1 above upper-ghost_cutoff (strict). */
void orc_ghosts(orc_sys *s)
{
  s->g = 0;
  const double xl = s->lo[0] + s->ghost_cutoff, xu = s->hi[0] - s->ghost_cutoff;
  const double yl = s->lo[1] + s->ghost_cutoff, yu = s->hi[1] - s->ghost_cutoff;
  const double zl = s->lo[2] + s->ghost_cutoff, zu = s->hi[2] - s->ghost_cutoff;
  for (int i = 0; i < s->n; ++i) {
    const double *p = s->pos + 3 * i;
    int xv = p[0] < xl ? -1 : (p[0] > xu ? 1 : 0);
    int yv = p[1] < yl ? -1 : (p[1] > yu ? 1 : 0);
    int zv = p[2] < zl ? -1 : (p[2] > zu ? 1 : 0);
    xv *= s->bc[0];
    yv *= s->bc[1];
    zv *= s->bc[2];
    const double gx = p[0] - xv * s->L[0], gy = p[1] - yv * s->L[1], gz = p[2] - zv * s->L[2];
    if (xv)
      push_ghost(s, i, gx, p[1], p[2]);
    if (yv)
      push_ghost(s, i, p[0], gy, p[2]);
    if (zv)
      push_ghost(s, i, p[0], p[1], gz);
    if (xv && yv)
      push_ghost(s, i, gx, gy, p[2]);
    if (xv && zv)
      push_ghost(s, i, gx, p[1], gz);
    if (yv && zv)
      push_ghost(s, i, p[0], gy, gz);
    if (xv && yv && zv)
      push_ghost(s, i, gx, gy, gz);
  }
}

/* ------------------------------------------------------------ neighbours -- */

static inline double dot3(const double *a) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

static inline const double *apos(const orc_sys *s, int j)
{
  return j < s->n ? s->pos + 3 * j : s->gpos + 3 * (j - s->n);
}

static inline int atype(const orc_sys *s, int j) { return j < s->n ? s->type[j] : s->type[s->gsrc[j - s->n]]; }
static inline int aid(const orc_sys *s, int j) { return j < s->n ? s->id[j] : s->id[s->gsrc[j - s->n]]; }
static inline int amol(const orc_sys *s, int j) { return j < s->n ? s->mol[j] : s->mol[s->gsrc[j - s->n]]; }

/* make_neigh_bin (neighborlist.cpp:581-697): 27 candidates x-outer, z-inner,
 * clamped on non-periodic axes, wrapped on periodic ones, de-duplicated
 * keeping first occurrences. */
static void make_stencil(orc_sys *s)
{
  const int nbx = s->nb[0], nby = s->nb[1], nbz = s->nb[2];
  s->nbins = nbx * nby * nbz;
  int *cnt = (int *)calloc(s->nbins, sizeof(int));
  int *cand = (int *)malloc(sizeof(int) * 27 * s->nbins);
  for (int x = 0; x < nbx; ++x)
    for (int y = 0; y < nby; ++y)
      for (int z = 0; z < nbz; ++z) {
        int lo[3] = {x - 1, y - 1, z - 1}, hi[3] = {x + 1, y + 1, z + 1};
        for (int d = 0; d < 3; ++d)
          if (s->bc[d] == 0) {
            if (lo[d] == -1)
              lo[d] = 0;
            if (hi[d] == s->nb[d])
              hi[d] = s->nb[d] - 1;
          }
        const int k = x + nbx * y + nbx * nby * z;
        int *c = cand + 27 * k;
        for (int a = lo[0]; a <= hi[0]; ++a)
          for (int b = lo[1]; b <= hi[1]; ++b)
            for (int e = lo[2]; e <= hi[2]; ++e) {
              int v[3] = {a, b, e};
              for (int d = 0; d < 3; ++d) {
                if (v[d] == s->nb[d])
                  v[d] = 0;
                if (v[d] == -1)
                  v[d] = s->nb[d] - 1;
              }
              const int lin = v[0] + nbx * v[1] + nbx * nby * v[2];
              int dup = 0;
              for (int q = 0; q < cnt[k]; ++q)
                dup |= c[q] == lin;
              if (!dup)
                c[cnt[k]++] = lin;
            }
      }
  free(s->st_off);
  free(s->st);
  s->st_off = (int *)malloc(sizeof(int) * (s->nbins + 1));
  s->st = (int *)malloc(sizeof(int) * 27 * s->nbins);
  int off = 0;
  for (int k = 0; k < s->nbins; ++k) {
    s->st_off[k] = off;
    memcpy(s->st + off, cand + 27 * k, sizeof(int) * cnt[k]);
    off += cnt[k];
  }
  s->st_off[s->nbins] = off;
  free(cnt);
  free(cand);
}

/* Neighborlist::init (neighborlist.cpp:313-377). */
static void nl_init(orc_sys *s)
{
  s->rce = s->nl_cutoff * s->nl_coef;
  s->thr_sq = 0.25 * (s->rce - s->nl_cutoff) * (s->rce - s->nl_cutoff);
  free(s->pos_old);
  s->pos_old = (double *)calloc(3 * (size_t)(s->n ? s->n : 1), sizeof(double));
  s->n_old = s->n;
  if (!s->build_cells)
    return;
  s->cinv = 1.0 / s->rce;
  for (int d = 0; d < 3; ++d)
    s->nb[d] = (int)ceil(s->L[d] * s->cinv);
  make_stencil(s);
}

/* binlist_index (neighborlist.cpp:408-430): floor, clamped to the grid. */
static void bin_index3(const orc_sys *s, const double *p, int ind[3])
{
  for (int d = 0; d < 3; ++d) {
    int v = (int)floor((p[d] - s->lo[d]) * s->cinv);
    if (v < 0)
      v = 0;
    if (v >= s->nb[d])
      v = s->nb[d] - 1;
    ind[d] = v;
  }
}

static int bin_of(const orc_sys *s, const double *p)
{
  int ind[3];
  bin_index3(s, p, ind);
  return ind[0] + s->nb[0] * ind[1] + s->nb[0] * s->nb[1] * ind[2];
}

/* update_cell_list (neighborlist.cpp:439-472): owned i ascending then ghosts n+g. */
static void update_cell_list(orc_sys *s)
{
  const int tot = s->n + s->g;
  int *bin = (int *)malloc(sizeof(int) * (tot ? tot : 1));
  free(s->bin_off);
  s->bin_off = (int *)calloc(s->nbins + 1, sizeof(int));
  for (int j = 0; j < tot; ++j) {
    bin[j] = bin_of(s, apos(s, j));
    s->bin_off[bin[j] + 1]++;
  }
  for (int k = 0; k < s->nbins; ++k)
    s->bin_off[k + 1] += s->bin_off[k];
  free(s->bin_items);
  s->bin_items = (int *)malloc(sizeof(int) * (tot ? tot : 1));
  int *cur = (int *)malloc(sizeof(int) * (s->nbins ? s->nbins : 1));
  memcpy(cur, s->bin_off, sizeof(int) * s->nbins);
  for (int j = 0; j < tot; ++j)
    s->bin_items[cur[bin[j]]++] = j;
  free(cur);
  free(bin);
}

/* update_is_needed (neighborlist.cpp:91-139). */
static int update_is_needed(const orc_sys *s)
{
  if (s->rebuild_test)
    return 1;
  if (s->n != s->n_old || s->g != s->g_old)
    return 1;
  for (int i = 0; i < s->n; ++i) {
    double d[3] = {s->pos[3 * i] - s->pos_old[3 * i], s->pos[3 * i + 1] - s->pos_old[3 * i + 1],
                   s->pos[3 * i + 2] - s->pos_old[3 * i + 2]};
    if (dot3(d) > s->thr_sq)
      return 1;
  }
  for (int i = 0; i < s->g; ++i) {
    double d[3] = {s->gpos[3 * i] - s->gpos_old[3 * i], s->gpos[3 * i + 1] - s->gpos_old[3 * i + 1],
                   s->gpos[3 * i + 2] - s->gpos_old[3 * i + 2]};
    if (dot3(d) > s->thr_sq)
      return 1;
  }
  return 0;
}

static void nl_push(orc_sys *s, long *w, int j)
{
  if (*w == s->nl_cap) {
    s->nl_cap = s->nl_cap ? 2 * s->nl_cap : 4096;
    s->nl_items = (int *)realloc(s->nl_items, sizeof(int) * s->nl_cap);
  }
  s->nl_items[(*w)++] = j;
}

static void save_old(orc_sys *s)
{
  free(s->pos_old);
  free(s->gpos_old);
  s->pos_old = (double *)dup_arr(s->pos, sizeof(double) * 3 * s->n);
  s->gpos_old = (double *)dup_arr(s->gpos, sizeof(double) * 3 * s->g);
  s->n_old = s->n;
  s->g_old = s->g;
}

/* all_atom_test_function (neighborlist.cpp:161-191): every owned j>i and every
 * ghost, no distance test; pos_old is not refreshed (the callers return early). */
static void all_atom_list(orc_sys *s)
{
  free(s->nl_off);
  s->nl_off = (long *)malloc(sizeof(long) * (s->n + 1));
  long w = 0;
  for (int i = 0; i < s->n; ++i) {
    s->nl_off[i] = w;
    for (int j = i + 1; j < s->n; ++j)
      nl_push(s, &w, j);
    for (int g = 0; g < s->g; ++g)
      nl_push(s, &w, g + s->n);
  }
  s->nl_off[s->n] = w;
}

/* update_verlet_list (neighborlist.cpp:257-307): brute force, j ascending,
 * strict dr*dr < cutoff_extra^2 with dr = pos_j - pos_i. */
static void verlet_brute(orc_sys *s)
{
  if (s->all_atom_test) {
    all_atom_list(s);
    return;
  }
  const double rc2 = s->rce * s->rce;
  free(s->nl_off);
  s->nl_off = (long *)malloc(sizeof(long) * (s->n + 1));
  long w = 0;
  for (int i = 0; i < s->n; ++i) {
    s->nl_off[i] = w;
    const double *pi = s->pos + 3 * i;
    for (int j = i + 1; j < s->n; ++j) {
      const double *pj = s->pos + 3 * j;
      double d[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
      if (dot3(d) < rc2)
        nl_push(s, &w, j);
    }
    for (int g = 0; g < s->g; ++g) {
      const double *pj = s->gpos + 3 * g;
      double d[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
      if (dot3(d) < rc2)
        nl_push(s, &w, g + s->n);
    }
  }
  s->nl_off[s->n] = w;
  save_old(s);
}

/* update_verlet_list_from_cell_list (neighborlist.cpp:515-573): stencil order,
 * owned j must be > i, dr = pos_i - pos_j, strict test. */
static void verlet_cells(orc_sys *s)
{
  if (s->all_atom_test) {
    all_atom_list(s);
    return;
  }
  const double rc2 = s->rce * s->rce;
  free(s->nl_off);
  s->nl_off = (long *)malloc(sizeof(long) * (s->n + 1));
  long w = 0;
  for (int i = 0; i < s->n; ++i) {
    s->nl_off[i] = w;
    const double *pi = s->pos + 3 * i;
    const int b = bin_of(s, pi);
    for (int q = s->st_off[b]; q < s->st_off[b + 1]; ++q) {
      const int c = s->st[q];
      for (int t = s->bin_off[c]; t < s->bin_off[c + 1]; ++t) {
        const int j = s->bin_items[t];
        if (j < s->n && j <= i)
          continue;
        const double *pj = apos(s, j);
        double d[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
        if (dot3(d) < rc2)
          nl_push(s, &w, j);
      }
    }
  }
  s->nl_off[s->n] = w;
  save_old(s);
}

/* Neighborlist::build (neighborlist.cpp:66-88). */
int orc_build(orc_sys *s, int update)
{
  if (!s->nl_on)
    return 0;
  if (s->build_cells)
    update_cell_list(s);
  if (!update)
    update = update_is_needed(s);
  if (update) {
    if (s->from_cells)
      verlet_cells(s);
    else
      verlet_brute(s);
    return 1;
  }
  return 0;
}

/* ----------------------------------------------------------- force fields -- */

void orc_reset_acc(orc_sys *s) { memset(s->acc, 0, sizeof(double) * 3 * s->n); } /* atom_data.cpp:308-316 */

/* Lj_cell_list::calculate_acceleration (lj_cell_list.cpp:85-175): 27-cell walk,
 * owned i<j and every ghost, skip dr^2 > rc^2, shifted force coefficient,
 * Newton-3 on owned partners only.  Contributes nothing to the virial. */
void orc_lj_acc(orc_sys *s)
{
  if (!s->build_cells)
    return; /* Lj_cell_list walks the cell list only */
  const double rc = s->lj_rc, rc2 = rc * rc;
  for (int i = 0; i < s->n; ++i) {
    const double *pi = s->pos + 3 * i;
    const int ti = s->type[i];
    const double mi = s->minv[ti];
    const int b = bin_of(s, pi);
    for (int q = s->st_off[b]; q < s->st_off[b + 1]; ++q) {
      const int c = s->st[q];
      for (int t = s->bin_off[c]; t < s->bin_off[c + 1]; ++t) {
        const int j = s->bin_items[t];
        const int ghost = j >= s->n;
        if (!ghost && !(i < j))
          continue;
        const double *pj = apos(s, j);
        const int tj = atype(s, j);
        const double mj = s->minv[tj];
        double dr[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
        const double dr2 = dot3(dr);
        if (dr2 > rc2)
          continue;
        const double e = s->eps[ti * s->nt + tj], sg = s->sig[ti * s->nt + tj];
        const double rcsi = 1 / (rc * rc);
        const double rhoc2 = sg * sg * rcsi;
        const double rhoc6 = rhoc2 * rhoc2 * rhoc2;
        const double rhoc12 = rhoc6 * rhoc6;
        const double dsi = 1 / dr2;
        const double rho2 = sg * sg * dsi;
        const double rho6 = rho2 * rho2 * rho2;
        const double rho12 = rho6 * rho6;
        const double coef = 4 * e *
                            (-12 * rho12 * dsi + 6 * rho6 * dsi + +12 * rhoc12 * rcsi -
                             6 * rhoc6 * rcsi);
        double f[3] = {coef * dr[0], coef * dr[1], coef * dr[2]};
        for (int d = 0; d < 3; ++d)
          s->acc[3 * i + d] += f[d] * mi;
        if (!ghost)
          for (int d = 0; d < 3; ++d)
            s->acc[3 * j + d] -= f[d] * mj;
      }
    }
  }
}

/* Lj::calculate_acceleration (lj.cpp:341-483): every Verlet-list entry;
 * same-molecule pairs skipped with ignore_intra_molecule; c = cutoff^2 or
 * cutoff_list[ti][tj]^2; skip dr^2 > c; jump_fix clamps dr^2 below jump_tol;
 * eps_ij = 0 skipped; the force shift at c; virial -coef*dr^2 when
 * id_i < id_j; Newton-3 on owned partners. */
void orc_lj_list_acc(orc_sys *s)
{
  const int nt = s->nt;
  const double cutoff_sq = s->ljl_rc * s->ljl_rc;
  double vir = 0;
  for (int i = 0; i < s->n; ++i) {
    const double *pi = s->pos + 3 * i;
    const int ti = s->type[i];
    const double mi = s->minv[ti];
    const int idi = s->id[i], moli = s->mol[i];
    for (long t = s->nl_off[i]; t < s->nl_off[i + 1]; ++t) {
      const int j = s->nl_items[t];
      const int ghost = j >= s->n;
      const double *pj = apos(s, j);
      const int tj = atype(s, j);
      if (s->ljl_ignore_intra && moli == amol(s, j))
        continue;
      double dr[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
      double dr2 = dot3(dr);
      double c = cutoff_sq;
      if (s->ljl_cut_on)
        c = s->ljl_cut[ti * nt + tj] * s->ljl_cut[ti * nt + tj];
      if (dr2 > c)
        continue;
      if (s->ljl_jump && dr2 < s->ljl_jump_tol)
        dr2 = s->ljl_jump_tol;
      const double e = s->ljl_eps[ti * nt + tj];
      if (e == 0)
        continue;
      const double sg = s->ljl_sig[ti * nt + tj];
      const double rcsi = 1 / c;
      const double rhoc2 = sg * sg * rcsi;
      const double rhoc6 = rhoc2 * rhoc2 * rhoc2;
      const double rhoc12 = rhoc6 * rhoc6;
      const double dsi = 1 / dr2;
      const double rho2 = sg * sg * dsi;
      const double rho6 = rho2 * rho2 * rho2;
      const double rho12 = rho6 * rho6;
      const double coef = 4 * e * (-12 * rho12 * dsi + 6 * rho6 * dsi + 12 * rhoc12 * rcsi -
                                   6 * rhoc6 * rcsi);
      if (idi < aid(s, j))
        vir += -coef * dr2;
      double f[3] = {coef * dr[0], coef * dr[1], coef * dr[2]};
      for (int d = 0; d < 3; ++d)
        s->acc[3 * i + d] += f[d] * mi;
      if (!ghost) {
        const double mj = s->minv[tj];
        for (int d = 0; d < 3; ++d)
          s->acc[3 * j + d] -= f[d] * mj;
      }
    }
  }
  s->virial += vir;
}

/* Electrostatic_short_range::calculate_acceleration (electrostatic_short_range.cpp:
 * 133-258): every Verlet-list entry; lambda of same-molecule pairs (0 skips);
 * skip dr^2 >= cutoff^2; coef = -lambda ke qi qj / (r^2 r) (1 - (r/rc)^(beta+1));
 * virial -coef*dr^2 when id_i < id_j; then the external field q_i E per atom. */
void orc_short_range_acc(orc_sys *s)
{
  const int nt = s->nt;
  const double cutoff = s->sr_rc, cutoff_sq = cutoff * cutoff, ke = s->sr_ke;
  double vir = 0;
  for (int i = 0; i < s->n; ++i) {
    const double *pi = s->pos + 3 * i;
    const int ti = s->type[i];
    const double qi = s->charge[ti], mi = s->minv[ti];
    const int idi = s->id[i], moli = s->mol[i];
    for (long t = s->nl_off[i]; t < s->nl_off[i + 1]; ++t) {
      const int j = s->nl_items[t];
      const int ghost = j >= s->n;
      const double *pj = apos(s, j);
      const int tj = atype(s, j);
      double lam = 1.0;
      if (s->sr_lambda_set && moli == amol(s, j)) {
        if (s->sr_lambda[ti * nt + tj] == 0.0)
          continue;
        lam *= s->sr_lambda[ti * nt + tj];
      }
      double dr[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
      const double dr2 = dot3(dr);
      if (dr2 >= cutoff_sq)
        continue;
      const double qj = s->charge[tj];
      const double rn = sqrt(dr2);
      const double coef = -lam * ke * qi * qj / (dr2 * rn) * (1.0 - pow(rn / cutoff, s->sr_beta + 1));
      if (idi < aid(s, j))
        vir += -coef * dr2;
      double f[3] = {coef * dr[0], coef * dr[1], coef * dr[2]};
      for (int d = 0; d < 3; ++d)
        s->acc[3 * i + d] += f[d] * mi;
      if (!ghost) {
        const double mj = s->minv[tj];
        for (int d = 0; d < 3; ++d)
          s->acc[3 * j + d] -= f[d] * mj;
      }
    }
    for (int d = 0; d < 3; ++d)
      s->acc[3 * i + d] += s->sr_ext[d] * qi * mi;
  }
  s->virial += vir;
}

/* ---------------------------------------------------- slab correction (ELC) -- */

/* give_slab_local_coordinates (electrostatic_ewald_slab_correction.h:55-77). */
static void elc_local(const orc_sys *s, const double *g, double *l)
{
  if (s->elc_axis == 0) {
    l[0] = g[1];
    l[1] = g[2];
    l[2] = g[0];
  } else if (s->elc_axis == 1) {
    l[0] = g[2];
    l[1] = g[0];
    l[2] = g[1];
  } else {
    l[0] = g[0];
    l[1] = g[1];
    l[2] = g[2];
  }
}

/* give_slab_global_coordinates (electrostatic_ewald_slab_correction.h:79-101),
 * as the reference maps it: for axis x it applies axis y's inverse and for
 * axis y axis x's (the two are not each other's local maps' inverses). */
static void elc_global(const orc_sys *s, const double *l, double *g)
{
  if (s->elc_axis == 0) {
    g[0] = l[1];
    g[1] = l[2];
    g[2] = l[0];
  } else if (s->elc_axis == 1) {
    g[0] = l[2];
    g[1] = l[0];
    g[2] = l[1];
  } else {
    g[0] = l[0];
    g[1] = l[1];
    g[2] = l[2];
  }
}

/* initialize (electrostatic_ewald_slab_correction/initialize.cpp:31-228): box
 * lengths in slab coordinates and the k grids once; then per call the dipole
 * sum and the chi sums.  The dipole field of a call comes from the previous
 * call's dipole sum (initialize.cpp:86: dipole_sum is read before it is
 * updated), zero on the first call. */
static void elc_initialize(orc_sys *s)
{
  const int kx = s->elc_kx, ky = s->elc_ky;
  if (!s->elc_init) {
    s->elc_init = 1;
    const double glx = s->L[0], gly = s->L[1], glz = s->L[2];
    if (s->elc_axis == 0) {
      s->elc_lx = gly;
      s->elc_ly = glz;
      s->elc_hz = glx;
    } else if (s->elc_axis == 1) {
      s->elc_lx = glx;
      s->elc_ly = glz;
      s->elc_hz = gly;
    } else {
      s->elc_lx = glx;
      s->elc_ly = gly;
      s->elc_hz = glz;
    }
    const double lx_ly_inv = 1.0 / (s->elc_lx * s->elc_ly);
    s->elc_sum_e_coef = -ORC_4PI * lx_ly_inv;
    /* make_slab_k_vectors (initialize.cpp:119-160) */
    free(s->elc_kxv);
    free(s->elc_kyv);
    free(s->elc_kpv);
    free(s->elc_kxc);
    free(s->elc_kyc);
    free(s->elc_kpc);
    s->elc_kxv = (double *)calloc(kx + 1, sizeof(double));
    s->elc_kxc = (double *)calloc(kx + 1, sizeof(double));
    s->elc_kyv = (double *)calloc(ky + 1, sizeof(double));
    s->elc_kyc = (double *)calloc(ky + 1, sizeof(double));
    s->elc_kpv = (double *)calloc(kx * ky + 1, sizeof(double));
    s->elc_kpc = (double *)calloc(kx * ky + 1, sizeof(double));
    const double lx_inv = 1.0 / s->elc_lx, ly_inv = 1.0 / s->elc_ly, lz = s->elc_lz;
    for (int ix = 0; ix < kx; ++ix) {
      s->elc_kxv[ix] = ORC_2PI * (ix + 1) * lx_inv;
      s->elc_kxc[ix] = 1.0 / (s->elc_kxv[ix] * (exp(s->elc_kxv[ix] * lz) - 1.0));
    }
    for (int iy = 0; iy < ky; ++iy) {
      s->elc_kyv[iy] = ORC_2PI * (iy + 1) * ly_inv;
      s->elc_kyc[iy] = 1.0 / (s->elc_kyv[iy] * (exp(s->elc_kyv[iy] * lz) - 1.0));
    }
    int ip = 0;
    for (int ix = 0; ix < kx; ++ix)
      for (int iy = 0; iy < ky; ++iy) {
        s->elc_kpv[ip] = sqrt((s->elc_kxv[ix] * s->elc_kxv[ix]) + (s->elc_kyv[iy] * s->elc_kyv[iy]));
        s->elc_kpc[ip] = 1.0 / (s->elc_kpv[ip] * (exp(s->elc_kpv[ip] * lz) - 1.0));
        ++ip;
      }
    const double l_xyz_inv = lx_ly_inv / lz;
    s->elc_dip_coef = s->elc_ke * ORC_4PI * l_xyz_inv;
    free(s->elc_chi_x);
    free(s->elc_chi_y);
    free(s->elc_chi_p);
    s->elc_chi_x = (double *)calloc(4 * kx + 1, sizeof(double));
    s->elc_chi_y = (double *)calloc(4 * ky + 1, sizeof(double));
    s->elc_chi_p = (double *)calloc(8 * kx * ky + 1, sizeof(double));
  }
  /* calculate_dipole_sum (initialize.cpp:75-117) */
  double ds[3] = {0, 0, 0};
  for (int j = 0; j < s->n; ++j) {
    const double q = s->charge[s->type[j]];
    for (int d = 0; d < 3; ++d)
      ds[d] += q * s->pos[3 * j + d];
  }
  double dfv[3];
  for (int d = 0; d < 3; ++d)
    dfv[d] = -s->elc_dip_sum[d] * s->elc_dip_coef;
  for (int d = 0; d < 3; ++d) {
    const int on = d == s->elc_axis;
    s->elc_dip_field[d] = on ? dfv[d] : 0.0;
    s->elc_dip_sum[d] = on ? ds[d] : 0.0;
  }
  /* make_slab_chi_vectors (initialize.cpp:162-268) */
  memset(s->elc_chi_x, 0, sizeof(double) * 4 * kx);
  memset(s->elc_chi_y, 0, sizeof(double) * 4 * ky);
  memset(s->elc_chi_p, 0, sizeof(double) * 8 * kx * ky);
  for (int j = 0; j < s->n; ++j) {
    const double q = s->charge[s->type[j]];
    double p[3];
    elc_local(s, s->pos + 3 * j, p);
    int ip = 0;
    for (int ix = 0; ix < kx; ++ix) {
      const double kxv = s->elc_kxv[ix];
      const double sinhkxz = sinh(kxv * p[2]), coshkxz = cosh(kxv * p[2]);
      const double sinkxx = sin(kxv * p[0]), coskxx = cos(kxv * p[0]);
      double *cx = s->elc_chi_x + 4 * ix;
      cx[0] += q * sinhkxz * sinkxx;
      cx[1] += q * sinhkxz * coskxx;
      cx[2] += q * coshkxz * sinkxx;
      cx[3] += q * coshkxz * coskxx;
      for (int iy = 0; iy < ky; ++iy) {
        const double kyv = s->elc_kyv[iy], kpv = s->elc_kpv[ip];
        const double sinhkpz = sinh(kpv * p[2]), coshkpz = cosh(kpv * p[2]);
        const double sinkyy = sin(kyv * p[1]), coskyy = cos(kyv * p[1]);
        double *cp = s->elc_chi_p + 8 * ip; /* [a][b][c] at 4a + 2b + c */
        cp[0] += q * sinhkpz * sinkxx * sinkyy;
        cp[1] += q * sinhkpz * sinkxx * coskyy;
        cp[2] += q * sinhkpz * coskxx * sinkyy;
        cp[4] += q * coshkpz * sinkxx * sinkyy;
        cp[3] += q * sinhkpz * coskxx * coskyy;
        cp[5] += q * coshkpz * sinkxx * coskyy;
        cp[6] += q * coshkpz * coskxx * sinkyy;
        cp[7] += q * coshkpz * coskxx * coskyy;
        if (ix == 0) {
          const double sinhkyz = sinh(kyv * p[2]), coshkyz = cosh(kyv * p[2]);
          double *cy = s->elc_chi_y + 4 * iy;
          cy[0] += q * sinhkyz * sinkyy;
          cy[1] += q * sinhkyz * coskyy;
          cy[2] += q * coshkyz * sinkyy;
          cy[3] += q * coshkyz * coskyy;
        }
        ++ip;
      }
    }
  }
}

/* Electrostatic_ewald_slab_correction::field (electrostatic_ewald_slab_correction/
 * field.cpp:29-136): the order-N scheme over the chi sums, in slab coordinates,
 * plus the dipole field. */
static void elc_field(const orc_sys *s, const double *r, double *out)
{
  const int kx = s->elc_kx, ky = s->elc_ky;
  double f[3] = {0, 0, 0};
  double p[3];
  elc_local(s, r, p);
  int ip = 0;
  for (int ix = 0; ix < kx; ++ix) {
    const double kxv = s->elc_kxv[ix];
    const double sinhkxz = sinh(kxv * p[2]), coshkxz = cosh(kxv * p[2]);
    const double sinkxx = sin(kxv * p[0]), coskxx = cos(kxv * p[0]);
    const double *c = s->elc_chi_x + 4 * ix;
    const double sx_chi = -kxv * (-(+c[0] * coskxx * sinhkxz) - (-c[1] * sinkxx * sinhkxz) +
                                  (+c[2] * coskxx * coshkxz) + (-c[3] * sinkxx * coshkxz));
    const double sz_chi = -kxv * (-(+c[0] * coshkxz * sinkxx) - (+c[1] * coshkxz * coskxx) +
                                  (+c[2] * sinhkxz * sinkxx) + (+c[3] * sinhkxz * coskxx));
    f[0] += s->elc_kxc[ix] * sx_chi;
    f[2] += s->elc_kxc[ix] * sz_chi;
    for (int iy = 0; iy < ky; ++iy) {
      const double kyv = s->elc_kyv[iy], kpv = s->elc_kpv[ip];
      const double sinhkpz = sinh(kpv * p[2]), coshkpz = cosh(kpv * p[2]);
      const double sinkyy = sin(kyv * p[1]), coskyy = cos(kyv * p[1]);
      const double *q = s->elc_chi_p + 8 * ip; /* 000 001 010 011 100 101 110 111 */
      const double sx = -kxv * (-(+q[0] * coskxx * sinhkpz * sinkyy) - (+q[1] * coskxx * sinhkpz * coskyy) -
                                (-q[2] * sinkxx * sinhkpz * sinkyy) + (+q[4] * coskxx * coshkpz * sinkyy) -
                                (-q[3] * sinkxx * sinhkpz * coskyy) + (+q[5] * coskxx * coshkpz * coskyy) +
                                (-q[6] * sinkxx * coshkpz * sinkyy) + (-q[7] * sinkxx * coshkpz * coskyy));
      const double sy = -kyv * (-(+q[0] * coskyy * sinhkpz * sinkxx) - (-q[1] * sinkyy * sinhkpz * sinkxx) -
                                (+q[2] * coskyy * sinhkpz * coskxx) + (+q[4] * coskyy * coshkpz * sinkxx) -
                                (-q[3] * sinkyy * sinhkpz * coskxx) + (-q[5] * sinkyy * coshkpz * sinkxx) +
                                (+q[6] * coskyy * coshkpz * coskxx) + (-q[7] * sinkyy * coshkpz * coskxx));
      const double sz = -kpv * (-(+q[0] * coshkpz * sinkxx * sinkyy) - (+q[1] * coshkpz * sinkxx * coskyy) -
                                (+q[2] * coshkpz * coskxx * sinkyy) + (+q[4] * sinhkpz * sinkxx * sinkyy) -
                                (+q[3] * coshkpz * coskxx * coskyy) + (+q[5] * sinhkpz * sinkxx * coskyy) +
                                (+q[6] * sinhkpz * coskxx * sinkyy) + (+q[7] * sinhkpz * coskxx * coskyy));
      f[0] += 2.0 * s->elc_kpc[ip] * sx;
      f[1] += 2.0 * s->elc_kpc[ip] * sy;
      f[2] += 2.0 * s->elc_kpc[ip] * sz;
      if (ix == 0) {
        const double sinhkyz = sinh(kyv * p[2]), coshkyz = cosh(kyv * p[2]);
        const double *cy = s->elc_chi_y + 4 * iy;
        const double syy = -kyv * (-(+cy[0] * coskyy * sinhkyz) - (-cy[1] * sinkyy * sinhkyz) +
                                   (+cy[2] * coskyy * coshkyz) + (-cy[3] * sinkyy * coshkyz));
        const double syz = -kyv * (-(+cy[0] * sinkyy * coshkyz) - (+cy[1] * coskyy * coshkyz) +
                                   (+cy[2] * sinkyy * sinhkyz) + (+cy[3] * coskyy * sinhkyz));
        f[1] += s->elc_kyc[iy] * syy;
        f[2] += s->elc_kyc[iy] * syz;
      }
      ++ip;
    }
  }
  const double sc = 2.0 * s->elc_sum_e_coef * s->elc_ke;
  for (int d = 0; d < 3; ++d)
    f[d] *= sc;
  double g[3];
  elc_global(s, f, g);
  for (int d = 0; d < 3; ++d)
    out[d] = g[d] + s->elc_dip_field[d];
}

/* Electrostatic_ewald_slab_correction::calculate_acceleration
 * (electrostatic_ewald_slab_correction.cpp:134-181): initialize, then
 * force = q_i field(r_i) per owned atom. */
void orc_elc_acc(orc_sys *s)
{
  elc_initialize(s);
  for (int i = 0; i < s->n; ++i) {
    const int ti = s->type[i];
    const double qi = s->charge[ti], mi = s->minv[ti];
    double e[3];
    elc_field(s, s->pos + 3 * i, e);
    for (int d = 0; d < 3; ++d) {
      const double force = qi * e[d];
      s->acc[3 * i + d] += force * mi;
    }
  }
}

/* energy (electrostatic_ewald_slab_correction/energy.cpp:32-138): the chi sums
 * of the last force evaluation, plus 0.5 dipole_coef |dipole_sum|^2. */
double orc_energy_elc(orc_sys *s)
{
  const int kx = s->elc_kx, ky = s->elc_ky;
  double sum_e = 0.0;
  for (int ip = 0; ip < kx * ky; ++ip) {
    const double *q = s->elc_chi_p + 8 * ip;
    const double sum_chi = -(q[0] * q[0]) - (q[1] * q[1]) - (q[2] * q[2]) + (q[4] * q[4]) - (q[3] * q[3]) +
                           (q[5] * q[5]) + (q[6] * q[6]) + (q[7] * q[7]);
    sum_e += 2.0 * s->elc_kpc[ip] * sum_chi;
  }
  for (int ix = 0; ix < kx; ++ix) {
    const double *c = s->elc_chi_x + 4 * ix;
    const double sum_chi = -(c[0] * c[0]) - (c[1] * c[1]) + (c[2] * c[2]) + (c[3] * c[3]);
    sum_e += s->elc_kxc[ix] * sum_chi;
  }
  for (int iy = 0; iy < ky; ++iy) {
    const double *c = s->elc_chi_y + 4 * iy;
    const double sum_chi = -(c[0] * c[0]) - (c[1] * c[1]) + (c[2] * c[2]) + (c[3] * c[3]);
    sum_e += s->elc_kyc[iy] * sum_chi;
  }
  sum_e *= s->elc_ke * s->elc_sum_e_coef;
  const double *d = s->elc_dip_sum;
  return sum_e + 0.5 * s->elc_dip_coef * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

/* potential(const Vector&) (electrostatic_ewald_slab_correction/potential.cpp:
 * 30-79) and field(const Vector&) at np points (state of the last force
 * evaluation). */
void orc_point_elc(orc_sys *s, int np, const double *pts, double *field, double *pot)
{
  const int kx = s->elc_kx, ky = s->elc_ky;
  for (int k = 0; k < np; ++k) {
    const double *r = pts + 3 * k;
    if (field)
      elc_field(s, r, field + 3 * k);
    if (!pot)
      continue;
    double sum_p = 0;
    for (int j = 0; j < s->n; ++j) {
      double dg[3] = {r[0] - s->pos[3 * j], r[1] - s->pos[3 * j + 1], r[2] - s->pos[3 * j + 2]};
      double p[3];
      elc_local(s, dg, p);
      const double qj = s->charge[s->type[j]];
      double sum_kx = 0, sum_ky = 0, sum_kp = 0;
      int ip = 0;
      for (int ix = 0; ix < kx; ++ix) {
        const double coshkxz = cosh(s->elc_kxv[ix] * p[2]);
        const double coskxx = cos(s->elc_kxv[ix] * p[0]);
        sum_kx += s->elc_kxc[ix] * coshkxz * coskxx;
        for (int iy = 0; iy < ky; ++iy) {
          const double coshkpz = cosh(s->elc_kpv[ip] * p[2]);
          const double coskyy = cos(s->elc_kyv[iy] * p[1]);
          sum_kp += s->elc_kpc[ip] * coshkpz * coskxx * coskyy;
          if (ix == 0)
            sum_ky += s->elc_kyc[iy] * cosh(s->elc_kyv[iy] * p[2]) * coskyy;
          ++ip;
        }
      }
      sum_p += qj * ((2.0 * sum_kp) + sum_kx + sum_ky);
    }
    pot[k] = sum_p * s->elc_sum_e_coef * s->elc_ke;
  }
}

/* Electrostatic_ewald_r::calculate_acceleration (electrostatic_ewald_r.cpp:115-227):
 * every Verlet-list entry (no cutoff test), exact-zero separation skipped,
 * virial -coef*r^2 when id_i < id_j. */
void orc_ewald_r_acc(orc_sys *s)
{
  const double alpha = s->er_alpha, a2 = alpha * alpha, ke = s->er_ke;
  double vir = 0;
  for (int i = 0; i < s->n; ++i) {
    const double *pi = s->pos + 3 * i;
    const int ti = s->type[i];
    const double qi = s->charge[ti], mi = s->minv[ti];
    const int idi = s->id[i], moli = s->mol[i];
    for (long t = s->nl_off[i]; t < s->nl_off[i + 1]; ++t) {
      const int j = s->nl_items[t];
      const int ghost = j >= s->n;
      const double *pj = apos(s, j);
      const int tj = atype(s, j);
      double lam = 1.0;
      if (s->lambda_set && moli == amol(s, j)) {
        if (s->lambda[ti * s->nt + tj] == 0.0)
          continue;
        lam *= s->lambda[ti * s->nt + tj];
      }
      double r[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
      if (r[0] == 0 && r[1] == 0 && r[2] == 0)
        continue;
      const double qj = s->charge[tj];
      const double r2 = dot3(r);
      const double rn = sqrt(r2);
      const double sum = (2 * alpha * ORC_PIS_INV * exp(-a2 * r2) + erfc(alpha * rn) / rn) * (1.0 / r2);
      const double coef = -lam * ke * qi * qj * sum;
      if (idi < aid(s, j))
        vir += -coef * r2;
      double f[3] = {coef * r[0], coef * r[1], coef * r[2]};
      for (int d = 0; d < 3; ++d)
        s->acc[3 * i + d] += f[d] * mi;
      if (!ghost) {
        const double mj = s->minv[tj];
        for (int d = 0; d < 3; ++d)
          s->acc[3 * j + d] -= f[d] * mj;
      }
    }
  }
  s->virial += vir;
}

/* Electrostatic_ewald_k::initialize / make_k_vectors (electrostatic_ewald_k/
 * initialize.cpp:31-92,184-222): lexicographic kx,ky,kz in [-kmax,kmax] without
 * 0, but only the first (2kx)(2ky)(2kz) vectors are used by every consumer. */
static void ek_init(orc_sys *s)
{
  double lx = s->L[0], ly = s->L[1], lz = s->L[2];
  if (s->slab_axis == 0)
    lx = s->slab_lz;
  if (s->slab_axis == 1)
    ly = s->slab_lz;
  if (s->slab_axis == 2)
    lz = s->slab_lz;
  const double lxi = 1.0 / lx, lyi = 1.0 / ly, lzi = 1.0 / lz;
  s->lxyz_inv = 1.0 / (lx * ly * lz);
  s->nk = (2 * s->kx) * (2 * s->ky) * (2 * s->kz);
  const int nfull = (2 * s->kx + 1) * (2 * s->ky + 1) * (2 * s->kz + 1) - 1;
  free(s->kvec);
  free(s->fk);
  free(s->sk);
  s->kvec = (double *)malloc(sizeof(double) * 3 * (nfull ? nfull : 1));
  s->fk = (double *)malloc(sizeof(double) * (nfull ? nfull : 1));
  s->sk = (double *)calloc(2 * (size_t)(s->nk ? s->nk : 1), sizeof(double));
  const double a2 = s->ek_alpha * s->ek_alpha, f4a2 = 1.0 / (4.0 * a2);
  int w = 0;
  for (int kx = -s->kx; kx <= s->kx; ++kx)
    for (int ky = -s->ky; ky <= s->ky; ++ky)
      for (int kz = -s->kz; kz <= s->kz; ++kz) {
        double k[3] = {ORC_2PI * kx * lxi, ORC_2PI * ky * lyi, ORC_2PI * kz * lzi};
        const double k2 = dot3(k);
        if (k2 == 0)
          continue;
        const double k2i = 1.0 / k2;
        memcpy(s->kvec + 3 * w, k, sizeof k);
        s->fk[w] = k2i * exp(-k2 * f4a2);
        ++w;
      }
  if (s->dipole)
    s->dip_coef = s->ek_ke * ORC_4PI * s->lxyz_inv / (1 + 2.0 * s->eps_dipole);
  s->ek_init = 1;
}

/* calculate_dipole_sum + calculate_potential_k_coef_cmplx
 * (electrostatic_ewald_k/initialize.cpp:94-146). */
static void ek_structure_factor(orc_sys *s, int k0, int k1)
{
  if (s->dipole) {
    s->dip_sum[0] = s->dip_sum[1] = s->dip_sum[2] = 0;
    for (int j = 0; j < s->n; ++j) {
      const double q = s->charge[s->type[j]];
      for (int d = 0; d < 3; ++d)
        s->dip_sum[d] += q * s->pos[3 * j + d];
    }
    for (int d = 0; d < 3; ++d)
      s->dip_field[d] = -s->dip_sum[d] * s->dip_coef;
  }
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int k = k0; k < k1; ++k) {
    const double *kv = s->kvec + 3 * k;
    double complex rho = 0;
    for (int j = 0; j < s->n; ++j) {
      const double q = s->charge[s->type[j]];
      const double *p = s->pos + 3 * j;
      const double th = kv[0] * p[0] + kv[1] * p[1] + kv[2] * p[2];
      rho += q * (cos(th) + I * sin(th));
    }
    s->sk[2 * k] = creal(rho);
    s->sk[2 * k + 1] = cimag(rho);
  }
}

/* Electrostatic_ewald_k::calculate_acceleration (electrostatic_ewald_k.cpp:161-234).
 * The reference loops k outer, i inner; per atom the k terms are added in the
 * same k order here.  With epsilon_dipole the dipole force is added once per
 * k-vector, as the reference's loop does (electrostatic_ewald_k.cpp:216-219). */
static void ek_force_pass_idx(orc_sys *s, int i0, int i1, const int *idx);
static void ek_force_pass(orc_sys *s, int i0, int i1) { ek_force_pass_idx(s, i0, i1, NULL); }

/* The force pass over atoms idx[i0..i1) (idx NULL: atoms i0..i1). */
static void ek_force_pass_idx(orc_sys *s, int i0, int i1, const int *idx)
{
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int ii = i0; ii < i1; ++ii) {
    const int i = idx ? idx[ii] : ii;
    const int ti = s->type[i];
    const double qi = s->charge[ti], mi = s->minv[ti];
    const double *p = s->pos + 3 * i;
    double a[3] = {s->acc[3 * i], s->acc[3 * i + 1], s->acc[3 * i + 2]};
    for (int k = 0; k < s->nk; ++k) {
      const double *kv = s->kvec + 3 * k;
      const double sr = s->sk[2 * k], si = -s->sk[2 * k + 1];
      const double th = kv[0] * p[0] + kv[1] * p[1] + kv[2] * p[2];
      const double c = cos(th), sn = sin(th);
      const double im = sr * sn + si * c;
      const double w = ORC_4PI * s->fk[k];
      double sk3[3] = {w * kv[0] * im, w * kv[1] * im, w * kv[2] * im};
      const double cc = s->ek_ke * qi * s->lxyz_inv;
      for (int d = 0; d < 3; ++d) {
        double f = cc * sk3[d];
        if (s->dipole)
          f += qi * s->dip_field[d];
        a[d] += f * mi;
      }
    }
    s->acc[3 * i] = a[0];
    s->acc[3 * i + 1] = a[1];
    s->acc[3 * i + 2] = a[2];
  }
}

void orc_ewald_k_acc(orc_sys *s)
{
  if (!s->ek_init)
    ek_init(s);
  ek_structure_factor(s, 0, s->nk);
  ek_force_pass(s, 0, s->n);
}

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

void orc_ewald_k_sample(orc_sys *s, int k0, int k1, int i0, int i1, double *t_s, double *t_f)
{
  if (!s->ek_init)
    ek_init(s);
  double t0 = now_s();
  ek_structure_factor(s, k0, k1);
  double t1 = now_s();
  ek_force_pass(s, i0, i1);
  double t2 = now_s();
  *t_s = t1 - t0;
  *t_f = t2 - t1;
}

void orc_forces_sampled(orc_sys *s, int m, const int *idx)
{
  orc_reset_acc(s);
  if (s->lj_on)
    orc_lj_acc(s);
  if (s->ljl_on)
    orc_lj_list_acc(s);
  if (s->er_on)
    orc_ewald_r_acc(s);
  if (s->sr_on)
    orc_short_range_acc(s);
  if (s->ek_on) {
    if (!s->ek_init)
      ek_init(s);
    ek_structure_factor(s, 0, s->nk);
    ek_force_pass_idx(s, 0, m, idx);
  }
  if (s->elc_on)
    orc_elc_acc(s);
}

int orc_setup_sampled(orc_sys *s, int m, const int *idx)
{
  s->virial = 0;
  if (s->nl_on)
    nl_init(s);
  orc_wrap(s);
  orc_ghosts(s);
  orc_build(s, 1);
  orc_forces_sampled(s, m, idx);
  return 0;
}

/* The force fields in the order the fixtures' scripts add them to the
 * simulator (md_simulator.cpp:370-373 evaluates them in that order). */
void orc_forces(orc_sys *s)
{
  orc_reset_acc(s);
  if (s->lj_on)
    orc_lj_acc(s);
  if (s->ljl_on)
    orc_lj_list_acc(s);
  if (s->er_on)
    orc_ewald_r_acc(s);
  if (s->sr_on)
    orc_short_range_acc(s);
  if (s->ek_on)
    orc_ewald_k_acc(s);
  if (s->elc_on)
    orc_elc_acc(s);
}

/* ------------------------------------------- temperature, pressure, thermostats -- */

/* atom_data k_b, n_r_df (-1: from the boundary conditions), temperature_process,
 * pressure_process (atom_data.cpp:121-151). */
void orc_set_thermo(orc_sys *s, double k_b, int n_r_df, int temperature_process, int pressure_process)
{
  s->k_b = k_b;
  s->n_r_df = n_r_df;
  s->temperature_process = temperature_process;
  s->pressure_process = pressure_process;
  s->temperature = s->pressure = 0.0; /* reset_temperature / reset_pressure */
}

void orc_set_step_index(orc_sys *s, long step) { s->step_index = step; }

/* get_n_r_df (atom_data.cpp:322-350), without stochastic forces. */
static int n_r_df(const orc_sys *s)
{
  if (s->n_r_df >= 0)
    return s->n_r_df;
  const int bc_sum = s->bc[0] + s->bc[1] + s->bc[2];
  return bc_sum == 0 ? 6 : (bc_sum == 3 ? 3 : 6 - bc_sum);
}

/* owned_velocity_cm (kinetic_energy.cpp:413-457): sum m v / sum m; also the
 * mass sum. */
static double velocity_cm(const orc_sys *s, double vcm[3])
{
  double f[3] = {0, 0, 0}, msum = 0.0;
  for (int i = 0; i < s->n; ++i) {
    const double m = s->mass[s->type[i]];
    msum += m;
    for (int d = 0; d < 3; ++d)
      f[d] += s->vel[3 * i + d] * m;
  }
  for (int d = 0; d < 3; ++d)
    vcm[d] = f[d] / msum;
  return msum;
}

/* kinetic_energy_mpi_domain (kinetic_energy.cpp:30-138): 1/2 sum m v'^2 with
 * v' = v for n_r_df 0 and v - v_cm for 1..3 (owned_velocity_cm, :413-457).
 * For 4..6 the reference forms the inertia tensor I about the mass centre
 * (:611-640, p_cm :391-411) and tests matrix_inverse_3d(I, I^-1) != 0
 * (:92-105), whose 0 is success: an invertible I (det != 0, the usual case)
 * takes the warning branch and leaves the kinetic energy 0; an exactly
 * singular one leaves I^-1 zero, so only v_cm is removed.  The rigid-rotation
 * term (I^-1 L) x (r - p_cm) is then zero in every case and L never matters. */
double orc_kinetic_energy(const orc_sys *s)
{
  const int nr = n_r_df(s);
  double vcm[3] = {0, 0, 0};
  if (nr >= 1) {
    const double msum = velocity_cm(s, vcm);
    if (nr >= 4) {
      double pcm[3] = {0, 0, 0}, A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int i = 0; i < s->n; ++i) {
        const double m = s->mass[s->type[i]];
        for (int d = 0; d < 3; ++d)
          pcm[d] += s->pos[3 * i + d] * m;
      }
      for (int d = 0; d < 3; ++d)
        pcm[d] = pcm[d] / msum;
      for (int i = 0; i < s->n; ++i) {
        const double m = s->mass[s->type[i]];
        const double p[3] = {s->pos[3 * i] - pcm[0], s->pos[3 * i + 1] - pcm[1],
                             s->pos[3 * i + 2] - pcm[2]};
        A[0][0] += m * (p[1] * p[1] + p[2] * p[2]);
        A[1][1] += m * (p[0] * p[0] + p[2] * p[2]);
        A[2][2] += m * (p[0] * p[0] + p[1] * p[1]);
        A[0][1] -= m * (p[0] * p[1]);
        A[1][2] -= m * (p[1] * p[2]);
        A[2][0] -= m * (p[2] * p[0]);
      }
      A[1][0] = A[0][1];
      A[2][1] = A[1][2];
      A[0][2] = A[2][0];
      /* matrix_inverse_3d's determinant (common_template_functions.h:168-172) */
      const double det = (A[0][0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2])) -
                         (A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])) +
                         (A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]));
      if (det != 0.0)
        return 0.0;
    }
  }
  double e = 0.0;
  for (int i = 0; i < s->n; ++i) {
    const double m = s->mass[s->type[i]];
    double v[3];
    for (int d = 0; d < 3; ++d)
      v[d] = nr >= 1 ? s->vel[3 * i + d] - vcm[d] : s->vel[3 * i + d];
    e += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  }
  return e * 0.5;
}

/* degree_of_freedoms_mpi_domain (kinetic_energy.cpp:681-718) without molecules' bonds. */
static int dof(const orc_sys *s) { return 3 * s->n - n_r_df(s); }

/* finalize_temperature (kinetic_energy.cpp:162-202): 2 KE / (k_b dof). */
static void finalize_temperature(orc_sys *s)
{
  if (!s->temperature_process)
    return;
  s->temperature = 2.0 * orc_kinetic_energy(s) / (s->k_b * dof(s));
}

/* finalize_pressure (kinetic_energy.cpp:210-292): n k_b T / V + virial / (3V). */
static void finalize_pressure(orc_sys *s)
{
  if (!s->pressure_process)
    return;
  const double volume = s->L[0] * s->L[1] * s->L[2];
  const double p2 = (0.0 + 0 + s->virial) / (3.0 * volume);
  const double p1 = ((double)s->n * s->k_b * s->temperature) / volume;
  s->pressure = p1 + p2;
}

double orc_temperature(const orc_sys *s) { return s->temperature; }
double orc_pressure(const orc_sys *s) { return s->pressure; }

/* constraint berendsen: temperature, coupling, dt, step, lambda_min, lambda_max
 * (berendsen.cpp:33-118, defaults berendsen.h:50-60). */
void orc_set_berendsen(orc_sys *s, double t0, double coupling, double dt, int step, double lmin,
                       double lmax)
{
  s->ber_on = 1;
  s->ber_t0 = t0;
  s->ber_coupling = coupling;
  s->ber_dt = dt;
  s->ber_step = step;
  s->ber_lmin = lmin;
  s->ber_lmax = lmax;
}

/* constraint nve: energy_per_dof, energy_tot, kbt, kb + temperature, step
 * (nve.cpp:33-132); unset values <= 0. */
void orc_set_nve(orc_sys *s, double e_dof, double e_tot, double kbt, double kb, double t, int step)
{
  s->nve_on = 1;
  s->nve_e_dof = e_dof;
  s->nve_e_tot = e_tot;
  s->nve_kbt = kbt;
  s->nve_kb = kb;
  s->nve_t = t;
  s->nve_step = step;
}

static void scale_velocities(orc_sys *s, double lambda)
{
  for (int q = 0; q < 3 * s->n; ++q)
    s->vel[q] *= lambda;
}

/* apply_thermostat of berendsen (berendsen.cpp:120-146) and nve (nve.cpp:134-178)
 * at timestep `ts`, in that order; returns whether any scaled. */
static int thermostats(orc_sys *s, long ts)
{
  int scaled = 0;
  if (s->ber_on && ts % s->ber_step == 0) {
    scaled = 1;
    const double t = s->temperature;
    double lambda = sqrt(1.0 + (s->ber_dt / s->ber_coupling) * ((s->ber_t0 / t) - 1.0));
    if (lambda < s->ber_lmin)
      lambda = s->ber_lmin;
    if (lambda > s->ber_lmax)
      lambda = s->ber_lmax;
    scale_velocities(s, lambda);
  }
  if (s->nve_on && ts % s->nve_step == 0) {
    scaled = 1;
    const double e_now = orc_kinetic_energy(s);
    double lambda = 1.0;
    if (s->nve_e_dof > 0.0)
      lambda = sqrt(s->nve_e_dof * dof(s) / e_now);
    else if (s->nve_e_tot > 0.0)
      lambda = sqrt(s->nve_e_tot / e_now);
    else if (s->nve_kbt > 0.0)
      lambda = sqrt(dof(s) * s->nve_kbt * 0.5 / e_now);
    else if (s->nve_t > 0.0)
      lambda = sqrt(dof(s) * s->nve_kb * s->nve_t * 0.5 / e_now);
    scale_velocities(s, lambda);
  }
  return scaled;
}

/* ---------------------------------------------------------------- driver -- */

int orc_setup(orc_sys *s)
{
  s->virial = 0;
  if (s->nl_on)
    nl_init(s);
  orc_wrap(s);
  orc_ghosts(s);
  orc_build(s, 1);
  orc_forces(s);
  finalize_temperature(s); /* md_simulator.cpp:543 */
  return 0;
}

/* integrate_velocity_verlet (md_simulator.cpp:555-617). */
int orc_step(orc_sys *s, double dt)
{
  s->virial = 0; /* Md_simulator::step resets the virial (md_simulator.cpp:408) */
  for (int i = 0; i < 3 * s->n; ++i)
    if (isnan(s->acc[i]))
      return 1;
  for (int i = 0; i < s->n; ++i)
    for (int d = 0; d < 3; ++d) {
      const int q = 3 * i + d;
      s->pos[q] += s->vel[q] * dt + 0.5 * s->acc[q] * dt * dt;
      s->acc_old[q] = s->acc[q];
    }
  /* re_calculate_acc (md_simulator.cpp:361-374) */
  int update = orc_wrap(s);
  orc_ghosts(s);
  orc_build(s, update);
  orc_forces(s);
  for (int q = 0; q < 3 * s->n; ++q)
    s->vel[q] += 0.5 * dt * (s->acc[q] + s->acc_old[q]);
  /* md_simulator.cpp:619-633: temperature, thermostats (and the temperature
   * again after they scaled), pressure */
  finalize_temperature(s);
  if (thermostats(s, s->step_index))
    finalize_temperature(s);
  finalize_pressure(s);
  s->step_index++;
  return 0;
}

/* ------------------------------------------------------ energies, queries -- */

/* Electrostatic_ewald_r::potential(const Vector&) (electrostatic_ewald_r/
 * potential.cpp:30-86) and field (field.cpp:31-96): every owned or ghost atom
 * in the 27 cells around r's bin, no cutoff, exact-zero separation skipped. */
static void point_r(orc_sys *s, const double *r, double *field, double *pot)
{
  const double alpha = s->er_alpha, a2 = alpha * alpha;
  double phi = 0, e[3] = {0, 0, 0};
  const int b = bin_of(s, r);
  for (int q = s->st_off[b]; q < s->st_off[b + 1]; ++q) {
    const int c = s->st[q];
    for (int t = s->bin_off[c]; t < s->bin_off[c + 1]; ++t) {
      const int j = s->bin_items[t];
      const double *pj = apos(s, j);
      const double qj = s->charge[atype(s, j)];
      double d[3] = {r[0] - pj[0], r[1] - pj[1], r[2] - pj[2]};
      if (d[0] == 0 && d[1] == 0 && d[2] == 0)
        continue;
      const double d2 = dot3(d), dn = sqrt(d2);
      phi += qj * erfc(alpha * dn) / dn;
      const double sr = 2 * alpha * ORC_PIS_INV * exp(-a2 * d2) + erfc(alpha * dn) / dn;
      for (int k = 0; k < 3; ++k)
        e[k] += qj * (sr * (d[k] / d2));
    }
  }
  if (pot)
    *pot = phi * s->er_ke;
  if (field)
    for (int k = 0; k < 3; ++k)
      field[k] = e[k] * s->er_ke;
}

double orc_energy_r(orc_sys *s)
{
  if (!s->build_cells)
    return NAN; /* the cell walk needs build_cell_list, as in the reference */
  update_cell_list(s);
  double e = 0;
  for (int j = 0; j < s->n; ++j) {
    double phi;
    point_r(s, s->pos + 3 * j, NULL, &phi);
    e += s->charge[s->type[j]] * phi;
  }
  return 0.5 * e; /* energy.cpp:30-51 */
}

void orc_point_r(orc_sys *s, int np, const double *pts, double *field, double *pot)
{
  if (!s->build_cells)
    return;
  update_cell_list(s);
  for (int p = 0; p < np; ++p)
    point_r(s, pts + 3 * p, field ? field + 3 * p : NULL, pot ? pot + p : NULL);
}

/* k_space_potential(const Vector&) (electrostatic_ewald_k/potential.cpp:107-129)
 * and k_space_field (field.cpp:80-115), plus the dipole field when enabled. */
static void point_k(const orc_sys *s, const double *r, double *field, double *pot)
{
  double sum = 0, fx = 0, fy = 0, fz = 0;
  for (int k = 0; k < s->nk; ++k) {
    const double *kv = s->kvec + 3 * k;
    const double th = kv[0] * r[0] + kv[1] * r[1] + kv[2] * r[2];
    const double c = cos(th), sn = sin(th);
    const double sr = s->sk[2 * k], si = s->sk[2 * k + 1];
    /* Re(S * exp(-i th)) */
    sum += s->fk[k] * (sr * c + si * sn);
    /* Im(f * conj(S) * exp(i th)) */
    const double im = s->fk[k] * (sr * sn - si * c);
    fx += kv[0] * im;
    fy += kv[1] * im;
    fz += kv[2] * im;
  }
  const double c0 = s->ek_ke * s->lxyz_inv * ORC_4PI;
  if (pot)
    *pot = ORC_4PI * s->lxyz_inv * sum * s->ek_ke;
  if (field) {
    field[0] = c0 * fx;
    field[1] = c0 * fy;
    field[2] = c0 * fz;
    if (s->dipole)
      for (int d = 0; d < 3; ++d)
        field[d] += s->dip_field[d];
  }
}

void orc_point_k(orc_sys *s, int np, const double *pts, double *field, double *pot)
{
  for (int p = 0; p < np; ++p)
    point_k(s, pts + 3 * p, field ? field + 3 * p : NULL, pot ? pot + p : NULL);
}

double orc_self_energy(orc_sys *s)
{
  double sum = 0;
  for (int j = 0; j < s->n; ++j) {
    const double q = s->charge[s->type[j]];
    sum += q * q;
  }
  return -ORC_PIS_INV * s->ek_alpha * sum * s->ek_ke; /* energy.cpp:31-50 */
}

/* Electrostatic_ewald_k::energy (electrostatic_ewald_k/energy.cpp:61-116); valid
 * right after a force evaluation at the same positions (S(k) is refreshed there). */
double orc_energy_k(orc_sys *s)
{
  double e = 0;
  for (int j = 0; j < s->n; ++j) {
    double phi;
    point_k(s, s->pos + 3 * j, NULL, &phi);
    e += s->charge[s->type[j]] * phi;
  }
  e = 0.5 * e + orc_self_energy(s);
  if (s->dipole)
    e += 0.5 * s->dip_coef * dot3(s->dip_sum);
  return e;
}

/* Shifted-force LJ energy consistent with the reference force (lj_cell_list.cpp:
 * 143-155): U(r) = 4e[(s/r)^12 - (s/r)^6 + g_c r^2/2] - (same at rc), g_c = 12s^12/rc^14
 * - 6s^6/rc^8.  The reference has no LJ energy; this pairs with its force. */
double orc_energy_lj(orc_sys *s)
{
  if (!s->build_cells)
    return NAN;
  update_cell_list(s);
  const double rc = s->lj_rc, rc2 = rc * rc;
  double e = 0;
  for (int i = 0; i < s->n; ++i) {
    const double *pi = s->pos + 3 * i;
    const int ti = s->type[i];
    const int b = bin_of(s, pi);
    for (int q = s->st_off[b]; q < s->st_off[b + 1]; ++q) {
      const int c = s->st[q];
      for (int t = s->bin_off[c]; t < s->bin_off[c + 1]; ++t) {
        const int j = s->bin_items[t];
        const int ghost = j >= s->n;
        if (!ghost && !(i < j))
          continue;
        const double *pj = apos(s, j);
        const int tj = atype(s, j);
        double dr[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
        const double r2 = dot3(dr);
        if (r2 > rc2)
          continue;
        const double ep = s->eps[ti * s->nt + tj], sg = s->sig[ti * s->nt + tj];
        const double s2 = sg * sg, s6 = s2 * s2 * s2, s12 = s6 * s6;
        const double gc = 12 * s12 / pow(rc, 14) - 6 * s6 / pow(rc, 8);
        const double u = s12 / pow(r2, 6) - s6 / pow(r2, 3) + 0.5 * gc * r2;
        const double uc = s12 / pow(rc2, 6) - s6 / pow(rc2, 3) + 0.5 * gc * rc2;
        /* a ghost pair is seen from both owned ends: count half each time */
        e += (ghost ? 0.5 : 1.0) * 4 * ep * (u - uc);
      }
    }
  }
  return e;
}

/* --------------------------------------------------------------- getters -- */

int orc_n(const orc_sys *s) { return s->n; }

static double *arr_of(const orc_sys *s, int what)
{
  switch (what) {
  case 0:
    return s->pos;
  case 1:
    return s->vel;
  case 2:
    return s->acc;
  default:
    return s->acc_old;
  }
}

void orc_get(const orc_sys *s, int what, double *out) { memcpy(out, arr_of(s, what), sizeof(double) * 3 * s->n); }
void orc_set(orc_sys *s, int what, const double *in) { memcpy(arr_of(s, what), in, sizeof(double) * 3 * s->n); }
int orc_num_ghosts(const orc_sys *s) { return s->g; }

void orc_get_ghosts(const orc_sys *s, double *gpos, int *gsrc)
{
  if (gpos)
    memcpy(gpos, s->gpos, sizeof(double) * 3 * s->g);
  if (gsrc)
    memcpy(gsrc, s->gsrc, sizeof(int) * s->g);
}

long orc_num_pairs(const orc_sys *s) { return s->nl_off ? s->nl_off[s->n] : 0; }

void orc_get_pairs(const orc_sys *s, int *ij)
{
  long w = 0;
  for (int i = 0; i < s->n; ++i)
    for (long t = s->nl_off[i]; t < s->nl_off[i + 1]; ++t) {
      ij[2 * w] = i;
      ij[2 * w + 1] = s->nl_items[t];
      ++w;
    }
}

void orc_get_nbins(const orc_sys *s, int nb[3])
{
  nb[0] = s->nb[0];
  nb[1] = s->nb[1];
  nb[2] = s->nb[2];
}

double orc_virial(const orc_sys *s) { return s->virial; }
void orc_reset_virial(orc_sys *s) { s->virial = 0; }
int orc_num_kvectors(const orc_sys *s) { return s->nk; }

void orc_get_kvectors(const orc_sys *s, double *kvec, double *fk)
{
  if (kvec)
    memcpy(kvec, s->kvec, sizeof(double) * 3 * s->nk);
  if (fk)
    memcpy(fk, s->fk, sizeof(double) * s->nk);
}

void orc_get_sk(const orc_sys *s, double *sk) { memcpy(sk, s->sk, sizeof(double) * 2 * s->nk); }

/* ------------------------------------------------------ force-matching adjoint --
 * FP64 reference of g = sum_i w_i . dF_i/dtheta (no counterpart in the reference;
 * the derivative of its force formulas).  Written independently of the GPU
 * kernels: pairs once per reference list entry with Newton's third law, so each
 * pair term is contracted with (w_i - w_j) (w_j = 0 for a ghost partner), and
 * k-space over the reference's full truncated set K without symmetry.  F = mass *
 * acceleration.  Outputs: g_eps[nt*nt], g_sig[nt*nt], g_q[nt]. */
void orc_fm_adjoint(orc_sys *s, const double *w, double *g_eps, double *g_sig, double *g_q)
{
  const int nt = s->nt, nt2 = nt * nt;
  for (int p = 0; p < nt2; ++p)
    g_eps[p] = g_sig[p] = 0;
  for (int t = 0; t < nt; ++t)
    g_q[t] = 0;
  /* LJ: the cell walk of Lj_cell_list (owned j > i and every ghost) */
  if (s->lj_on && s->build_cells) {
    update_cell_list(s);
    const double rc = s->lj_rc, rc2 = rc * rc;
    const double rci2 = 1.0 / rc2, c6 = rci2 * rci2 * rci2, c8 = c6 * rci2, c14 = c6 * c6 * rci2;
    for (int i = 0; i < s->n; ++i) {
      const double *pi = s->pos + 3 * i;
      const int ti = s->type[i];
      const int b = bin_of(s, pi);
      for (int q = s->st_off[b]; q < s->st_off[b + 1]; ++q) {
        const int c = s->st[q];
        for (int t = s->bin_off[c]; t < s->bin_off[c + 1]; ++t) {
          const int j = s->bin_items[t];
          const int ghost = j >= s->n;
          if (!ghost && !(i < j))
            continue;
          const double *pj = apos(s, j);
          const int tj = atype(s, j);
          double dr[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
          const double r2 = dot3(dr);
          if (r2 > rc2)
            continue;
          double dw[3] = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
          if (!ghost)
            for (int d = 0; d < 3; ++d)
              dw[d] -= w[3 * j + d];
          const double wr = dw[0] * dr[0] + dw[1] * dr[1] + dw[2] * dr[2];
          const int p = ti * nt + tj;
          const double e = s->eps[p], sg = s->sig[p];
          const double ri2 = 1.0 / r2, r6 = ri2 * ri2 * ri2, r8 = r6 * ri2, r14 = r6 * r6 * ri2;
          const double s2 = sg * sg, s5 = s2 * s2 * sg, s6 = s5 * sg, s11 = s6 * s5, s12 = s6 * s6;
          g_eps[p] += wr * 4.0 * (-12 * s12 * r14 + 6 * s6 * r8 + 12 * s12 * c14 - 6 * s6 * c8);
          g_sig[p] += wr * 4.0 * e * (-144 * s11 * r14 + 36 * s5 * r8 + 144 * s11 * c14 - 36 * s5 * c8);
        }
      }
    }
  }
  /* Ewald_r: every reference Verlet-list entry */
  if (s->er_on && s->nl_off) {
    const double alpha = s->er_alpha, a2 = alpha * alpha;
    for (int i = 0; i < s->n; ++i) {
      const double *pi = s->pos + 3 * i;
      const int ti = s->type[i];
      for (long t = s->nl_off[i]; t < s->nl_off[i + 1]; ++t) {
        const int j = s->nl_items[t];
        const int ghost = j >= s->n;
        const double *pj = apos(s, j);
        const int tj = atype(s, j);
        double lam = 1.0;
        if (s->lambda_set && s->mol[i] == amol(s, j)) {
          if (s->lambda[ti * nt + tj] == 0.0)
            continue;
          lam = s->lambda[ti * nt + tj];
        }
        double r[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
        if (r[0] == 0 && r[1] == 0 && r[2] == 0)
          continue;
        double dw[3] = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
        if (!ghost)
          for (int d = 0; d < 3; ++d)
            dw[d] -= w[3 * j + d];
        const double wr = dw[0] * r[0] + dw[1] * r[1] + dw[2] * r[2];
        const double r2 = dot3(r), rn = sqrt(r2);
        const double sr = (2 * alpha * ORC_PIS_INV * exp(-a2 * r2) + erfc(alpha * rn) / rn) / r2;
        const double base = -lam * s->er_ke * sr * wr;
        g_q[ti] += base * s->charge[tj];
        g_q[tj] += base * s->charge[ti];
      }
    }
  }
  /* Ewald_k over the full truncated set K */
  if (s->ek_on) {
    if (!s->ek_init)
      ek_init(s);
    const double C = ORC_4PI * s->ek_ke * s->lxyz_inv;
    double *G = (double *)calloc(3 * (size_t)(s->n ? s->n : 1), sizeof(double));
    double *gk = (double *)calloc((size_t)nt, sizeof(double));
#pragma omp parallel num_threads(g_threads)
    {
      double *gk_loc = (double *)calloc((size_t)nt, sizeof(double));
      double complex *ss = (double complex *)malloc(sizeof(double complex) * (nt ? nt : 1));
#pragma omp for schedule(static)
      for (int k = 0; k < s->nk; ++k) {
        const double *kv = s->kvec + 3 * k;
        double complex S = 0, T = 0;
        for (int t = 0; t < nt; ++t)
          ss[t] = 0;
        for (int j = 0; j < s->n; ++j) {
          const double *p = s->pos + 3 * j;
          const double th = kv[0] * p[0] + kv[1] * p[1] + kv[2] * p[2];
          const double complex e = cos(th) + I * sin(th);
          const double q = s->charge[s->type[j]];
          S += q * e;
          ss[s->type[j]] += e;
          T += q * (w[3 * j] * kv[0] + w[3 * j + 1] * kv[1] + w[3 * j + 2] * kv[2]) * e;
        }
        for (int t = 0; t < nt; ++t)
          gk_loc[t] += s->fk[k] * cimag(conj(ss[t]) * T);
      }
#pragma omp critical
      for (int t = 0; t < nt; ++t)
        gk[t] += gk_loc[t];
      free(gk_loc);
      free(ss);
    }
    /* per-atom field G_i = C sum_k f k Im(conj(S_k) e^{ik.r_i}) with S_k from ek_structure_factor */
    ek_structure_factor(s, 0, s->nk);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int i = 0; i < s->n; ++i) {
      const double *p = s->pos + 3 * i;
      double g[3] = {0, 0, 0};
      for (int k = 0; k < s->nk; ++k) {
        const double *kv = s->kvec + 3 * k;
        const double th = kv[0] * p[0] + kv[1] * p[1] + kv[2] * p[2];
        const double im = s->sk[2 * k] * sin(th) - s->sk[2 * k + 1] * cos(th);
        for (int d = 0; d < 3; ++d)
          g[d] += s->fk[k] * kv[d] * im;
      }
      for (int d = 0; d < 3; ++d)
        G[3 * i + d] = C * g[d];
    }
    for (int i = 0; i < s->n; ++i)
      g_q[s->type[i]] += w[3 * i] * G[3 * i] + w[3 * i + 1] * G[3 * i + 1] + w[3 * i + 2] * G[3 * i + 2];
    for (int t = 0; t < nt; ++t)
      g_q[t] += C * gk[t];
    free(G);
    free(gk);
  }
}
