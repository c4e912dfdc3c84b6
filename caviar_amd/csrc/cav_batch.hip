// Batched replicas: many independent systems of one force field (LJ +
// Ewald_r + Ewald_k, Verlet list from cells), each with its own atom count and
// periodic box, advanced together.  Every kernel covers all replicas at once
// (grid row = replica for per-atom work, a global cell or work-item table for
// the rest), each replica keeps its own rebuild / ghost / overflow decisions in
// its own flag words, and one velocity-Verlet step of all replicas is one
// launch of a captured hipGraph.  The Ewald_k tiles are picked by a timing
// autotuner on the batch itself (cav_batch_autotune).
//
// Layout: slot arrays (positions, types, charges, cells) hold replica r's owned
// atoms then its ghosts from reps[r].base; owned arrays (velocities,
// accelerations, Verlet rows, phase tables) hold its owned atoms from
// reps[r].obase.  Per replica the arithmetic is the single-system path's, so
// each replica reproduces a single cav_ctx run of the same system.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cav_internal.hpp"

namespace cav {

int bfail(cav_batch *b, int code, const std::string &msg)
{
  if (b) b->err = msg;
  return code;
}

int bhip_check(cav_batch *b, hipError_t e, const char *what)
{
  return bfail(b, CAV_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ----------------------------------------------------------------- kernels --

__global__ void kb_begin(int B, int *flags)
{
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r >= B) return;
  int *f = flags + 16 * r;
  f[F_WRAPPED] = 0;
  f[F_REBUILD] = f[F_FORCE];
  f[F_FORCE] = 0;
  f[F_NAN] = 0;
  f[F_OVERFLOW] = 0;
  f[F_MAXNB] = 0;
  f[F_GOVER] = 0;
}

__global__ void kb_set_all(int B, int *flags, int idx, int v)
{
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < B) flags[16 * r + idx] = v;
}

// exchange_owned for every replica (setup)
__global__ void kb_wrap(const RepDev *rep, double *x, double *y, double *z, int *flags)
{
  const RepDev &r = rep[blockIdx.y];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= r.n) return;
  double *p[3] = {x + r.base + i, y + r.base + i, z + r.base + i};
  bool wrapped = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double v = *p[d];
    if (v < r.lo[d]) {
      v = xadd(v, r.L[d]);
      wrapped = true;
    } else if (v > r.hi[d]) {
      v = xsub(v, r.L[d]);
      wrapped = true;
    }
    *p[d] = v;
  }
  if (wrapped) {
    flags[16 * blockIdx.y + F_WRAPPED] = 1;
    flags[16 * blockIdx.y + F_REBUILD] = 1;
  }
}

// drift + wrap + NaN check, as the single-system k_drift (md_simulator.cpp:555-590)
__global__ void kb_drift(const RepDev *rep, double dt, double *x, double *y, double *z,
                         const double *vx, const double *vy, const double *vz, const double *ax,
                         const double *ay, const double *az, double *aox, double *aoy, double *aoz,
                         int *flags)
{
  const RepDev &r = rep[blockIdx.y];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= r.n) return;
  int *f = flags + 16 * blockIdx.y;
  const long o = r.obase + i, s = r.base + i;
  const double a[3] = {ax[o], ay[o], az[o]};
  if (isnan(a[0]) || isnan(a[1]) || isnan(a[2])) f[F_NAN] = 1;
  const double v[3] = {vx[o], vy[o], vz[o]};
  double *p[3] = {x + s, y + s, z + s};
  bool wrapped = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double h = xmul(xmul(xmul(0.5, a[d]), dt), dt);
    double q = xadd(*p[d], xadd(xmul(v[d], dt), h));
    if (q < r.lo[d]) {
      q = xadd(q, r.L[d]);
      wrapped = true;
    } else if (q > r.hi[d]) {
      q = xsub(q, r.L[d]);
      wrapped = true;
    }
    *p[d] = q;
  }
  aox[o] = a[0];
  aoy[o] = a[1];
  aoz[o] = a[2];
  if (wrapped) {
    f[F_WRAPPED] = 1;
    f[F_REBUILD] = 1;
  }
}

__global__ void kb_kick(long nown, double hdt, double *vx, double *vy, double *vz, const double *ax,
                        const double *ay, const double *az, const double *aox, const double *aoy,
                        const double *aoz)
{
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nown) return;
  vx[i] = xadd(vx[i], xmul(hdt, xadd(ax[i], aox[i])));
  vy[i] = xadd(vy[i], xmul(hdt, xadd(ay[i], aoy[i])));
  vz[i] = xadd(vz[i], xmul(hdt, xadd(az[i], aoz[i])));
}

// positions (slots) and velocities/accelerations (owned) <-> the snapshot
__global__ void kb_snapshot(const RepDev *rep, long nown, int save, double *x, double *y, double *z,
                            double *vx, double *vy, double *vz, double *ax, double *ay, double *az,
                            double *snap)
{
  const RepDev &r = rep[blockIdx.y];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= r.n) return;
  const long o = r.obase + i, s = r.base + i;
  double *a[9] = {x + s, y + s, z + s, vx + o, vy + o, vz + o, ax + o, ay + o, az + o};
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    double *sv = snap + k * nown + o;
    if (save)
      *sv = *a[k];
    else
      *a[k] = *sv;
  }
}

}  // namespace cav

using namespace cav;

// ------------------------------------------------------------------- host --

static void batch_free(cav_batch *b)
{
  void *ptrs[] = {b->d_rep, b->cell_rep, b->st_off, b->st, b->x, b->y, b->z, b->xo, b->yo, b->zo,
                  b->type, b->id, b->q, b->xq, b->tid, b->vx, b->vy, b->vz, b->ax, b->ay, b->az,
                  b->aox, b->aoy, b->aoz, b->type_o, b->q_o, b->snap, b->bin_of, b->bin_cnt,
                  b->bin_start, b->bin_items, b->gblk, b->nl, b->nn, b->flags, b->d_minv, b->d_lj,
                  b->part, b->vir, b->d_W, b->tab, b->ztab, b->bmat, b->bmat_h, b->sk_part, b->sk,
                  b->epart, b->s_items, b->s_items_o, b->f_items, b->rchunks, b->cpos, b->d_raw_eps,
                  b->d_raw_sig, b->d_charge, b->nlb, b->xqs, b->tids, b->litems, b->binpos,
                  b->order, b->ocnt, b->ostart};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (b->h_flags) (void)hipHostFree(b->h_flags);
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
}

template <typename X>
static int balloc(cav_batch *b, X **p, size_t count)
{
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  BAT_HIP(hipMalloc(p, sizeof(X) * std::max<size_t>(count, 1)));
  return CAV_OK;
}

// P5 image count of one atom in a box (host estimate of the ghost capacity)
static int images_of(const double *p, const double *lo, const double *hi, double gc)
{
  int s[3];
  for (int d = 0; d < 3; ++d) s[d] = (p[d] < lo[d] + gc) || (p[d] > hi[d] - gc);
  return s[0] + s[1] + s[2] + (s[0] & s[1]) + (s[0] & s[2]) + (s[1] & s[2]) + (s[0] & s[1] & s[2]);
}

// Replica table, stencils, cells and every buffer, from the host state.
static int batch_layout(cav_batch *b)
{
  b->B = (int)b->hreps.size();
  b->reps.assign(b->B, RepDev{});
  b->rce = b->nl_cutoff * b->nl_coef;
  b->rce2 = b->rce * b->rce;
  b->thr_sq = 0.25 * (b->rce - b->nl_cutoff) * (b->rce - b->nl_cutoff);
  b->cinv = 1.0 / b->rce;
  long slots = 0, own = 0, tabs = 0;
  int cells = 0, maxn = 0;
  b->h_cell_rep.clear();
  b->h_st_off.clear();
  b->h_st.clear();
  size_t list_need = 16;
  for (int r = 0; r < b->B; ++r) {
    cav_batch::HostRep &h = b->hreps[r];
    RepDev &d = b->reps[r];
    for (int k = 0; k < 3; ++k) {
      d.lo[k] = h.lo[k];
      d.hi[k] = h.hi[k];
      d.L[k] = h.hi[k] - h.lo[k];
      d.il[k] = 1.0 / d.L[k];
    }
    if (b->ghost_cutoff >= std::min(d.L[0], std::min(d.L[1], d.L[2])))
      return bfail(b, CAV_ERR_ARG, "replica box shorter than the ghost cutoff");
    if (h.gcap <= 0) {
      long g = 0;
      for (int i = 0; i < h.n; ++i) g += images_of(&h.pos[3 * (size_t)i], d.lo, d.hi, b->ghost_cutoff);
      h.gcap = (int)std::min<long>(7L * h.n, (long)(1.25 * g) + 64);
    }
    d.n = h.n;
    d.gcap = h.gcap;
    d.base = (int)slots;
    d.obase = (int)own;
    d.tbase = (int)tabs;
    slots += ((long)h.n + h.gcap + 63) / 64 * 64;
    own += h.n;
    tabs += ((long)h.n + 31) / 32 * 32;
    maxn = std::max(maxn, h.n);
    // cells of the reference's width (cutoff_extra) on the reference's grid
    // (neighborlist.cpp:581-697), plus one layer of cells past each face that
    // holds the ghosts where they lie instead of clamping them into the edge
    // cells; the ghosts are the periodic images, so a cell's 27-stencil is
    // its plain neighbourhood in this grid (no wrap).  Small dense replicas
    // put several times their own atoms in the ghost shell, so clamping it
    // into the edge cells made every edge atom test the whole shell.
    d.cbase = cells;
    int nbins = 1;
    for (int k = 0; k < 3; ++k) {
      d.nb[k] = std::max(1, (int)std::ceil(d.L[k] * b->cinv)) + 2;
      nbins *= d.nb[k];
    }
    d.nbins = nbins;
    const int nbx = d.nb[0], nby = d.nb[1], nbz = d.nb[2];
    for (int k = 0; k < nbins; ++k) {
      const int cx = k % nbx, cy = (k / nbx) % nby, cz = k / (nbx * nby);
      b->h_st_off.push_back((int)b->h_st.size());
      for (int a = cx - 1; a <= cx + 1; ++a)
        for (int c = cy - 1; c <= cy + 1; ++c)
          for (int e = cz - 1; e <= cz + 1; ++e)
            if (a >= 0 && a < nbx && c >= 0 && c < nby && e >= 0 && e < nbz)
              b->h_st.push_back(cells + a + nbx * c + nbx * nby * e);
      b->h_cell_rep.push_back(r);
    }
    cells += nbins;
    const double vol = d.L[0] * d.L[1] * d.L[2];
    const double expect = 4.18879020478639 * b->rce2 * b->rce * (h.n / vol);
    list_need = std::max<size_t>(list_need, (size_t)(2.0 * expect) + 32);
  }
  b->h_st_off.push_back((int)b->h_st.size());
  b->nslots = slots;
  b->nown = own;
  b->ntab = tabs;
  b->ncells = cells;
  b->maxn = maxn;
  b->scan_blocks = grid_for(maxn);
  b->pair_blocks = grid_for((long)maxn * kLPA);
  for (int r = 0; r < b->B; ++r) b->reps[r].pbase = r * b->pair_blocks;
  if (b->cap < (int)list_need) b->cap = (int)((list_need + 3) / 4 * 4);
  CAV_TRY(batch_ewald_k_plan(b));

  // device buffers
  CAV_TRY(balloc(b, &b->d_rep, b->B));
  CAV_TRY(balloc(b, &b->cell_rep, b->ncells));
  CAV_TRY(balloc(b, &b->st_off, b->h_st_off.size()));
  CAV_TRY(balloc(b, &b->st, b->h_st.size()));
  double **sd[] = {&b->x, &b->y, &b->z, &b->xo, &b->yo, &b->zo, &b->q};
  for (double **p : sd) CAV_TRY(balloc(b, p, slots));
  CAV_TRY(balloc(b, &b->type, slots));
  CAV_TRY(balloc(b, &b->id, slots));
  CAV_TRY(balloc(b, &b->xq, slots));
  CAV_TRY(balloc(b, &b->tid, slots));
  CAV_TRY(balloc(b, &b->bin_of, slots));
  CAV_TRY(balloc(b, &b->bin_items, slots));
  CAV_TRY(balloc(b, &b->cpos, slots));
  double **od[] = {&b->vx, &b->vy, &b->vz, &b->ax, &b->ay, &b->az, &b->aox, &b->aoy, &b->aoz, &b->q_o};
  for (double **p : od) CAV_TRY(balloc(b, p, own));
  CAV_TRY(balloc(b, &b->type_o, own));
  CAV_TRY(balloc(b, &b->snap, 9 * (size_t)own));
  CAV_TRY(balloc(b, &b->bin_cnt, 2 * (size_t)b->ncells));
  CAV_TRY(balloc(b, &b->bin_start, (size_t)b->ncells + b->B));
  CAV_TRY(balloc(b, &b->gblk, (size_t)b->B * b->scan_blocks));
  CAV_TRY(balloc(b, &b->nl, (size_t)own * b->cap));
  CAV_TRY(balloc(b, &b->nlb, (size_t)own * b->cap));
  CAV_TRY(balloc(b, &b->nn, own));
  CAV_TRY(balloc(b, &b->xqs, slots));
  CAV_TRY(balloc(b, &b->tids, slots));
  CAV_TRY(balloc(b, &b->litems, slots));
  CAV_TRY(balloc(b, &b->binpos, slots));
  CAV_TRY(balloc(b, &b->order, own));
  CAV_TRY(balloc(b, &b->ocnt, b->ncells));
  CAV_TRY(balloc(b, &b->ostart, b->ncells));
  CAV_TRY(balloc(b, &b->flags, 16 * (size_t)b->B));
  CAV_TRY(balloc(b, &b->part, (size_t)b->B * b->pair_blocks));
  CAV_TRY(balloc(b, &b->vir, b->B));
  if (b->h_flags) (void)hipHostFree(b->h_flags);
  BAT_HIP(hipHostMalloc((void **)&b->h_flags, sizeof(int) * 16 * b->B));
  BAT_HIP(hipMemset(b->flags, 0, sizeof(int) * 16 * b->B));
  BAT_HIP(hipMemset(b->xo, 0, sizeof(double) * slots));
  BAT_HIP(hipMemset(b->yo, 0, sizeof(double) * slots));
  BAT_HIP(hipMemset(b->zo, 0, sizeof(double) * slots));
  BAT_HIP(hipMemcpy(b->d_rep, b->reps.data(), sizeof(RepDev) * b->B, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->cell_rep, b->h_cell_rep.data(), sizeof(int) * b->ncells, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->st_off, b->h_st_off.data(), sizeof(int) * b->h_st_off.size(), hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->st, b->h_st.data(), sizeof(int) * b->h_st.size(), hipMemcpyHostToDevice));

  // per-type tables: 1/m and the LJ coefficient blocks [c12 | c6 | cshift]
  const int nt = b->nt, nt2 = nt * nt;
  std::vector<double> lj(3 * (size_t)std::max(nt2, 1), 0.0);
  if (b->lj_on) {
    const double rc = b->lj_rc, rcsi = 1 / (rc * rc);
    for (int p = 0; p < nt2; ++p) {
      const double e = b->eps[p], s = b->sig[p];
      const double s2 = s * s, s6 = s2 * s2 * s2, s12 = s6 * s6;
      const double rhoc2 = s * s * rcsi, rhoc6 = rhoc2 * rhoc2 * rhoc2, rhoc12 = rhoc6 * rhoc6;
      lj[p] = 48 * e * s12;
      lj[nt2 + p] = 24 * e * s6;
      lj[2 * nt2 + p] = 4 * e * (12 * rhoc12 * rcsi - 6 * rhoc6 * rcsi);
    }
  }
  CAV_TRY(balloc(b, &b->d_lj, lj.size()));
  CAV_TRY(balloc(b, &b->d_minv, nt));
  std::vector<double> re(std::max(nt2, 1), 0.0), rs(std::max(nt2, 1), 0.0);
  if (b->lj_on) {
    re = b->eps;
    rs = b->sig;
  }
  CAV_TRY(balloc(b, &b->d_raw_eps, re.size()));
  CAV_TRY(balloc(b, &b->d_raw_sig, rs.size()));
  CAV_TRY(balloc(b, &b->d_charge, nt));
  BAT_HIP(hipMemcpy(b->d_raw_eps, re.data(), sizeof(double) * re.size(), hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->d_raw_sig, rs.data(), sizeof(double) * rs.size(), hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->d_charge, b->charge.data(), sizeof(double) * nt, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->d_lj, lj.data(), sizeof(double) * lj.size(), hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->d_minv, b->minv.data(), sizeof(double) * nt, hipMemcpyHostToDevice));

  // atoms: owned slots, owned arrays, ids = local index
  std::vector<double> sx(slots, 0.0), sy(slots, 0.0), sz(slots, 0.0), sq(slots, 0.0);
  std::vector<int> st(slots, 0), sid(slots, 0);
  std::vector<double> v[3], qo(own);
  for (int k = 0; k < 3; ++k) v[k].assign(own, 0.0);
  std::vector<int> to(own);
  for (int r = 0; r < b->B; ++r) {
    const cav_batch::HostRep &h = b->hreps[r];
    const RepDev &d = b->reps[r];
    for (int i = 0; i < h.n; ++i) {
      const long s = d.base + i, o = d.obase + i;
      sx[s] = h.pos[3 * (size_t)i];
      sy[s] = h.pos[3 * (size_t)i + 1];
      sz[s] = h.pos[3 * (size_t)i + 2];
      st[s] = to[o] = h.type[i];
      sid[s] = i;
      sq[s] = qo[o] = b->charge[h.type[i]];
      for (int k = 0; k < 3; ++k) v[k][o] = h.vel[3 * (size_t)i + k];
    }
  }
  BAT_HIP(hipMemcpy(b->x, sx.data(), sizeof(double) * slots, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->y, sy.data(), sizeof(double) * slots, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->z, sz.data(), sizeof(double) * slots, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->q, sq.data(), sizeof(double) * slots, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->type, st.data(), sizeof(int) * slots, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->id, sid.data(), sizeof(int) * slots, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->vx, v[0].data(), sizeof(double) * own, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->vy, v[1].data(), sizeof(double) * own, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->vz, v[2].data(), sizeof(double) * own, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->q_o, qo.data(), sizeof(double) * own, hipMemcpyHostToDevice));
  BAT_HIP(hipMemcpy(b->type_o, to.data(), sizeof(int) * own, hipMemcpyHostToDevice));
  CAV_TRY(batch_ewald_k_alloc(b));
  return CAV_OK;
}

// forces of every replica: LJ + Ewald_r (overwrite), Ewald_k (add)
static int batch_forces(cav_batch *b)
{
  CAV_TRY(batch_pairs(b));
  CAV_TRY(batch_ewald_k(b));
  return CAV_OK;
}

static int batch_sync_flags(cav_batch *b)
{
  BAT_HIP(hipMemcpyAsync(b->h_flags, b->flags, sizeof(int) * 16 * b->B, hipMemcpyDeviceToHost, b->stream));
  BAT_HIP(hipStreamSynchronize(b->stream));
  return CAV_OK;
}

// One step of every replica (graph-capturable: no host decisions inside).
static int batch_step_body(cav_batch *b, double dt)
{
  const dim3 own_grid(grid_for(b->maxn), b->B);
  kb_snapshot<<<own_grid, kBlock, 0, b->stream>>>(b->d_rep, b->nown, 1, b->x, b->y, b->z, b->vx, b->vy,
                                                   b->vz, b->ax, b->ay, b->az, b->snap);
  kb_begin<<<grid_for(b->B), kBlock, 0, b->stream>>>(b->B, b->flags);
  kb_drift<<<own_grid, kBlock, 0, b->stream>>>(b->d_rep, dt, b->x, b->y, b->z, b->vx, b->vy, b->vz,
                                                b->ax, b->ay, b->az, b->aox, b->aoy, b->aoz, b->flags);
  CAV_TRY(batch_neighbor(b, false));
  CAV_TRY(batch_forces(b));
  kb_kick<<<grid_for(b->nown), kBlock, 0, b->stream>>>(b->nown, 0.5 * dt, b->vx, b->vy, b->vz, b->ax,
                                                        b->ay, b->az, b->aox, b->aoy, b->aoz);
  BAT_HIP(hipMemcpyAsync(b->h_flags, b->flags, sizeof(int) * 16 * b->B, hipMemcpyDeviceToHost, b->stream));
  BAT_HIP(hipGetLastError());
  return CAV_OK;
}

static void drop_graph(cav_batch *b)
{
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

static int capture_step(cav_batch *b, double dt)
{
  drop_graph(b);
  BAT_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = batch_step_body(b, dt);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc != CAV_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return bhip_check(b, e, "hipStreamEndCapture");
  b->graph = g;
  BAT_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
  b->graph_dt = dt;
  return CAV_OK;
}

// Capacity trouble after a step: grow the Verlet rows or the ghost slots of
// the replicas that overflowed.  Returns whether the state was rebuilt.
static int batch_grow(cav_batch *b, bool *relaid)
{
  *relaid = false;
  int need = 0;
  bool gover = false;
  for (int r = 0; r < b->B; ++r) {
    const int *f = b->h_flags + 16 * r;
    if (f[F_OVERFLOW]) need = std::max(need, f[F_MAXNB]);
    if (f[F_GOVER]) gover = true;
  }
  if (!gover && !need) return CAV_OK;
  // back to the state before the step, on the host
  const dim3 own_grid(grid_for(b->maxn), b->B);
  kb_snapshot<<<own_grid, kBlock, 0, b->stream>>>(b->d_rep, b->nown, 0, b->x, b->y, b->z, b->vx, b->vy,
                                                   b->vz, b->ax, b->ay, b->az, b->snap);
  std::vector<double> hx(b->nslots), hy(b->nslots), hz(b->nslots), hv[3];
  BAT_HIP(hipMemcpyAsync(hx.data(), b->x, sizeof(double) * b->nslots, hipMemcpyDeviceToHost, b->stream));
  BAT_HIP(hipMemcpyAsync(hy.data(), b->y, sizeof(double) * b->nslots, hipMemcpyDeviceToHost, b->stream));
  BAT_HIP(hipMemcpyAsync(hz.data(), b->z, sizeof(double) * b->nslots, hipMemcpyDeviceToHost, b->stream));
  double *vs[3] = {b->vx, b->vy, b->vz};
  for (int k = 0; k < 3; ++k) {
    hv[k].resize(b->nown);
    BAT_HIP(hipMemcpyAsync(hv[k].data(), vs[k], sizeof(double) * b->nown, hipMemcpyDeviceToHost, b->stream));
  }
  BAT_HIP(hipStreamSynchronize(b->stream));
  for (int r = 0; r < b->B; ++r) {
    cav_batch::HostRep &h = b->hreps[r];
    const RepDev &d = b->reps[r];
    for (int i = 0; i < h.n; ++i) {
      h.pos[3 * (size_t)i] = hx[d.base + i];
      h.pos[3 * (size_t)i + 1] = hy[d.base + i];
      h.pos[3 * (size_t)i + 2] = hz[d.base + i];
      for (int k = 0; k < 3; ++k) h.vel[3 * (size_t)i + k] = hv[k][d.obase + i];
    }
    if (b->h_flags[16 * r + F_GOVER]) h.gcap = std::min(7 * h.n, 2 * h.gcap + 64);
  }
  if (need) b->cap = (need + need / 4 + 16 + 3) / 4 * 4;
  *relaid = true;
  return CAV_OK;
}

static int batch_setup_impl(cav_batch *b)
{
  drop_graph(b);
  for (int attempt = 0; attempt < 6; ++attempt) {
    CAV_TRY(batch_layout(b));
    kb_set_all<<<grid_for(b->B), kBlock, 0, b->stream>>>(b->B, b->flags, F_NGHOST_OLD, -1);
    kb_begin<<<grid_for(b->B), kBlock, 0, b->stream>>>(b->B, b->flags);
    kb_wrap<<<dim3(grid_for(b->maxn), b->B), kBlock, 0, b->stream>>>(b->d_rep, b->x, b->y, b->z, b->flags);
    CAV_TRY(batch_neighbor(b, true));
    CAV_TRY(batch_forces(b));
    CAV_TRY(batch_sync_flags(b));
    bool grow = false;
    for (int r = 0; r < b->B; ++r) {
      const int *f = b->h_flags + 16 * r;
      if (f[F_GOVER]) {
        b->hreps[r].gcap = std::min(7 * b->hreps[r].n, 2 * b->hreps[r].gcap + 64);
        grow = true;
      }
      if (f[F_OVERFLOW]) {
        b->cap = std::max(b->cap, (f[F_MAXNB] + f[F_MAXNB] / 4 + 16 + 3) / 4 * 4);
        grow = true;
      }
    }
    if (!grow) {
      b->setup_done = true;
      return CAV_OK;
    }
  }
  return bfail(b, CAV_ERR_CAPACITY, "batch capacities could not be met");
}

static int batch_run(cav_batch *b, double dt, int nsteps, bool graph)
{
  for (int s = 0; s < nsteps; ++s) {
    bool redo = true;
    for (int attempt = 0; redo && attempt < 6; ++attempt) {
      if (graph) {
        if (!b->gexec || b->graph_dt != dt) CAV_TRY(capture_step(b, dt));
        BAT_HIP(hipGraphLaunch(b->gexec, b->stream));
      } else {
        CAV_TRY(batch_step_body(b, dt));
      }
      BAT_HIP(hipStreamSynchronize(b->stream));
      bool nan = false;
      for (int r = 0; r < b->B; ++r) nan |= b->h_flags[16 * r + F_NAN] != 0;
      if (nan) {
        const dim3 own_grid(grid_for(b->maxn), b->B);
        kb_snapshot<<<own_grid, kBlock, 0, b->stream>>>(b->d_rep, b->nown, 0, b->x, b->y, b->z, b->vx,
                                                         b->vy, b->vz, b->ax, b->ay, b->az, b->snap);
        BAT_HIP(hipStreamSynchronize(b->stream));
        return bfail(b, CAV_ERR_NAN, "NaN acceleration before the drift");
      }
      bool relaid = false;
      CAV_TRY(batch_grow(b, &relaid));
      redo = relaid;
      if (relaid) CAV_TRY(batch_setup_impl(b));  // rebuilds from the restored state
    }
    if (redo) return bfail(b, CAV_ERR_CAPACITY, "batch capacities could not be met");
  }
  return CAV_OK;
}

// ---------------------------------------------------------------- autotune --

static float time_ewald_k(cav_batch *b, int reps)
{
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int i = 0; i < 2; ++i) (void)batch_ewald_k(b);
  (void)hipEventRecord(e0, b->stream);
  for (int i = 0; i < reps; ++i) (void)batch_ewald_k(b);
  (void)hipEventRecord(e1, b->stream);
  (void)hipEventSynchronize(e1);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

extern "C" {

int cav_batch_create(int device, int precision, cav_batch **out)
{
  if (!out) return CAV_ERR_ARG;
  *out = nullptr;
  if (precision != CAV_FP64 && precision != CAV_MIXED) return CAV_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return CAV_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return CAV_ERR_HIP;
  cav_batch *b = new cav_batch;
  b->device = device;
  b->precision = precision;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return CAV_ERR_HIP;
  }
  *out = b;
  return CAV_OK;
}

int cav_batch_destroy(cav_batch *b)
{
  if (!b) return CAV_ERR_ARG;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  hipStream_t st = b->stream;
  batch_free(b);
  (void)hipStreamDestroy(st);
  delete b;
  return CAV_OK;
}

const char *cav_batch_last_error(const cav_batch *b) { return b ? b->err.c_str() : "null batch"; }

int cav_batch_set_types(cav_batch *b, int nt, const double *mass, const double *charge)
{
  if (!b || nt <= 0 || !mass || !charge) return CAV_ERR_ARG;
  b->nt = nt;
  b->mass.assign(mass, mass + nt);
  b->charge.assign(charge, charge + nt);
  b->minv.resize(nt);
  for (int t = 0; t < nt; ++t) b->minv[t] = mass[t] == 0.0 ? 0.0 : 1.0 / mass[t];
  b->setup_done = false;
  return CAV_OK;
}

int cav_batch_set_neighbor(cav_batch *b, double ghost_cutoff, double cutoff, double cutoff_extra_coef)
{
  if (!b || !(cutoff > 0) || cutoff_extra_coef < 1.0 || ghost_cutoff < 0) return CAV_ERR_ARG;
  b->ghost_cutoff = ghost_cutoff;
  b->nl_cutoff = cutoff;
  b->nl_coef = cutoff_extra_coef;
  b->nl_on = true;
  b->setup_done = false;
  return CAV_OK;
}

int cav_batch_set_lj(cav_batch *b, double cutoff, const double *epsilon, const double *sigma)
{
  if (!b || b->nt <= 0 || cutoff < 0 || !epsilon || !sigma) return CAV_ERR_ARG;
  b->lj_on = true;
  b->lj_rc = cutoff;
  b->eps.assign(epsilon, epsilon + b->nt * b->nt);
  b->sig.assign(sigma, sigma + b->nt * b->nt);
  b->setup_done = false;
  return CAV_OK;
}

int cav_batch_set_ewald_r(cav_batch *b, double alpha, double k_electrostatic)
{
  if (!b || alpha < 0 || k_electrostatic < 0) return CAV_ERR_ARG;
  b->er_on = true;
  b->er_alpha = alpha;
  b->er_ke = k_electrostatic;
  b->setup_done = false;
  return CAV_OK;
}

int cav_batch_set_ewald_k(cav_batch *b, double alpha, double k_electrostatic, int kx_max, int ky_max,
                          int kz_max)
{
  if (!b || !(alpha > 0) || k_electrostatic < 0 || kx_max < 0 || ky_max < 0 || kz_max < 0)
    return CAV_ERR_ARG;
  b->ek_on = true;
  b->ek_alpha = alpha;
  b->ek_ke = k_electrostatic;
  b->kx = kx_max;
  b->ky = ky_max;
  b->kz = kz_max;
  b->setup_done = false;
  return CAV_OK;
}

int cav_batch_add_replica(cav_batch *b, const double lower[3], const double upper[3], int n,
                          const int *type, const double *pos, const double *vel, int *index)
{
  if (!b || !lower || !upper || n <= 0 || !type || !pos) return CAV_ERR_ARG;
  if (b->nt <= 0) return bfail(b, CAV_ERR_STATE, "set the types before adding replicas");
  cav_batch::HostRep h;
  for (int k = 0; k < 3; ++k) {
    h.lo[k] = lower[k];
    h.hi[k] = upper[k];
    if (!(upper[k] > lower[k])) return bfail(b, CAV_ERR_ARG, "replica box: upper must exceed lower");
  }
  h.n = n;
  h.type.assign(type, type + n);
  for (int t : h.type)
    if (t < 0 || t >= b->nt) return bfail(b, CAV_ERR_ARG, "atom type out of range");
  h.pos.assign(pos, pos + 3 * (size_t)n);
  h.vel.assign(3 * (size_t)n, 0.0);
  if (vel) h.vel.assign(vel, vel + 3 * (size_t)n);
  if (index) *index = (int)b->hreps.size();
  b->hreps.push_back(std::move(h));
  b->setup_done = false;
  return CAV_OK;
}

int cav_batch_num_replicas(const cav_batch *b, int *n)
{
  if (!b || !n) return CAV_ERR_ARG;
  *n = (int)b->hreps.size();
  return CAV_OK;
}

int cav_batch_setup(cav_batch *b)
{
  if (!b) return CAV_ERR_ARG;
  if (b->hreps.empty()) return bfail(b, CAV_ERR_STATE, "no replicas");
  if (!b->nl_on) return bfail(b, CAV_ERR_STATE, "set the neighbour list before setup");
  (void)hipSetDevice(b->device);
  return batch_setup_impl(b);
}

int cav_batch_step(cav_batch *b, double dt, int nsteps, int use_graph)
{
  if (!b) return CAV_ERR_ARG;
  if (!b->setup_done) return bfail(b, CAV_ERR_STATE, "cav_batch_setup must run first");
  (void)hipSetDevice(b->device);
  return batch_run(b, dt, nsteps, use_graph != 0);
}

int cav_batch_download(cav_batch *b, int replica, int which, double *out)
{
  if (!b || !out || replica < 0 || replica >= b->B) return CAV_ERR_ARG;
  if (!b->setup_done) return bfail(b, CAV_ERR_STATE, "cav_batch_setup must run first");
  const RepDev &d = b->reps[replica];
  std::vector<double> a[3];
  double *src[3];
  bool slot = false;
  switch (which) {
  case CAV_POS: src[0] = b->x; src[1] = b->y; src[2] = b->z; slot = true; break;
  case CAV_VEL: src[0] = b->vx; src[1] = b->vy; src[2] = b->vz; break;
  case CAV_ACC: src[0] = b->ax; src[1] = b->ay; src[2] = b->az; break;
  case CAV_ACC_OLD: src[0] = b->aox; src[1] = b->aoy; src[2] = b->aoz; break;
  default: return CAV_ERR_ARG;
  }
  const long off = slot ? d.base : d.obase;
  for (int k = 0; k < 3; ++k) {
    a[k].resize(d.n);
    BAT_HIP(hipMemcpyAsync(a[k].data(), src[k] + off, sizeof(double) * d.n, hipMemcpyDeviceToHost, b->stream));
  }
  BAT_HIP(hipStreamSynchronize(b->stream));
  for (int i = 0; i < d.n; ++i)
    for (int k = 0; k < 3; ++k) out[3 * (size_t)i + k] = a[k][i];
  return CAV_OK;
}

int cav_batch_virial(cav_batch *b, int replica, double *v)
{
  if (!b || !v || replica < 0 || replica >= b->B) return CAV_ERR_ARG;
  if (!b->setup_done) return bfail(b, CAV_ERR_STATE, "cav_batch_setup must run first");
  BAT_HIP(hipMemcpyAsync(v, b->vir + replica, sizeof(double), hipMemcpyDeviceToHost, b->stream));
  BAT_HIP(hipStreamSynchronize(b->stream));
  return CAV_OK;
}

// The replica's Verlet pairs with the reference's half-list membership
// (owned j > i, every ghost as n + g), as cav_get_pairs.
int cav_batch_get_pairs(cav_batch *b, int replica, int *ij, long capacity, long *npairs)
{
  if (!b || !npairs || replica < 0 || replica >= b->B) return CAV_ERR_ARG;
  if (!b->setup_done) return bfail(b, CAV_ERR_STATE, "cav_batch_setup must run first");
  const RepDev &d = b->reps[replica];
  std::vector<int> nl((size_t)d.n * b->cap), nn(d.n);
  BAT_HIP(hipMemcpyAsync(nl.data(), b->nl + (size_t)d.obase * b->cap, sizeof(int) * nl.size(),
                         hipMemcpyDeviceToHost, b->stream));
  BAT_HIP(hipMemcpyAsync(nn.data(), b->nn + d.obase, sizeof(int) * d.n, hipMemcpyDeviceToHost, b->stream));
  BAT_HIP(hipStreamSynchronize(b->stream));
  long k = 0;
  for (int i = 0; i < d.n; ++i)
    for (int e = 0; e < nn[i]; ++e) {
      const int j = nl[(size_t)i * b->cap + e];
      if (j < d.n && j <= i) continue;
      if (ij && k < capacity) {
        ij[2 * k] = i;
        ij[2 * k + 1] = j;
      }
      ++k;
    }
  *npairs = k;
  return CAV_OK;
}

// Autotuner: times the batched Ewald_k passes for each candidate tile shape
// (split form: row tiles per S-pass wave, atoms per S chunk, force-pass atom
// tiles per wave; plain form: atoms per S chunk) on the batch itself and keeps
// the fastest.  Results: cav_batch_tiles.
int cav_batch_autotune(cav_batch *b)
{
  if (!b) return CAV_ERR_ARG;
  if (!b->setup_done) return bfail(b, CAV_ERR_STATE, "cav_batch_setup must run first");
  if (!b->ek_on) return CAV_OK;
  (void)hipSetDevice(b->device);
  struct Cand { int rt, stage, chunk, fnt; };
  std::vector<Cand> cands;
  // the split S pass stages one 32-atom table block at a time
  const int rts[3] = {1, 2, 3}, chunks[3] = {256, 512, 1024};
  if (b->precision == CAV_MIXED) {
    for (int rt : rts)
      for (int ch : chunks)
        for (int nt : {1, 2}) cands.push_back({rt, 32, ch, nt});
  } else {
    for (int ch : chunks) cands.push_back({b->s_rt, b->s_stage, ch, b->f_nt});
  }
  float best = 1e30f;
  Cand pick = {b->s_rt, b->s_stage, b->s_chunk, b->f_nt};
  for (const Cand &c : cands) {
    b->s_rt = c.rt;
    b->s_stage = c.stage;
    b->s_chunk = c.chunk;
    b->f_nt = c.fnt;
    CAV_TRY(batch_ewald_k_alloc(b));
    const float ms = time_ewald_k(b, 5);
    BAT_HIP(hipGetLastError());
    if (ms < best) {
      best = ms;
      pick = c;
    }
  }
  b->s_rt = pick.rt;
  b->s_stage = pick.stage;
  b->s_chunk = pick.chunk;
  b->f_nt = pick.fnt;
  b->tuned = true;
  drop_graph(b);
  // rebuild buffers for the chosen tiles and restore the forces of the current state
  CAV_TRY(batch_ewald_k_alloc(b));
  CAV_TRY(batch_forces(b));
  BAT_HIP(hipStreamSynchronize(b->stream));
  return CAV_OK;
}

int cav_batch_tiles(const cav_batch *b, int *s_rows, int *s_stage, int *s_chunk, int *f_tiles)
{
  if (!b) return CAV_ERR_ARG;
  if (s_rows) *s_rows = b->s_rt;
  if (s_stage) *s_stage = b->s_stage;
  if (s_chunk) *s_chunk = b->s_chunk;
  if (f_tiles) *f_tiles = b->f_nt;
  return CAV_OK;
}

int cav_batch_num_atoms(const cav_batch *b, int replica, int *n)
{
  if (!b || !n || replica < 0 || replica >= (int)b->hreps.size()) return CAV_ERR_ARG;
  *n = b->hreps[replica].n;
  return CAV_OK;
}

}  // extern "C"
