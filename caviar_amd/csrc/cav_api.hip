// C ABI of caviar_amd: context lifetime, settings, upload/download, and the
// velocity-Verlet step driver (all device work on the context's one stream).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cav_internal.hpp"

namespace cav {

int fail(cav_ctx *c, int code, const std::string &msg)
{
  if (c) c->err = msg;
  return code;
}

int hip_check(cav_ctx *c, hipError_t e, const char *what)
{
  return fail(c, CAV_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void timer_begin(cav_ctx *c, Timer &t)
{
  if (!c->timing) return;
  if (!t.a) {
    (void)hipEventCreate(&t.a);
    (void)hipEventCreate(&t.b);
  }
  (void)hipEventRecord(t.a, c->stream);
}

void timer_end(cav_ctx *c, Timer &t)
{
  if (!c->timing) return;
  (void)hipEventRecord(t.b, c->stream);
  t.pending = true;
}

static void timer_harvest(Timer &t)
{
  if (!t.pending) return;
  float ms = 0;
  if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
    t.total_ms += ms;
    t.launches += 1;
  }
  t.pending = false;
}

// ------------------------------------------------------- integration kernels --

// integrate_velocity_verlet drift (md_simulator.cpp:584-594) with the PBC wrap
// of exchange_owned fused in (domain.cpp:418-468).  Every operation is
// rounded as the reference's expression is: r += (v*dt + ((0.5*a)*dt)*dt).
// snap (or nullptr): the step snapshot of the owned positions, velocities and
// accelerations before the drift ([x | y | z | vx | vy | vz | ax | ay | az],
// n each), written from the values the drift reads anyway.
__global__ void k_drift(int n, double dt, const double *lo, const double *hi, const double *L,
                        int bx, int by, int bz, double *x, double *y, double *z, const double *vx,
                        const double *vy, const double *vz, const double *ax, const double *ay,
                        const double *az, double *aox, double *aoy, double *aoz, int *flags,
                        double *snap)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double a[3] = {ax[i], ay[i], az[i]};
  if (isnan(a[0]) || isnan(a[1]) || isnan(a[2])) flags[F_NAN] = 1;
  const double v[3] = {vx[i], vy[i], vz[i]};
  double *p[3] = {x + i, y + i, z + i};
  if (snap) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      snap[(long)d * n + i] = *p[d];
      snap[(long)(3 + d) * n + i] = v[d];
      snap[(long)(6 + d) * n + i] = a[d];
    }
  }
  const int b[3] = {bx, by, bz};
  bool wrapped = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double h = xmul(xmul(xmul(0.5, a[d]), dt), dt);
    double r = xadd(*p[d], xadd(xmul(v[d], dt), h));
    if (b[d] == 1) {
      if (r < lo[d]) {
        r = xadd(r, L[d]);
        wrapped = true;
      } else if (r > hi[d]) {
        r = xsub(r, L[d]);
        wrapped = true;
      }
    }
    *p[d] = r;
  }
  aox[i] = a[0];
  aoy[i] = a[1];
  aoz[i] = a[2];
  if (wrapped) {
    flags[F_WRAPPED] = 1;
    flags[F_REBUILD] = 1;
  }
}

// exchange_owned_single_md_domain alone (setup and recalculate).
__global__ void k_wrap(int n, const double *lo, const double *hi, const double *L, int bx, int by,
                       int bz, double *x, double *y, double *z, int *flags)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double *p[3] = {x + i, y + i, z + i};
  const int b[3] = {bx, by, bz};
  bool wrapped = false;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (b[d] != 1) continue;
    double r = *p[d];
    if (r < lo[d]) {
      r = xadd(r, L[d]);
      wrapped = true;
    } else if (r > hi[d]) {
      r = xsub(r, L[d]);
      wrapped = true;
    }
    *p[d] = r;
  }
  if (wrapped) {
    flags[F_WRAPPED] = 1;
    flags[F_REBUILD] = 1;
  }
}

// v += (0.5*dt) * (a + a_old)  (md_simulator.cpp:610-617)
__global__ void k_kick(int n, double hdt, double *vx, double *vy, double *vz, const double *ax,
                       const double *ay, const double *az, const double *aox, const double *aoy,
                       const double *aoz)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  vx[i] = xadd(vx[i], xmul(hdt, xadd(ax[i], aox[i])));
  vy[i] = xadd(vy[i], xmul(hdt, xadd(ay[i], aoy[i])));
  vz[i] = xadd(vz[i], xmul(hdt, xadd(az[i], aoz[i])));
}

// The kick after a concurrent Ewald_k: first a += ek (the k-space
// accelerations, the single addition the in-order schedule makes), then the kick.
__global__ void k_kick_join(int n, double hdt, double *vx, double *vy, double *vz, double *ax,
                            double *ay, double *az, const double *ekx, const double *eky,
                            const double *ekz, const double *aox, const double *aoy,
                            const double *aoz)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double a[3] = {xadd(ax[i], ekx[i]), xadd(ay[i], eky[i]), xadd(az[i], ekz[i])};
  ax[i] = a[0];
  ay[i] = a[1];
  az[i] = a[2];
  vx[i] = xadd(vx[i], xmul(hdt, xadd(a[0], aox[i])));
  vy[i] = xadd(vy[i], xmul(hdt, xadd(a[1], aoy[i])));
  vz[i] = xadd(vz[i], xmul(hdt, xadd(a[2], aoz[i])));
}

// a += e per component, one rounded addition (the concurrent Ewald_k's join).
__global__ void k_add3(int n, double *ax, double *ay, double *az, const double *ex,
                       const double *ey, const double *ez)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  ax[i] = xadd(ax[i], ex[i]);
  ay[i] = xadd(ay[i], ey[i]);
  az[i] = xadd(az[i], ez[i]);
}

__global__ void k_begin_step(int *flags)
{
  flags[F_WRAPPED] = 0;
  flags[F_REBUILD] = flags[F_FORCE];
  flags[F_FORCE] = 0;
  flags[F_NAN] = 0;
  flags[F_OVERFLOW] = 0;
  flags[F_MAXNB] = 0;
}

__global__ void k_set_flag(int *flags, int idx, int v) { flags[idx] = v; }

__global__ void k_set_scalar(double *p, double v) { *p = v; }

__global__ void k_zero3(int n, double *a, double *b, double *c)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] = b[i] = c[i] = 0.0;
}

__global__ void k_charge_of_type(int n, const int *type, const double *charge, double *q)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) q[i] = charge[type[i]];
}

// The same over the ghost slots n .. n + flags[F_NGHOST) of the last exchange
// (their types are copies of their sources'), bounded by the slot capacity.
__global__ void k_ghost_charge_of_type(int n, int gcap, const int *flags, const int *type,
                                       const double *charge, double *q)
{
  const int g = blockIdx.x * kBlock + threadIdx.x;
  const int ng = min(flags[F_NGHOST], gcap);
  if (g < ng) q[n + g] = charge[type[n + g]];
}

}  // namespace cav

using namespace cav;

// --------------------------------------------------------------- helpers --

template <typename X>
static int dmalloc(cav_ctx *c, X **p, size_t count)
{
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  CAV_HIP(hipMalloc((void **)p, sizeof(X) * (count ? count : 1)));
  return CAV_OK;
}

static void free_all(cav_ctx *c)
{
  void *ptrs[] = {c->x,       c->y,      c->z,         c->xo,       c->yo,       c->zo,
                  c->vx,      c->vy,     c->vz,        c->ax,       c->ay,       c->az,
                  c->aox,     c->aoy,    c->aoz,       c->snap,     c->type,     c->id,
                  c->mol,     c->q,      c->gsrc,      c->flags,    c->blk,      c->bin_of,
                  c->bin_cnt, c->bin_start, c->bin_items, c->d_st_off, c->d_st,  c->scan_tmp,
                  c->d_xst_off, c->d_xst,
                  c->nl,      c->nn,     c->d_minv,    c->d_charge, c->d_eps,    c->d_lambda,
                  c->d_ljl,   c->d_sr_lambda, c->d_elc_k, c->d_elc_chi, c->d_elc_dip,
                  c->d_thermo, c->d_mass,
                  c->part,    c->scal,   c->tab,       c->sk_part,  c->sk,       c->bmat,
                  c->d_W,     c->kacc,      c->order,    c->ord_blk,  c->xq,
                  c->tid,     c->d_raw_eps, c->d_raw_sig, c->ztab, c->bmat_h, c->cpos, c->sk_sum,
                  c->binpos,  c->xqs,    c->tids,       c->nll,    c->nlb,    c->lstart, c->litems,
                  c->ekx,     c->eky,    c->ekz,        c->part_k};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_flags) (void)hipHostFree(c->h_flags);
}

// box arrays on the device (lo, hi, L) live in the scalar buffer tail
static const int kBoxSlot = 16;

static int upload_params(cav_ctx *c)
{
  const int nt = c->nt, nt2 = nt * nt;
  CAV_TRY(dmalloc(c, &c->d_minv, nt));
  CAV_TRY(dmalloc(c, &c->d_charge, nt));
  CAV_TRY(dmalloc(c, &c->d_mass, nt));
  CAV_HIP(hipMemcpyAsync(c->d_minv, c->minv.data(), sizeof(double) * nt, hipMemcpyHostToDevice, c->stream));
  CAV_HIP(hipMemcpyAsync(c->d_mass, c->mass.data(), sizeof(double) * nt, hipMemcpyHostToDevice, c->stream));
  CAV_HIP(hipMemcpyAsync(c->d_charge, c->charge.data(), sizeof(double) * nt, hipMemcpyHostToDevice, c->stream));
  // LJ table blocks: [c12 | c6 | cshift | u12 | u6 | uc], see the pair kernels
  std::vector<double> lj(6 * (size_t)std::max(nt2, 1), 0.0);
  if (c->lj_on) {
    const double rc = c->lj_rc, rcsi = 1 / (rc * rc);
    for (int p = 0; p < nt2; ++p) {
      const double e = c->eps[p], s = c->sig[p];
      const double s2 = s * s, s6 = s2 * s2 * s2, s12 = s6 * s6;
      // lj_cell_list.cpp:145-155: 4e(-12 rho^12/r^2 + 6 rho^6/r^2 + 12 rho_c^12/rc^2 - 6 rho_c^6/rc^2)
      const double rhoc2 = s * s * rcsi, rhoc6 = rhoc2 * rhoc2 * rhoc2, rhoc12 = rhoc6 * rhoc6;
      lj[p] = 48 * e * s12;
      lj[nt2 + p] = 24 * e * s6;
      lj[2 * nt2 + p] = 4 * e * (12 * rhoc12 * rcsi - 6 * rhoc6 * rcsi);
      lj[3 * nt2 + p] = 4 * e * s12;
      lj[4 * nt2 + p] = 4 * e * s6;
      const double rc2 = rc * rc, ri6 = 1 / (rc2 * rc2 * rc2);
      lj[5 * nt2 + p] = 4 * e * s12 * ri6 * ri6 - 4 * e * s6 * ri6 + 0.5 * lj[2 * nt2 + p] * rc2;
    }
  }
  std::vector<double> re(std::max(nt2, 1), 0.0), rs(std::max(nt2, 1), 0.0);
  if (c->lj_on) {
    re = c->eps;
    rs = c->sig;
  }
  CAV_TRY(dmalloc(c, &c->d_raw_eps, re.size()));
  CAV_TRY(dmalloc(c, &c->d_raw_sig, rs.size()));
  CAV_HIP(hipMemcpyAsync(c->d_raw_eps, re.data(), sizeof(double) * re.size(), hipMemcpyHostToDevice, c->stream));
  CAV_HIP(hipMemcpyAsync(c->d_raw_sig, rs.data(), sizeof(double) * rs.size(), hipMemcpyHostToDevice, c->stream));
  CAV_TRY(dmalloc(c, &c->d_eps, lj.size()));
  CAV_HIP(hipMemcpyAsync(c->d_eps, lj.data(), sizeof(double) * lj.size(), hipMemcpyHostToDevice, c->stream));
  std::vector<double> lam(std::max(nt2, 1), 1.0);
  if (c->lambda_set) lam = c->lambda;
  CAV_TRY(dmalloc(c, &c->d_lambda, lam.size()));
  CAV_HIP(hipMemcpyAsync(c->d_lambda, lam.data(), sizeof(double) * lam.size(), hipMemcpyHostToDevice, c->stream));
  // force_field lj: coef(r) = -c12 r^-14 + c6 r^-8 + cshift with the shift at
  // the pair's own cutoff c (lj.cpp:421-434); rc2 = c, or -1 where eps = 0
  // (lj.cpp:416-418 skips those pairs)
  std::vector<double> ljl(4 * (size_t)std::max(nt2, 1), 0.0);
  if (c->ljl_on) {
    for (int p = 0; p < nt2; ++p) {
      const double e = c->ljl_eps[p], s = c->ljl_sig[p];
      const double cc = c->ljl_cut_on ? c->ljl_cut[p] * c->ljl_cut[p] : c->ljl_rc * c->ljl_rc;
      if (e == 0) {
        ljl[3 * nt2 + p] = -1.0;
        continue;
      }
      const double s2 = s * s, s6 = s2 * s2 * s2, s12 = s6 * s6;
      const double rcsi = 1 / cc;
      const double rhoc2 = s * s * rcsi, rhoc6 = rhoc2 * rhoc2 * rhoc2, rhoc12 = rhoc6 * rhoc6;
      ljl[p] = 48 * e * s12;
      ljl[nt2 + p] = 24 * e * s6;
      ljl[2 * nt2 + p] = 4 * e * (12 * rhoc12 * rcsi - 6 * rhoc6 * rcsi);
      ljl[3 * nt2 + p] = cc;
    }
  }
  CAV_TRY(dmalloc(c, &c->d_ljl, ljl.size()));
  CAV_HIP(hipMemcpyAsync(c->d_ljl, ljl.data(), sizeof(double) * ljl.size(), hipMemcpyHostToDevice, c->stream));
  std::vector<double> srl(std::max(nt2, 1), 1.0);
  if (c->sr_lambda_set) srl = c->sr_lambda;
  CAV_TRY(dmalloc(c, &c->d_sr_lambda, srl.size()));
  CAV_HIP(hipMemcpyAsync(c->d_sr_lambda, srl.data(), sizeof(double) * srl.size(), hipMemcpyHostToDevice, c->stream));
  return CAV_OK;
}

// Neighborlist::init + make_neigh_bin (neighborlist.cpp:313-406,581-697).
static int neighbor_init(cav_ctx *c)
{
  if (!c->nl_on) return CAV_OK;
  if (c->nl_cutoff <= 0.0) return fail(c, CAV_ERR_ARG, "neighborlist cutoff has to be positive");
  if (c->nl_coef < 1.0) return fail(c, CAV_ERR_ARG, "cutoff_extra_coef must be >= 1");
  c->rce = c->nl_cutoff * c->nl_coef;
  c->rce2 = c->rce * c->rce;
  c->thr_sq = 0.25 * (c->rce - c->nl_cutoff) * (c->rce - c->nl_cutoff);
  c->cinv = 1.0 / c->rce;
  long nbins = 1;
  for (int d = 0; d < 3; ++d) {
    c->nb[d] = (int)std::ceil(c->L[d] * c->cinv);
    if (c->nb[d] < 1) c->nb[d] = 1;
    nbins *= c->nb[d];
  }
  if (nbins > (1L << 26)) return fail(c, CAV_ERR_ARG, "cell grid too large");
  c->nbins = (int)nbins;
  const int nbx = c->nb[0], nby = c->nb[1], nbz = c->nb[2];
  std::vector<int> cnt(c->nbins, 0), cand(27 * (size_t)c->nbins);
  for (int x = 0; x < nbx; ++x)
    for (int y = 0; y < nby; ++y)
      for (int z = 0; z < nbz; ++z) {
        int lo[3] = {x - 1, y - 1, z - 1}, hi[3] = {x + 1, y + 1, z + 1};
        for (int d = 0; d < 3; ++d)
          if (c->bc[d] == 0) {
            if (lo[d] == -1) lo[d] = 0;
            if (hi[d] == c->nb[d]) hi[d] = c->nb[d] - 1;
          }
        const int k = x + nbx * y + nbx * nby * z;
        int *cc = cand.data() + 27 * (size_t)k;
        for (int a = lo[0]; a <= hi[0]; ++a)
          for (int b = lo[1]; b <= hi[1]; ++b)
            for (int e = lo[2]; e <= hi[2]; ++e) {
              int v[3] = {a, b, e};
              for (int d = 0; d < 3; ++d) {
                if (v[d] == c->nb[d]) v[d] = 0;
                if (v[d] == -1) v[d] = c->nb[d] - 1;
              }
              const int lin = v[0] + nbx * v[1] + nbx * nby * v[2];
              bool dup = false;
              for (int q = 0; q < cnt[k]; ++q) dup |= cc[q] == lin;
              if (!dup) cc[cnt[k]++] = lin;
            }
      }
  c->st_off.assign(c->nbins + 1, 0);
  c->st.clear();
  for (int k = 0; k < c->nbins; ++k) {
    c->st_off[k] = (int)c->st.size();
    c->st.insert(c->st.end(), cand.begin() + 27 * (size_t)k, cand.begin() + 27 * (size_t)k + cnt[k]);
  }
  c->st_off[c->nbins] = (int)c->st.size();
  CAV_TRY(dmalloc(c, &c->d_st_off, c->st_off.size()));
  CAV_TRY(dmalloc(c, &c->d_st, c->st.size()));
  CAV_HIP(hipMemcpyAsync(c->d_st_off, c->st_off.data(), sizeof(int) * c->st_off.size(), hipMemcpyHostToDevice, c->stream));
  CAV_HIP(hipMemcpyAsync(c->d_st, c->st.data(), sizeof(int) * c->st.size(), hipMemcpyHostToDevice, c->stream));
  // the list build's grid: one more layer of cells past every face, where the
  // ghosts lie (the periodic images, so no stencil wraps) and where positions
  // outside a non-periodic box are clamped; each cell's stencil is its 3x3x3
  // neighbourhood in the grid
  const int gx = nbx + 2, gy = nby + 2, gz = nbz + 2;
  const long xnbins = (long)gx * gy * gz;
  if (xnbins > (1L << 26)) return fail(c, CAV_ERR_ARG, "cell grid too large");
  c->xnbins = (int)xnbins;
  std::vector<int> xoff(xnbins + 1), xst;
  xst.reserve(27 * (size_t)xnbins);
  for (long k = 0; k < xnbins; ++k) {
    const int cx = (int)(k % gx), cy = (int)((k / gx) % gy), cz = (int)(k / ((long)gx * gy));
    xoff[k] = (int)xst.size();
    for (int a = cx - 1; a <= cx + 1; ++a)
      for (int b = cy - 1; b <= cy + 1; ++b)
        for (int e = cz - 1; e <= cz + 1; ++e)
          if (a >= 0 && a < gx && b >= 0 && b < gy && e >= 0 && e < gz)
            xst.push_back(a + gx * b + gx * gy * e);
  }
  xoff[xnbins] = (int)xst.size();
  CAV_TRY(dmalloc(c, &c->d_xst_off, xoff.size()));
  CAV_TRY(dmalloc(c, &c->d_xst, xst.size()));
  CAV_HIP(hipMemcpy(c->d_xst_off, xoff.data(), sizeof(int) * xoff.size(), hipMemcpyHostToDevice));
  CAV_HIP(hipMemcpy(c->d_xst, xst.data(), sizeof(int) * xst.size(), hipMemcpyHostToDevice));
  const int bins = std::max(c->nbins, c->xnbins);
  CAV_TRY(dmalloc(c, &c->bin_cnt, bins));
  CAV_TRY(dmalloc(c, &c->bin_start, bins + 1));
  CAV_TRY(dmalloc(c, &c->lstart, bins + 1));
  CAV_TRY(dmalloc(c, &c->scan_tmp, grid_for(bins) + 1));
  return CAV_OK;
}

static int alloc_list(cav_ctx *c, int cap)
{
  c->nl_cap = cap;
  CAV_TRY(dmalloc(c, &c->nl, (size_t)cap * c->n));
  CAV_TRY(dmalloc(c, &c->nll, (size_t)cap * c->n));
  CAV_TRY(dmalloc(c, &c->nlb, (size_t)cap * c->n));
  return CAV_OK;
}

static int sync_flags(cav_ctx *c)
{
  CAV_HIP(hipMemcpyAsync(c->h_flags, c->flags, sizeof(int) * F_COUNT, hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  // a conditional list build that ran decides whether the staging positions
  // match the current list
  if (c->h_flags[F_REBUILD]) {
    c->nll_built = c->nll_pending;
    c->nlb_built = c->nlb_pending;
  }
  return CAV_OK;
}

static int box_dev(cav_ctx *c, const double **lo, const double **hi, const double **L)
{
  *lo = c->scal + kBoxSlot;
  *hi = c->scal + kBoxSlot + 3;
  *L = c->scal + kBoxSlot + 6;
  return CAV_OK;
}

// New parameter values (same structure) set after setup: upload them, refresh
// the per-atom charges of owned atoms and of the current ghosts (a force call
// without a new exchange reads the ghosts' charges: the reference reads
// charge[type_j] for ghosts too), and the k-space weights.
int cav::refresh_params(cav_ctx *c)
{
  if (!c->params_dirty) return CAV_OK;
  CAV_TRY(upload_params(c));
  k_charge_of_type<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->type, c->d_charge, c->q);
  if (c->gcap > 0)
    k_ghost_charge_of_type<<<grid_for(c->gcap), kBlock, 0, c->stream>>>(c->n, c->gcap, c->flags,
                                                                          c->type, c->d_charge, c->q);
  CAV_HIP(hipGetLastError());
  if (c->ek_on) {
    CAV_TRY(ewald_k_plan(c));
    CAV_HIP(hipMemcpyAsync(c->d_W, c->ek.W.data(), sizeof(double) * c->ek.W.size(),
                           hipMemcpyHostToDevice, c->stream));
  }
  c->params_dirty = false;
  return CAV_OK;
}

// Starts the step's Ewald_k on the second stream once the drift has run: the
// passes read only the owned positions and charges, and write their own
// accelerations (zeroed first) and partial sums, so they run concurrently
// with the ghost exchange, neighbour build and pair kernel of the main stream.
static int ewald_k_fork(cav_ctx *c)
{
  CAV_HIP(hipEventRecord(c->ev_fork, c->stream));
  CAV_HIP(hipStreamWaitEvent(c->kstream, c->ev_fork, 0));
  struct Side {  // the Ewald_k launchers see the side stream and buffers
    cav_ctx *c;
    explicit Side(cav_ctx *c) : c(c) { flip(); }
    ~Side() { flip(); }
    void flip()
    {
      std::swap(c->stream, c->kstream);
      std::swap(c->ax, c->ekx);
      std::swap(c->ay, c->eky);
      std::swap(c->az, c->ekz);
      std::swap(c->part, c->part_k);
    }
  } side(c);
  k_zero3<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->ax, c->ay, c->az);
  CAV_HIP(hipGetLastError());
  c->defer_dipole = true;
  const int st = launch_ewald_k(c);
  c->defer_dipole = false;
  CAV_TRY(st);
  CAV_HIP(hipEventRecord(c->ev_join, c->stream));
  return CAV_OK;
}

// The main stream waits for the side Ewald_k; the kick then adds its
// accelerations (k_kick_join): per atom the short-range sum plus the k-space
// one, as in forces().
static int ewald_k_join(cav_ctx *c)
{
  CAV_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
  return CAV_OK;
}

// The force phase: reset acceleration + every enabled field (md_simulator.cpp:370-373).
// `kspace` false leaves Ewald_k to the concurrent side stream.
static int forces(cav_ctx *c, bool kspace = true);
static int neighbor_phase(cav_ctx *c, bool forced);

// The autotuner's unit of work: the Ewald_k passes as cav_step runs them,
// i.e. on the side stream concurrently with a list build and the pair kernel
// (in order when the overlap is off).  Accelerations are left inconsistent:
// the caller recomputes them.
int cav::tuning_phase(cav_ctx *c)
{
  if (!(c->overlap && c->n > 0)) return launch_ewald_k(c);
  CAV_TRY(ewald_k_fork(c));
  CAV_TRY(neighbor_phase(c, true));
  CAV_TRY(forces(c, false));
  return ewald_k_join(c);
}
int forces_all(cav_ctx *c) { return forces(c); }
static int forces(cav_ctx *c, bool kspace)
{
  const bool pair = (c->lj_on && !c->lj_cellwalk) || c->er_on;
  if (pair)
    CAV_TRY(launch_pairs(c, true, true, false));
  else
    CAV_TRY(launch_zero_acc(c));
  if (c->lj_on && c->lj_cellwalk) CAV_TRY(launch_lj_cellwalk(c));
  if (c->ljl_on || c->sr_on) CAV_TRY(launch_pairs_list(c, true, true, true, !pair));
  if (c->ek_on && kspace) CAV_TRY(launch_ewald_k(c));
  if (c->elc_on && kspace) CAV_TRY(launch_elc(c));
  return CAV_OK;
}

// Neighbour phase for one step: ghosts already made; rebuild decision, cells,
// list, old positions.  `forced` requests an unconditional rebuild.
static int neighbor_phase(cav_ctx *c, bool forced)
{
  if (!c->nl_on) return CAV_OK;
  if (forced) {
    k_set_flag<<<1, 1, 0, c->stream>>>(c->flags, F_REBUILD, 1);
  } else {
    CAV_TRY(launch_rebuild_check(c));
  }
  timer_begin(c, c->t_nb);
  // cells are needed on rebuild steps, and every step by the LJ cell walk
  CAV_TRY(launch_bins(c, !c->lj_cellwalk, true));
  timer_end(c, c->t_nb);
  timer_begin(c, c->t_list);
  CAV_TRY(launch_list(c, true));
  if (forced) {  // the build is known to run
    c->nll_built = c->nll_pending;
    c->nlb_built = c->nlb_pending;
  }
  CAV_TRY(launch_save_old(c, true));
  timer_end(c, c->t_list);
  return CAV_OK;
}

static int check_capacity(cav_ctx *c, bool *redo)
{
  *redo = false;
  if (c->h_flags[F_OVERFLOW]) {
    const int need = c->h_flags[F_MAXNB];
    int cap = need + need / 4 + 16;
    cap = (cap + 3) / 4 * 4;
    CAV_TRY(alloc_list(c, cap));
    *redo = true;
  }
  return CAV_OK;
}

struct Nine {
  double *p[9];
};

// The owned positions, velocities and accelerations to or from the snapshot,
// in one launch (nine separate copies cost nine launches per step).
__global__ void k_snapshot(int n, Nine a, double *snap, int save)
{
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int k = 0; k < 9; ++k) {
    if (save)
      snap[(long)k * n + i] = a.p[k][i];
    else
      a.p[k][i] = snap[(long)k * n + i];
  }
}

static int snapshot(cav_ctx *c, bool save)
{
  const Nine arr = {{c->x, c->y, c->z, c->vx, c->vy, c->vz, c->ax, c->ay, c->az}};
  k_snapshot<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, arr, c->snap, save ? 1 : 0);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

static void harvest_timers(cav_ctx *c)
{
  if (!c->timing) return;
  timer_harvest(c->t_pair);
  timer_harvest(c->t_sk);
  timer_harvest(c->t_force);
  timer_harvest(c->t_tab);
  timer_harvest(c->t_nb);
  timer_harvest(c->t_int);
  timer_harvest(c->t_list);
}

// ------------------------------------------------------------------- C ABI --

extern "C" {

int cav_create(int device, cav_ctx **out)
{
  if (!out) return CAV_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return CAV_ERR_HIP;
  if (device < 0 || device >= ndev) return CAV_ERR_ARG;
  cav_ctx *c = new cav_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return CAV_ERR_HIP;
  }
  if (hipMalloc(&c->flags, sizeof(int) * F_COUNT) != hipSuccess ||
      hipHostMalloc(&c->h_flags, sizeof(int) * F_COUNT) != hipSuccess ||
      hipMalloc(&c->scal, sizeof(double) * 32) != hipSuccess) {
    delete c;
    return CAV_ERR_HIP;
  }
  (void)hipMemset(c->flags, 0, sizeof(int) * F_COUNT);
  (void)hipMemset(c->scal, 0, sizeof(double) * 32);
  *out = c;
  return CAV_OK;
}

int cav_destroy(cav_ctx *c)
{
  if (!c) return CAV_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->kstream) (void)hipStreamSynchronize(c->kstream);
  free_all(c);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->kstream) (void)hipStreamDestroy(c->kstream);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return CAV_OK;
}

const char *cav_last_error(const cav_ctx *c) { return c ? c->err.c_str() : "null context"; }

int cav_set_precision(cav_ctx *c, int p)
{
  if (!c || (p != CAV_FP64 && p != CAV_MIXED)) return fail(c, CAV_ERR_ARG, "precision must be CAV_FP64 or CAV_MIXED");
  c->precision = p;
  c->setup_done = false;
  return CAV_OK;
}

int cav_set_box(cav_ctx *c, const double lower[3], const double upper[3], const int bc[3])
{
  if (!c || !lower || !upper || !bc) return fail(c, CAV_ERR_ARG, "null argument");
  for (int d = 0; d < 3; ++d) {
    if (!(upper[d] > lower[d])) return fail(c, CAV_ERR_ARG, "box upper must exceed lower");
    if (bc[d] != 0 && bc[d] != 1) return fail(c, CAV_ERR_ARG, "boundary_condition must be 0 or 1");
    c->lo[d] = lower[d];
    c->hi[d] = upper[d];
    c->L[d] = upper[d] - lower[d];  // domain.cpp:470-477
    c->bc[d] = bc[d];
  }
  c->box_set = true;
  c->setup_done = false;
  return CAV_OK;
}

int cav_set_ghost_cutoff(cav_ctx *c, double gc)
{
  if (!c || gc < 0) return fail(c, CAV_ERR_ARG, "ghost_cutoff have to non-negative.");
  c->ghost_cutoff = gc;
  c->setup_done = false;
  return CAV_OK;
}

int cav_set_types(cav_ctx *c, int nt, const double *mass, const double *charge)
{
  if (!c || nt <= 0 || !mass || !charge) return fail(c, CAV_ERR_ARG, "types need mass and charge");
  const bool keep = c->setup_done && nt == c->nt;
  c->nt = nt;
  c->mass.assign(mass, mass + nt);
  c->charge.assign(charge, charge + nt);
  c->minv.resize(nt);
  for (int t = 0; t < nt; ++t) c->minv[t] = mass[t] == 0.0 ? 0.0 : 1.0 / mass[t];  // atom_data.cpp:236-239
  if (keep)
    c->params_dirty = true;  // values only: refreshed before the next force evaluation
  else
    c->setup_done = false;
  return CAV_OK;
}

int cav_set_neighbor(cav_ctx *c, double cutoff, double coef, int from_cells, int build_cells,
                     int all_atom_test, int rebuild_test)
{
  if (!c) return CAV_ERR_ARG;
  if (cutoff <= 0.0) return fail(c, CAV_ERR_ARG, "cutoff have to non-negative.");
  if (coef < 1.0) return fail(c, CAV_ERR_ARG, "cutoff_extra_coef must be larger than 1.0");
  c->nl_on = true;
  c->nl_cutoff = cutoff;
  c->nl_coef = coef;
  c->from_cells = from_cells;
  c->build_cells = build_cells || from_cells;  // neighborlist.cpp:216-220
  c->all_atom_test = all_atom_test;
  c->rebuild_test = rebuild_test;
  c->setup_done = false;
  return CAV_OK;
}

int cav_set_lj(cav_ctx *c, double rc, const double *eps, const double *sigma)
{
  if (!c || !eps || !sigma || c->nt <= 0) return fail(c, CAV_ERR_ARG, "set types before lj");
  if (rc < 0) return fail(c, CAV_ERR_ARG, "Force field cutoff have to non-negative.");
  const int nt2 = c->nt * c->nt;
  for (int p = 0; p < nt2; ++p)
    if (eps[p] < 0 || sigma[p] < 0) return fail(c, CAV_ERR_ARG, "Epsilon and sigma have to be non-negative.");
  const bool keep = c->setup_done && c->lj_on && rc == c->lj_rc;
  c->lj_on = true;
  c->lj_rc = rc;
  c->eps.assign(eps, eps + nt2);
  c->sig.assign(sigma, sigma + nt2);
  if (keep)
    c->params_dirty = true;
  else
    c->setup_done = false;
  return CAV_OK;
}

int cav_set_lj_list(cav_ctx *c, double cutoff, const double *eps, const double *sig,
                    const double *eps_atom, const double *sig_atom, const double *cutoff_list,
                    int wca, const double *lambda_s, const double *lambda_e,
                    int ignore_intra_molecule, int jump_fix, double jump_tol)
{
  if (!c || c->nt <= 0) return fail(c, CAV_ERR_ARG, "set types before lj");
  if (cutoff < 0) return fail(c, CAV_ERR_ARG, "Force field cutoff have to non-negative.");
  const int nt = c->nt, nt2 = nt * nt;
  const bool by_array = eps || sig, by_atom = eps_atom || sig_atom;
  if (by_array && by_atom) return fail(c, CAV_ERR_ARG, "lj: (input_by_atom && input_by_array)");
  if (!by_array && !by_atom) return fail(c, CAV_ERR_ARG, "lj: no epsilon/sigma given");
  for (int p = 0; p < nt2; ++p) {
    if ((eps && eps[p] < 0) || (sig && sig[p] < 0) || (cutoff_list && cutoff_list[p] < 0) ||
        (lambda_s && lambda_s[p] < 0) || (lambda_e && lambda_e[p] < 0))
      return fail(c, CAV_ERR_ARG, "lj: epsilon, sigma, cutoff_list and lambdas have to be non-negative.");
  }
  // Lj::verify_settings (lj.cpp:162-339): type matrices by array, or by atom
  // with Lorentz-Berthelot off the diagonal (lj.cpp:207-222); WCA cutoffs from
  // the unscaled sigma (lj.cpp:278-294); then lambda_s scales sigma and
  // lambda_e epsilon (lj.cpp:310-338)
  std::vector<double> e(nt2, 0.0), s(nt2, 0.0), cut(nt2, 0.0);
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) {
      const int p = i * nt + j;
      if (by_atom) {
        const double ea = eps_atom ? eps_atom[i] : 0, eb = eps_atom ? eps_atom[j] : 0;
        const double sa = sig_atom ? sig_atom[i] : 0, sb = sig_atom ? sig_atom[j] : 0;
        s[p] = i == j ? sa : 0.5 * (sa + sb);
        e[p] = i == j ? ea : std::sqrt(ea * eb);
      } else {
        e[p] = eps ? eps[p] : 0;
        s[p] = sig ? sig[p] : 0;
      }
    }
  if (wca) {
    const double cut_coef = std::pow(2.0, 1.0 / 6.0);
    for (int p = 0; p < nt2; ++p) cut[p] = cut_coef * s[p];
  } else if (cutoff_list) {
    cut.assign(cutoff_list, cutoff_list + nt2);
  }
  for (int p = 0; p < nt2; ++p) {
    if (lambda_s) s[p] *= lambda_s[p];
    if (lambda_e) e[p] *= lambda_e[p];
  }
  const bool cut_on = wca || cutoff_list;
  const bool keep = c->setup_done && c->ljl_on;
  c->ljl_on = true;
  c->ljl_rc = cutoff;
  c->ljl_eps = e;
  c->ljl_sig = s;
  c->ljl_cut = cut;
  c->ljl_cut_on = cut_on;
  c->ljl_ignore_intra = ignore_intra_molecule != 0;
  c->ljl_jump = jump_fix != 0;
  c->ljl_jump_tol = jump_tol;
  if (keep)
    c->params_dirty = true;
  else
    c->setup_done = false;
  return CAV_OK;
}

int cav_set_short_range(cav_ctx *c, double cutoff, double ke, double beta, const double ext[3],
                        const double *lambda)
{
  if (!c) return CAV_ERR_ARG;
  if (cutoff < 0) return fail(c, CAV_ERR_ARG, "Force field cutoff have to non-negative.");
  if (ke < 0) return fail(c, CAV_ERR_ARG, "k_electrostatic has to be non-negative.");
  if (lambda && c->nt <= 0) return fail(c, CAV_ERR_ARG, "set types before lambda");
  if (lambda)
    for (int p = 0; p < c->nt * c->nt; ++p)
      if (lambda[p] < 0) return fail(c, CAV_ERR_ARG, "lambda has to be non-negative.");
  const bool keep = c->setup_done && c->sr_on && (lambda != nullptr) == c->sr_lambda_set;
  c->sr_on = true;
  c->sr_rc = cutoff;
  c->sr_ke = ke;
  c->sr_beta = beta;
  for (int d = 0; d < 3; ++d) c->sr_ext[d] = ext ? ext[d] : 0.0;
  c->sr_lambda_set = lambda != nullptr;
  if (lambda) c->sr_lambda.assign(lambda, lambda + c->nt * c->nt);
  if (keep)
    c->params_dirty = true;
  else
    c->setup_done = false;
  return CAV_OK;
}

int cav_set_elc(cav_ctx *c, int axis, int kx, int ky, double lz, double ke)
{
  if (!c) return CAV_ERR_ARG;
  if (axis < 0 || axis > 2) return fail(c, CAV_ERR_ARG, "expected 'x', 'y' or 'z'.");
  if (kx < 0 || ky < 0) return fail(c, CAV_ERR_ARG, "'k_max' have to non-negative.");
  if (lz < 0) return fail(c, CAV_ERR_ARG, "'lz' has to be non-negative.");
  if (ke < 0) return fail(c, CAV_ERR_ARG, "k_electrostatic has to be non-negative.");
  c->elc_on = true;
  c->elc_axis = axis;
  c->elc_kx = kx;
  c->elc_ky = ky;
  c->elc_lz = lz;
  c->elc_ke = ke;
  c->setup_done = false;
  return CAV_OK;
}

int cav_set_thermo(cav_ctx *c, double k_b, int n_r_df, int temperature_process, int pressure_process)
{
  if (!c) return CAV_ERR_ARG;
  if (n_r_df < -1 || n_r_df > 6) return fail(c, CAV_ERR_ARG, "n_r_df must be -1 or 0..6");
  c->k_b = k_b;
  c->n_r_df = n_r_df;
  c->temperature_process = temperature_process != 0;
  c->pressure_process = pressure_process != 0;
  c->thermo_on = c->temperature_process || c->pressure_process || c->ber_on || c->nve_on;
  return CAV_OK;
}

int cav_set_berendsen(cav_ctx *c, double temperature, double coupling, double dt, int step,
                      double lambda_min, double lambda_max)
{
  if (!c) return CAV_ERR_ARG;
  if (temperature <= 0 || coupling <= 0 || dt <= 0 || step <= 0 || lambda_min <= 0 || lambda_max <= 0)
    return fail(c, CAV_ERR_ARG, "berendsen: temperature, coupling, dt, step and lambdas have to be positive");
  c->ber_on = true;
  c->ber_t0 = temperature;
  c->ber_coupling = coupling;
  c->ber_dt = dt;
  c->ber_step = step;
  c->ber_lmin = lambda_min;
  c->ber_lmax = lambda_max;
  c->thermo_on = true;
  return CAV_OK;
}

int cav_set_nve(cav_ctx *c, double energy_per_dof, double energy_tot, double kbt, double kb,
                double temperature, int step)
{
  if (!c) return CAV_ERR_ARG;
  if (step <= 0) return fail(c, CAV_ERR_ARG, "step have to non-negative.");
  // Nve::verify_settings (constraint/nve.cpp:97-132)
  const bool kbt_set = kbt > 0, t_set = temperature > 0, kb_set = kb > 0;
  if ((t_set && kb_set) != (t_set || kb_set)) return fail(c, CAV_ERR_ARG, "kb and temperature should be set together.");
  const bool e_set = energy_per_dof > 0 || energy_tot > 0, kt_set = (t_set && kb_set) || kbt_set;
  if (!(e_set || kt_set)) return fail(c, CAV_ERR_ARG, "nve: an energy or a temperature has to be set");
  if (e_set && kt_set) return fail(c, CAV_ERR_ARG, "Only one of 'energy' and 'temperature' can be set.");
  if (energy_per_dof > 0 && energy_tot > 0)
    return fail(c, CAV_ERR_ARG, "Only one of 'energy_per_dof' and 'energy_tot' can be set.");
  if (kbt_set && t_set && kb_set) return fail(c, CAV_ERR_ARG, "Only one of 'kbt' and 'temperatue kb' can be set.");
  c->nve_on = true;
  c->nve_e_dof = energy_per_dof;
  c->nve_e_tot = energy_tot;
  c->nve_kbt = kbt;
  c->nve_kb = kb;
  c->nve_t = temperature;
  c->nve_step = step;
  c->thermo_on = true;
  return CAV_OK;
}

int cav_set_step_index(cav_ctx *c, long step)
{
  if (!c) return CAV_ERR_ARG;
  c->step_index = step;
  return CAV_OK;
}

int cav_thermo(cav_ctx *c, double *kinetic_energy, double *temperature, double *pressure)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  (void)hipSetDevice(c->device);
  return thermo_read(c, kinetic_energy, temperature, pressure);
}

int cav_set_ewald_r(cav_ctx *c, double alpha, double ke, const double *lambda)
{
  if (!c) return CAV_ERR_ARG;
  if (alpha < 0) return fail(c, CAV_ERR_ARG, "alpha have to non-negative.");
  if (ke < 0) return fail(c, CAV_ERR_ARG, "k_electrostatic has to be non-negative.");
  const bool keep = c->setup_done && c->er_on && (lambda != nullptr) == c->lambda_set;
  c->er_on = true;
  c->er_alpha = alpha;
  c->er_ke = ke;
  c->lambda_set = lambda != nullptr;
  if (lambda) {
    if (c->nt <= 0) return fail(c, CAV_ERR_ARG, "set types before lambda");
    c->lambda.assign(lambda, lambda + c->nt * c->nt);
  }
  if (keep)
    c->params_dirty = true;
  else
    c->setup_done = false;
  return CAV_OK;
}

int cav_set_ewald_k(cav_ctx *c, double alpha, double ke, int kx, int ky, int kz, int dipole,
                    double eps_dipole, int slab_axis, double slab_lz)
{
  if (!c) return CAV_ERR_ARG;
  if (alpha <= 0) return fail(c, CAV_ERR_ARG, "Electrostatic_ewald_k needs a positive alpha");
  if (kx < 0 || ky < 0 || kz < 0) return fail(c, CAV_ERR_ARG, "'k_max' have to non-negative.");
  if (ke < 0) return fail(c, CAV_ERR_ARG, "'k_electrostatic' has to be non-negative.");
  if (slab_axis < -1 || slab_axis > 2) return fail(c, CAV_ERR_ARG, "slab_normal_axis must be x, y or z");
  const bool keep = c->setup_done && c->ek_on && kx == c->kx && ky == c->ky && kz == c->kz &&
                    slab_axis == c->slab_axis && (dipole != 0) == (c->dipole != 0);
  c->ek_on = true;
  c->ek_alpha = alpha;
  c->ek_ke = ke;
  c->kx = kx;
  c->ky = ky;
  c->kz = kz;
  c->dipole = dipole;
  c->eps_dipole = eps_dipole;
  c->slab_axis = slab_axis;
  c->slab_lz = slab_lz;
  if (keep)
    c->params_dirty = true;
  else
    c->setup_done = false;
  return CAV_OK;
}

int cav_upload_atoms(cav_ctx *c, int n, const int *id, const int *type, const int *mol,
                     const double *pos, const double *vel)
{
  if (!c || n < 0 || (n && (!id || !type || !pos || !vel))) return fail(c, CAV_ERR_ARG, "null atom arrays");
  if (c->nt <= 0) return fail(c, CAV_ERR_ARG, "set types before atoms");
  for (int i = 0; i < n; ++i)
    if (type[i] < 0 || type[i] >= c->nt) return fail(c, CAV_ERR_ARG, "atom type outside the declared types");
  (void)hipSetDevice(c->device);
  c->n = n;
  c->gcap = 7 * n;  // at most seven images per owned atom
  c->h_id.assign(id, id + n);
  c->h_type.assign(type, type + n);
  c->h_mol.assign(n, -1);
  if (mol) c->h_mol.assign(mol, mol + n);
  const size_t tot = (size_t)n + c->gcap;
  double **pa[] = {&c->x, &c->y, &c->z, &c->xo, &c->yo, &c->zo, &c->q};
  for (double **p : pa) CAV_TRY(dmalloc(c, p, tot));
  double **oa[] = {&c->vx, &c->vy, &c->vz, &c->ax, &c->ay, &c->az, &c->aox, &c->aoy, &c->aoz};
  for (double **p : oa) CAV_TRY(dmalloc(c, p, n));
  CAV_TRY(dmalloc(c, &c->snap, 9 * (size_t)n));
  CAV_TRY(dmalloc(c, &c->type, tot));
  CAV_TRY(dmalloc(c, &c->id, tot));
  CAV_TRY(dmalloc(c, &c->mol, tot));
  CAV_TRY(dmalloc(c, &c->gsrc, c->gcap));
  CAV_TRY(dmalloc(c, &c->bin_of, tot));
  CAV_TRY(dmalloc(c, &c->bin_items, tot));
  CAV_TRY(dmalloc(c, &c->cpos, tot));
  CAV_TRY(dmalloc(c, &c->binpos, tot));
  CAV_TRY(dmalloc(c, &c->litems, tot));
  // no list build's bin order exists for the new slot arrays yet
  k_set_flag<<<1, 1, 0, c->stream>>>(c->flags, F_NGHOST_OLD, -1);
  CAV_TRY(dmalloc(c, &c->xqs, tot));
  CAV_TRY(dmalloc(c, &c->tids, tot));
  CAV_TRY(dmalloc(c, &c->nn, n));
  CAV_TRY(dmalloc(c, &c->order, n));
  CAV_TRY(dmalloc(c, &c->xq, tot));
  CAV_TRY(dmalloc(c, &c->tid, tot));
  CAV_TRY(dmalloc(c, &c->ord_blk, grid_for((long)tot) + 1));
  c->nblk = grid_for(n) + 1;
  CAV_TRY(dmalloc(c, &c->blk, c->nblk));
  std::vector<double> soa(3 * (size_t)n), qv(n);
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) soa[(size_t)k * n + i] = pos[3 * i + k];
    qv[i] = c->charge[type[i]];
  }
  const size_t b = sizeof(double) * n;
  CAV_HIP(hipMemcpy(c->x, soa.data(), b, hipMemcpyHostToDevice));
  CAV_HIP(hipMemcpy(c->y, soa.data() + n, b, hipMemcpyHostToDevice));
  CAV_HIP(hipMemcpy(c->z, soa.data() + 2 * (size_t)n, b, hipMemcpyHostToDevice));
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) soa[(size_t)k * n + i] = vel[3 * i + k];
  CAV_HIP(hipMemcpy(c->vx, soa.data(), b, hipMemcpyHostToDevice));
  CAV_HIP(hipMemcpy(c->vy, soa.data() + n, b, hipMemcpyHostToDevice));
  CAV_HIP(hipMemcpy(c->vz, soa.data() + 2 * (size_t)n, b, hipMemcpyHostToDevice));
  CAV_HIP(hipMemset(c->ax, 0, b));
  CAV_HIP(hipMemset(c->ay, 0, b));
  CAV_HIP(hipMemset(c->az, 0, b));
  CAV_HIP(hipMemset(c->aox, 0, b));
  CAV_HIP(hipMemset(c->aoy, 0, b));
  CAV_HIP(hipMemset(c->aoz, 0, b));
  CAV_HIP(hipMemcpy(c->q, qv.data(), b, hipMemcpyHostToDevice));
  CAV_HIP(hipMemcpy(c->type, type, sizeof(int) * n, hipMemcpyHostToDevice));
  CAV_HIP(hipMemcpy(c->id, id, sizeof(int) * n, hipMemcpyHostToDevice));
  CAV_HIP(hipMemcpy(c->mol, c->h_mol.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  c->setup_done = false;
  return CAV_OK;
}

int cav_num_atoms(const cav_ctx *c, int *n)
{
  if (!c || !n) return CAV_ERR_ARG;
  *n = c->n;
  return CAV_OK;
}

static double *soa_of(cav_ctx *c, int which, int k)
{
  double *a[4][3] = {{c->x, c->y, c->z}, {c->vx, c->vy, c->vz}, {c->ax, c->ay, c->az}, {c->aox, c->aoy, c->aoz}};
  return a[which][k];
}

int cav_download(cav_ctx *c, int which, double *out)
{
  if (!c || !out || which < 0 || which > 3) return fail(c, CAV_ERR_ARG, "bad download request");
  (void)hipSetDevice(c->device);
  std::vector<double> soa(3 * (size_t)c->n);
  for (int k = 0; k < 3; ++k)
    CAV_HIP(hipMemcpyAsync(soa.data() + (size_t)k * c->n, soa_of(c, which, k), sizeof(double) * c->n,
                           hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->n; ++i)
    for (int k = 0; k < 3; ++k) out[3 * i + k] = soa[(size_t)k * c->n + i];
  return CAV_OK;
}

int cav_upload(cav_ctx *c, int which, const double *in)
{
  if (!c || !in || which < 0 || which > 3) return fail(c, CAV_ERR_ARG, "bad upload request");
  (void)hipSetDevice(c->device);
  std::vector<double> soa(3 * (size_t)c->n);
  for (int i = 0; i < c->n; ++i)
    for (int k = 0; k < 3; ++k) soa[(size_t)k * c->n + i] = in[3 * i + k];
  for (int k = 0; k < 3; ++k)
    CAV_HIP(hipMemcpyAsync(soa_of(c, which, k), soa.data() + (size_t)k * c->n, sizeof(double) * c->n,
                           hipMemcpyHostToDevice, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  return CAV_OK;
}

static int wrap_and_ghosts(cav_ctx *c)
{
  const double *lo, *hi, *L;
  box_dev(c, &lo, &hi, &L);
  k_wrap<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, lo, hi, L, c->bc[0], c->bc[1], c->bc[2],
                                                   c->x, c->y, c->z, c->flags);
  CAV_HIP(hipGetLastError());
  return launch_ghosts(c);
}

int cav_setup(cav_ctx *c)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->box_set) return fail(c, CAV_ERR_STATE, "set the box before setup");
  if (c->nt <= 0) return fail(c, CAV_ERR_STATE, "set types before setup");
  if ((c->lj_on || c->er_on || c->ljl_on || c->sr_on) && !c->nl_on)
    return fail(c, CAV_ERR_STATE, "lj_cell_list / lj / ewald_r / short_range need a neighborlist");
  if (c->lj_on && !c->build_cells) return fail(c, CAV_ERR_STATE, "lj_cell_list needs build_cell_list");
  (void)hipSetDevice(c->device);
  // box on the device (lo, hi, L)
  double box[9];
  for (int d = 0; d < 3; ++d) {
    box[d] = c->lo[d];
    box[3 + d] = c->hi[d];
    box[6 + d] = c->L[d];
  }
  CAV_HIP(hipMemcpy(c->scal + kBoxSlot, box, sizeof(box), hipMemcpyHostToDevice));
  CAV_TRY(upload_params(c));
  CAV_TRY(neighbor_init(c));
  // LJ over the list needs rc_lj <= the list's own cutoff (the skin guarantee of
  // the rebuild test); otherwise it walks the cells every step like the reference.
  c->lj_cellwalk = c->lj_on && (c->lj_rc > c->nl_cutoff || c->all_atom_test);
  const long cells = (c->nbins + 7) / 8 * 8;  // one partial per cell of the cell-staged kernel
  // one partial per block of the row-gather pair kernel (kLPA lanes per atom,
  // grid rounded up to a multiple of 8), per cell of the cell-staged kernel
  const long pair_blocks = (grid_for((long)kLPA * c->n) + 7) / 8 * 8;
  const int npart = (int)std::max<long>(std::max<long>(3L * grid_for(std::max<long>((long)c->n, 1L << 16)),
                                                       pair_blocks),
                                        cells) + 64;
  CAV_TRY(dmalloc(c, &c->part, npart));
  c->npart = npart;
  if (c->ek_on) {
    CAV_TRY(ewald_k_plan(c));
    const int need = 3 * grid_for((long)c->ek.Rpad * c->ek.nZ) + 64;
    if (need > c->npart) {
      CAV_TRY(dmalloc(c, &c->part, need));
      c->npart = need;
    }
    CAV_TRY(ewald_k_alloc(c));
    // the concurrent Ewald_k of a step: second stream, fork/join events, its
    // accelerations and partial sums
    CAV_TRY(dmalloc(c, &c->part_k, c->npart));
    CAV_TRY(dmalloc(c, &c->ekx, std::max(c->n, 1)));
    CAV_TRY(dmalloc(c, &c->eky, std::max(c->n, 1)));
    CAV_TRY(dmalloc(c, &c->ekz, std::max(c->n, 1)));
    if (!c->kstream) {
      CAV_HIP(hipStreamCreate(&c->kstream));
      CAV_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
      CAV_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
  }
  if (c->elc_on) CAV_TRY(elc_plan(c));
  CAV_TRY(thermo_alloc(c));
  // initial list capacity from the mean density, refined by the first build
  if (c->nl_on) {
    const double vol = c->L[0] * c->L[1] * c->L[2];
    const double expect = 4.18879020478639 * c->rce * c->rce * c->rce * (c->n / vol);
    int cap = (int)(2.0 * expect) + 32;
    if (c->all_atom_test) cap = 8 * c->n + 8;
    cap = (cap + 3) / 4 * 4;
    CAV_TRY(alloc_list(c, cap));
  }
  k_set_flag<<<1, 1, 0, c->stream>>>(c->flags, F_NGHOST_OLD, -1);
  for (int attempt = 0; attempt < 4; ++attempt) {
    if (attempt) CAV_TRY(elc_rewind(c));  // the retried evaluation is the same one
    k_begin_step<<<1, 1, 0, c->stream>>>(c->flags);
    CAV_TRY(wrap_and_ghosts(c));
    CAV_TRY(neighbor_phase(c, true));
    CAV_TRY(forces(c));
    CAV_TRY(sync_flags(c));
    bool redo = false;
    CAV_TRY(check_capacity(c, &redo));
    if (!redo) break;
  }
  if (c->h_flags[F_OVERFLOW]) return fail(c, CAV_ERR_CAPACITY, "Verlet list capacity could not be met");
  CAV_TRY(thermo_temperature(c));  // Md_simulator::setup (md_simulator.cpp:543)
  harvest_timers(c);
  c->setup_done = true;
  c->list_valid = c->nl_on;
  return CAV_OK;
}

int cav_step(cav_ctx *c, double dt, int nsteps)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run before cav_step");
  (void)hipSetDevice(c->device);
  CAV_TRY(refresh_params(c));
  const double *lo, *hi, *L;
  box_dev(c, &lo, &hi, &L);
  for (int s = 0; s < nsteps; ++s) {
    bool redo = true;
    for (int attempt = 0; redo && attempt < 4; ++attempt) {
      // the step snapshot (restored on NaN or overflow) is written by the drift
      k_begin_step<<<1, 1, 0, c->stream>>>(c->flags);
      timer_begin(c, c->t_int);
      k_drift<<<grid_for(c->n), kBlock, 0, c->stream>>>(
          c->n, dt, lo, hi, L, c->bc[0], c->bc[1], c->bc[2], c->x, c->y, c->z, c->vx, c->vy, c->vz,
          c->ax, c->ay, c->az, c->aox, c->aoy, c->aoz, c->flags, c->snap);
      timer_end(c, c->t_int);
      CAV_HIP(hipGetLastError());
      const bool side = c->overlap && c->ek_on && c->n > 0;
      if (side) CAV_TRY(ewald_k_fork(c));
      CAV_TRY(launch_ghosts(c));
      CAV_TRY(neighbor_phase(c, false));
      CAV_TRY(forces(c, !side));
      if (side && (c->dipole || c->elc_on)) {
        // a = ((a_sr + a_k) + a_dip) + a_elc, as in order: add, dipole, slab
        // correction, then the kick
        CAV_TRY(ewald_k_join(c));
        k_add3<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->ax, c->ay, c->az, c->ekx,
                                                         c->eky, c->ekz);
        if (c->dipole) CAV_TRY(launch_dipole(c));
        if (c->elc_on) CAV_TRY(launch_elc(c));
        k_kick<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, 0.5 * dt, c->vx, c->vy, c->vz,
                                                         c->ax, c->ay, c->az, c->aox, c->aoy,
                                                         c->aoz);
      } else if (side) {
        CAV_TRY(ewald_k_join(c));
        k_kick_join<<<grid_for(c->n), kBlock, 0, c->stream>>>(
            c->n, 0.5 * dt, c->vx, c->vy, c->vz, c->ax, c->ay, c->az, c->ekx, c->eky, c->ekz,
            c->aox, c->aoy, c->aoz);
      } else {
        k_kick<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, 0.5 * dt, c->vx, c->vy, c->vz,
                                                         c->ax, c->ay, c->az, c->aox, c->aoy,
                                                         c->aoz);
      }
      CAV_HIP(hipGetLastError());
      CAV_TRY(thermo_end_of_step(c, c->step_index));
      CAV_TRY(sync_flags(c));
      harvest_timers(c);
      if (c->h_flags[F_NAN]) {
        CAV_TRY(snapshot(c, false));
        CAV_TRY(elc_rewind(c));
        CAV_HIP(hipStreamSynchronize(c->stream));
        return fail(c, CAV_ERR_NAN, "NaN acceleration before the drift");
      }
      CAV_TRY(check_capacity(c, &redo));
      if (redo) {
        CAV_TRY(snapshot(c, false));
        CAV_TRY(elc_rewind(c));
        k_set_flag<<<1, 1, 0, c->stream>>>(c->flags, F_FORCE, 1);
      }
    }
    if (redo) return fail(c, CAV_ERR_CAPACITY, "Verlet list capacity could not be met");
    ++c->step_index;
  }
  return CAV_OK;
}

int cav_recalculate(cav_ctx *c, int force_rebuild)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  (void)hipSetDevice(c->device);
  CAV_TRY(refresh_params(c));
  k_begin_step<<<1, 1, 0, c->stream>>>(c->flags);
  CAV_TRY(wrap_and_ghosts(c));
  CAV_TRY(neighbor_phase(c, force_rebuild != 0));
  CAV_TRY(forces(c));
  CAV_TRY(sync_flags(c));
  bool redo = false;
  CAV_TRY(check_capacity(c, &redo));
  if (redo) {
    CAV_TRY(neighbor_phase(c, true));
    CAV_TRY(elc_rewind(c));
    CAV_TRY(forces(c));
    CAV_TRY(sync_flags(c));
  }
  harvest_timers(c);
  return CAV_OK;
}

int cav_neighbor_build(cav_ctx *c, int update, int *rebuilt)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  k_set_flag<<<1, 1, 0, c->stream>>>(c->flags, F_REBUILD, 0);
  CAV_TRY(neighbor_phase(c, update != 0));
  CAV_TRY(sync_flags(c));
  if (rebuilt) *rebuilt = c->h_flags[F_REBUILD];
  return CAV_OK;
}

int cav_exchange(cav_ctx *c, int *wrapped)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  (void)hipSetDevice(c->device);
  k_set_flag<<<1, 1, 0, c->stream>>>(c->flags, F_WRAPPED, 0);
  CAV_TRY(wrap_and_ghosts(c));
  CAV_TRY(sync_flags(c));
  if (wrapped) *wrapped = c->h_flags[F_WRAPPED] != 0;
  return CAV_OK;
}

int cav_reset_acceleration(cav_ctx *c)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  k_zero3<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->ax, c->ay, c->az);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

int cav_reset_virial(cav_ctx *c)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  k_set_scalar<<<1, 1, 0, c->stream>>>(c->scal, 0.0);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

int cav_forces(cav_ctx *c, int mask)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  if (mask & ~(CAV_FF_LJ | CAV_FF_EWALD_R | CAV_FF_EWALD_K | CAV_FF_LJ_LIST | CAV_FF_SHORT_RANGE | CAV_FF_ELC))
    return fail(c, CAV_ERR_ARG, "unknown force-field bit in mask");
  (void)hipSetDevice(c->device);
  CAV_TRY(refresh_params(c));
  const bool lj = (mask & CAV_FF_LJ) && c->lj_on, er = (mask & CAV_FF_EWALD_R) && c->er_on;
  const bool ljl = (mask & CAV_FF_LJ_LIST) && c->ljl_on, sr = (mask & CAV_FF_SHORT_RANGE) && c->sr_on;
  if ((lj && !c->lj_cellwalk) || er) CAV_TRY(launch_pairs(c, lj, er, true));
  if (lj && c->lj_cellwalk) CAV_TRY(launch_lj_cellwalk(c));
  if (ljl || sr) CAV_TRY(launch_pairs_list(c, ljl, sr, true, true));
  if ((mask & CAV_FF_EWALD_K) && c->ek_on) CAV_TRY(launch_ewald_k(c));
  if ((mask & CAV_FF_ELC) && c->elc_on) CAV_TRY(launch_elc(c));
  CAV_HIP(hipStreamSynchronize(c->stream));
  return CAV_OK;
}

int cav_autotune(cav_ctx *c)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  if (!c->ek_on) return CAV_OK;
  (void)hipSetDevice(c->device);
  CAV_TRY(ewald_k_autotune(c));
  // the tuning launches added k-space terms: recompute the forces of the current state
  CAV_TRY(elc_rewind(c));
  CAV_TRY(forces(c));
  CAV_TRY(sync_flags(c));
  return CAV_OK;
}

int cav_tiles(const cav_ctx *c, int *s_rows, int *s_stage, int *s_chunk, int *f_tiles,
              int *f_waves)
{
  if (!c) return CAV_ERR_ARG;
  if (s_rows) *s_rows = c->tune_rt;
  if (s_stage) *s_stage = c->tune_stage;
  if (s_chunk) *s_chunk = c->tune_chunk;
  if (f_tiles) *f_tiles = c->tune_fnt;
  if (f_waves) *f_waves = c->tune_fw;
  return CAV_OK;
}

int cav_set_tiles(cav_ctx *c, int s_rows, int s_chunk, int f_tiles, int f_waves)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  if (s_rows < 1 || s_rows > 4 || s_chunk <= 0 || s_chunk > 4096 || s_chunk % 32 || f_tiles < 1 ||
      f_tiles > 2 || (f_waves != 8 && f_waves != 16))
    return fail(c, CAV_ERR_ARG, "tiles out of range");
  if (!c->ek_on) return CAV_OK;
  (void)hipSetDevice(c->device);
  if (c->precision == CAV_MIXED) {
    c->tune_rt = s_rows;
    c->tune_stage = 32;
    c->tune_fnt = f_tiles;
    c->tune_fw = f_waves;
  }
  c->tune_chunk = s_chunk;
  c->tuned = true;
  CAV_TRY(ewald_k_alloc(c));
  CAV_TRY(elc_rewind(c));
  CAV_TRY(forces(c));
  CAV_TRY(sync_flags(c));
  return CAV_OK;
}

int cav_set_overlap(cav_ctx *c, int on)
{
  if (!c) return CAV_ERR_ARG;
  c->overlap = on != 0;
  return CAV_OK;
}

int cav_set_pair_kernel(cav_ctx *c, int kind)
{
  if (!c) return CAV_ERR_ARG;
  if (kind < CAV_PAIRS_AUTO || kind > CAV_PAIRS_CELLS)
    return fail(c, CAV_ERR_ARG, "pair kernel must be CAV_PAIRS_AUTO, _ROWS or _CELLS");
  c->pair_kind = kind;
  return CAV_OK;
}

int cav_pair_kernel(const cav_ctx *c, int *kind)
{
  if (!c || !kind) return CAV_ERR_ARG;
  *kind = cell_staged(c) ? CAV_PAIRS_CELLS : CAV_PAIRS_ROWS;
  return CAV_OK;
}

int cav_num_ghosts(cav_ctx *c, int *g)
{
  if (!c || !g) return CAV_ERR_ARG;
  CAV_TRY(sync_flags(c));
  *g = c->h_flags[F_NGHOST];
  return CAV_OK;
}

int cav_get_ghosts(cav_ctx *c, double *gpos, int *gsrc)
{
  if (!c) return CAV_ERR_ARG;
  CAV_TRY(sync_flags(c));
  const int G = c->h_flags[F_NGHOST];
  std::vector<double> t(3 * (size_t)G);
  double *arr[3] = {c->x, c->y, c->z};
  for (int k = 0; k < 3; ++k)
    CAV_HIP(hipMemcpyAsync(t.data() + (size_t)k * G, arr[k] + c->n, sizeof(double) * G, hipMemcpyDeviceToHost, c->stream));
  if (gsrc) CAV_HIP(hipMemcpyAsync(gsrc, c->gsrc, sizeof(int) * G, hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  if (gpos)
    for (int g = 0; g < G; ++g)
      for (int k = 0; k < 3; ++k) gpos[3 * g + k] = t[(size_t)k * G + g];
  return CAV_OK;
}

static int list_host(cav_ctx *c, std::vector<int> &nl, std::vector<int> &nn)
{
  nl.resize((size_t)c->nl_cap * c->n);
  nn.resize(c->n);
  CAV_TRY(launch_nl_fill(c));
  CAV_HIP(hipMemcpyAsync(nl.data(), c->nl, sizeof(int) * nl.size(), hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipMemcpyAsync(nn.data(), c->nn, sizeof(int) * nn.size(), hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  return CAV_OK;
}

int cav_num_pairs(cav_ctx *c, long *np)
{
  if (!c || !np) return CAV_ERR_ARG;
  if (!c->nl_on || !c->setup_done) return fail(c, CAV_ERR_STATE, "no neighbour list");
  std::vector<int> nl, nn;
  CAV_TRY(list_host(c, nl, nn));
  long cnt = 0;
  for (int i = 0; i < c->n; ++i)
    for (int k = 0; k < nn[i]; ++k) {
      const int j = nl[(size_t)i * c->nl_cap + k];
      cnt += (j >= c->n || j > i);
    }
  *np = cnt;
  return CAV_OK;
}

int cav_get_pairs(cav_ctx *c, int *ij, long capacity, long *np)
{
  if (!c || !ij) return CAV_ERR_ARG;
  if (!c->nl_on || !c->setup_done) return fail(c, CAV_ERR_STATE, "no neighbour list");
  std::vector<int> nl, nn;
  CAV_TRY(list_host(c, nl, nn));
  long w = 0;
  for (int i = 0; i < c->n; ++i)
    for (int k = 0; k < nn[i]; ++k) {
      const int j = nl[(size_t)i * c->nl_cap + k];
      if (!(j >= c->n || j > i)) continue;
      if (w < capacity) {
        ij[2 * w] = i;
        ij[2 * w + 1] = j;
      }
      ++w;
    }
  if (np) *np = w;
  if (w > capacity) return fail(c, CAV_ERR_CAPACITY, "pair buffer too small");
  return CAV_OK;
}

int cav_list_stats(cav_ctx *c, double *ppa, double *lj_ppa)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->nl_on || !c->setup_done) return fail(c, CAV_ERR_STATE, "no neighbour list");
  std::vector<int> nl, nn;
  CAV_TRY(list_host(c, nl, nn));
  std::vector<double> x(c->n + (size_t)c->gcap), y(x.size()), z(x.size());
  CAV_TRY(sync_flags(c));
  const size_t tot = (size_t)c->n + c->h_flags[F_NGHOST];
  CAV_HIP(hipMemcpy(x.data(), c->x, sizeof(double) * tot, hipMemcpyDeviceToHost));
  CAV_HIP(hipMemcpy(y.data(), c->y, sizeof(double) * tot, hipMemcpyDeviceToHost));
  CAV_HIP(hipMemcpy(z.data(), c->z, sizeof(double) * tot, hipMemcpyDeviceToHost));
  long half = 0, lj = 0;
  const double rc2 = c->lj_rc * c->lj_rc;
  for (int i = 0; i < c->n; ++i)
    for (int k = 0; k < nn[i]; ++k) {
      const int j = nl[(size_t)i * c->nl_cap + k];
      if (!(j >= c->n || j > i)) continue;
      ++half;
      const double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
      if (c->lj_on && dx * dx + dy * dy + dz * dz <= rc2) ++lj;
    }
  if (ppa) *ppa = c->n ? (double)half / c->n : 0;
  if (lj_ppa) *lj_ppa = c->n ? (double)lj / c->n : 0;
  return CAV_OK;
}

int cav_virial(cav_ctx *c, double *v)
{
  if (!c || !v) return CAV_ERR_ARG;
  CAV_HIP(hipMemcpyAsync(v, c->scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  return CAV_OK;
}

int cav_num_kvectors(cav_ctx *c, int *nk)
{
  if (!c || !nk) return CAV_ERR_ARG;
  if (!c->ek_on || !c->setup_done) return fail(c, CAV_ERR_STATE, "no Ewald_k");
  *nk = (int)c->ek.ref_r.size();
  return CAV_OK;
}

int cav_get_sk(cav_ctx *c, double *out)
{
  if (!c || !out) return CAV_ERR_ARG;
  if (!c->ek_on || !c->setup_done) return fail(c, CAV_ERR_STATE, "no Ewald_k");
  const EwaldKPlan &p = c->ek;
  std::vector<double> sk((size_t)p.Rpad * 2 * p.nZ);
  CAV_HIP(hipMemcpyAsync(sk.data(), c->sk, sizeof(double) * sk.size(), hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  for (size_t k = 0; k < p.ref_r.size(); ++k) {
    const double *s = sk.data() + (size_t)p.ref_r[k] * 2 * p.nZ + 2 * p.ref_c[k];
    out[2 * k] = s[0];
    out[2 * k + 1] = p.ref_conj[k] ? -s[1] : s[1];
  }
  return CAV_OK;
}

int cav_energy(cav_ctx *c, double *e_lj, double *e_r, double *e_k)
{
  if (!c) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  (void)hipSetDevice(c->device);
  CAV_TRY(refresh_params(c));
  if (e_k) {
    if (!c->ek_on) return fail(c, CAV_ERR_STATE, "no Ewald_k");
    CAV_TRY(launch_energy_k(c, e_k));
  }
  if (e_r || e_lj) {
    if (!c->nl_on) return fail(c, CAV_ERR_STATE, "no neighbour list");
    // the cell walk needs cells of the current positions and ghosts
    CAV_TRY(launch_bins(c, false));
  }
  if (e_r) {
    if (!c->er_on) return fail(c, CAV_ERR_STATE, "no Ewald_r");
    if (!c->build_cells) return fail(c, CAV_ERR_STATE, "Ewald_r energy walks the cell list: needs build_cell_list");
    CAV_TRY(launch_energy_r(c, e_r));
  }
  if (e_lj) {
    if (!c->lj_on) return fail(c, CAV_ERR_STATE, "no lj_cell_list");
    CAV_TRY(launch_energy_lj(c, e_lj));
  }
  return CAV_OK;
}

int cav_energy_elc(cav_ctx *c, double *e)
{
  if (!c || !e) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  if (!c->elc_on) return fail(c, CAV_ERR_STATE, "no electrostatic_ewald_slab_correction");
  (void)hipSetDevice(c->device);
  return energy_elc(c, e);
}

int cav_point_field(cav_ctx *c, int which, int np, const double *pts, double *field, double *pot)
{
  if (!c || np < 0 || (np && !pts)) return CAV_ERR_ARG;
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run first");
  (void)hipSetDevice(c->device);
  CAV_TRY(refresh_params(c));
  if (which == 0) {
    if (!c->er_on || !c->build_cells) return fail(c, CAV_ERR_STATE, "Ewald_r point queries need build_cell_list");
    CAV_TRY(launch_bins(c, false));
    return point_field_r(c, np, pts, field, pot);
  }
  if (which == 1) {
    if (!c->ek_on) return fail(c, CAV_ERR_STATE, "no Ewald_k");
    return point_field_k(c, np, pts, field, pot);
  }
  if (which == 2) {
    if (!c->elc_on) return fail(c, CAV_ERR_STATE, "no electrostatic_ewald_slab_correction");
    return point_field_elc(c, np, pts, field, pot);
  }
  return fail(c, CAV_ERR_ARG, "which must be 0 (ewald_r), 1 (ewald_k) or 2 (slab correction)");
}

int cav_enable_timing(cav_ctx *c, int on)
{
  if (!c) return CAV_ERR_ARG;
  c->timing = on != 0;
  Timer *ts[] = {&c->t_pair, &c->t_sk, &c->t_force, &c->t_tab, &c->t_nb, &c->t_int, &c->t_list};
  for (Timer *t : ts) {
    t->total_ms = 0;
    t->launches = 0;
    t->pending = false;
  }
  return CAV_OK;
}

int cav_kernel_time(cav_ctx *c, const char *name, double *ms, int *launches)
{
  if (!c || !name) return CAV_ERR_ARG;
  Timer *t = nullptr;
  if (!strcmp(name, "pair")) t = &c->t_pair;
  else if (!strcmp(name, "ewald_k_sk")) t = &c->t_sk;
  else if (!strcmp(name, "ewald_k_force")) t = &c->t_force;
  else if (!strcmp(name, "ewald_k_tables")) t = &c->t_tab;
  else if (!strcmp(name, "neighbor")) t = &c->t_nb;
  else if (!strcmp(name, "integrate")) t = &c->t_int;
  else if (!strcmp(name, "list")) t = &c->t_list;
  if (!t) return fail(c, CAV_ERR_ARG, "unknown kernel family");
  if (ms) *ms = t->launches ? t->total_ms / t->launches : 0.0;
  if (launches) *launches = t->launches;
  return CAV_OK;
}

}  // extern "C"
