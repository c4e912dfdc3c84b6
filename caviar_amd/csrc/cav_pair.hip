// Fused LJ + Ewald real-space pair kernel over the full Verlet rows, plus the
// cell-walk kernels (LJ when its cutoff exceeds the list guarantee, the
// real-space energy and point queries).  One lane per owned atom, no atomics:
// every atom gathers its own row, so FP64 results are reproducible run to run.
// No MFMA: the pair loop is a gather over irregular rows, not a contraction.
#include <algorithm>

#include "cav_internal.hpp"

namespace cav {

__device__ __forceinline__ double dot_rn_p(double a, double b, double c)
{
  return __dadd_rn(__dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)), __dmul_rn(c, c));
}

// erfc(x) = t * P(t) * exp(-x^2), t = 1/(1 + p x): the Hastings form of
// Abramowitz & Stegun 7.1.26 with a degree-8 polynomial whose coefficients
// were fitted here by weighted least squares on [0, 6] (relative error 1.3e-8
// in exact arithmetic, 2.6e-7 evaluated in FP32).  exp(-x^2) is shared with the
// Gaussian term of the Ewald force.
__device__ __forceinline__ float erfc_hastings(float x, float expmx2)
{
  const float t = __frcp_rn(fmaf(0.47047f, x, 1.0f));
  float p = -5.325673686e-02f;
  p = fmaf(p, t, 2.980402004e-01f);
  p = fmaf(p, t, -6.267941092e-01f);
  p = fmaf(p, t, 5.387638720e-01f);
  p = fmaf(p, t, -1.765473040e-01f);
  p = fmaf(p, t, 2.700992388e-01f);
  p = fmaf(p, t, 2.166366678e-01f);
  p = fmaf(p, t, 2.677430713e-01f);
  p = fmaf(p, t, 2.653150927e-01f);
  return t * p * expmx2;
}

struct PairArgs {
  int n, nt, cap;
  const double *x, *y, *z, *q, *minv;
  const int *type, *id, *mol, *nl, *nn;
  const int *order;   // owned atoms in cell order
  const double4 *xq;  // packed (x, y, z, q) of owned + ghost atoms
  const int2 *tid;    // packed (type, id)
  // LJ per type pair: coef(r) = -c12 r^-14 + c6 r^-8 + cshift
  const double *c12, *c6, *cshift;
  double lj_rc2;
  // Ewald real space
  double alpha, a2, ke, two_a_rpi;
  const double *lambda;
  double *ax, *ay, *az;
  double *part;
  int accumulate;  // add into the acceleration instead of overwriting it
  const RepDev *rep;  // batched replicas (blockIdx.y) or nullptr
};

constexpr int kLPA = 16;  // lanes per atom in the pair kernel

// Per-pair force coefficient (force on i = coef * (r_j - r_i)) for T = double
// (reference arithmetic) or float (mixed form).  kLPA lanes share one atom's
// row (coalesced row reads, balanced work); their partial forces are summed
// with a fixed shuffle tree, so FP64 results stay reproducible.
template <typename T, bool LJ, bool ER, bool LAM>
__global__ __launch_bounds__(kBlock) void k_pairs(PairArgs a)
{
  if (a.rep) {  // one replica per grid row: shift every array to it
    const RepDev &r = a.rep[blockIdx.y];
    const int base = r.base, obase = r.obase;
    a.n = r.n;
    a.xq += base;
    a.tid += base;
    a.nl += (long)obase * a.cap;
    a.nn += obase;
    a.ax += obase;
    a.ay += obase;
    a.az += obase;
    a.part += r.pbase;
  }
  long blk = blockIdx.x;
  if (!a.rep) {
    // XCD-aware order (the grid is a multiple of 8 and dispatch is round-robin
    // over the 8 XCDs): each XCD gets a contiguous run of cell-ordered atoms,
    // so neighbouring atoms' gathers share one L2
    const long per = gridDim.x >> 3;
    blk = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  }
  const long gid = blk * kBlock + threadIdx.x;
  const int t = (int)(gid / kLPA), sub = (int)(gid % kLPA);
  double vir = 0.0;
  T fx = 0, fy = 0, fz = 0;
  int i = 0, ti = 0;
  if (t < a.n) {
    i = a.order ? a.order[t] : t;
    const double4 pi4 = a.xq[i];
    const int2 tii = a.tid[i];
    ti = tii.x;
    const int idi = tii.y;
    const int moli = LAM ? a.mol[i] : 0;
    const T qi = (T)pi4.w;
    const T keqi = (T)(a.ke) * qi;
    const int cnt = a.nn[i];
    const int *row = a.nl + (long)i * a.cap;
    for (int k = sub; k < cnt; k += kLPA) {
      const int j = row[k];
      const double4 pj = a.xq[j];
      const double dxd = pj.x - pi4.x, dyd = pj.y - pi4.y, dzd = pj.z - pi4.z;
      const T dx = (T)dxd, dy = (T)dyd, dz = (T)dzd;
      T r2;
      if constexpr (sizeof(T) == 8)
        r2 = dot_rn_p(dxd, dyd, dzd);  // the exact rounding order of the reference test
      else
        r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
      T coef = 0;
      int2 tj = make_int2(0, 0);
      if constexpr (LJ || ER) tj = a.tid[j];
      // type-pair entry as the reference indexes it: [lower owned index][partner]
      const int p = (j < a.n && j < i) ? tj.x * a.nt + ti : ti * a.nt + tj.x;
      if constexpr (LJ) {
        if (!(r2 > (T)a.lj_rc2)) {  // lj_cell_list.cpp:141-142: skip dr^2 > rc^2
          const T ri2 = (T)1 / r2;
          const T ri4 = ri2 * ri2, ri8 = ri4 * ri4;
          coef += -(T)a.c12[p] * (ri8 * ri4 * ri2) + (T)a.c6[p] * ri8 + (T)a.cshift[p];
        }
      }
      if constexpr (ER) {
        // electrostatic_ewald_r.cpp:162-211: every listed pair, exact zero skipped
        if (!(dxd == 0 && dyd == 0 && dzd == 0)) {
          T lam = 1;
          bool skip = false;
          if constexpr (LAM) {
            if (moli == a.mol[j]) {
              const T l = (T)a.lambda[p];
              if (l == 0) skip = true;
              lam = l;
            }
          }
          if (!skip) {
            const T qj = (T)pj.w;
            T ce;
            if constexpr (sizeof(T) == 8) {
              const double rn = sqrt(r2);
              const double s = (a.two_a_rpi * exp(-a.a2 * r2) + erfc(a.alpha * rn) / rn) * (1.0 / r2);
              ce = -lam * a.ke * (double)qi * qj * s;
            } else {
              const float rinv = __frsqrt_rn(r2);
              const float rn = r2 * rinv;
              const float e = __expf(-(float)a.a2 * r2);
              const float ec = erfc_hastings((float)a.alpha * rn, e);
              const float s = fmaf((float)a.two_a_rpi, e, ec * rinv) * (rinv * rinv);
              ce = -lam * keqi * qj * s;
            }
            coef += ce;
            // virial once per reference list entry (owned j > i or ghost), id_i < id_j
            if ((j >= a.n || j > i) && idi < tj.y) vir += -(double)ce * (double)r2;
          }
        }
      }
      fx = fma(coef, dx, fx);
      fy = fma(coef, dy, fy);
      fz = fma(coef, dz, fz);
    }
  }
#pragma unroll
  for (int o = 1; o < kLPA; o <<= 1) {
    fx += __shfl_xor(fx, o, 64);
    fy += __shfl_xor(fy, o, 64);
    fz += __shfl_xor(fz, o, 64);
  }
  if (t < a.n && sub == 0) {
    const double mi = a.minv[ti];
    if (a.accumulate) {
      a.ax[i] += (double)fx * mi;
      a.ay[i] += (double)fy * mi;
      a.az[i] += (double)fz * mi;
    } else {
      a.ax[i] = (double)fx * mi;
      a.ay[i] = (double)fy * mi;
      a.az[i] = (double)fz * mi;
    }
  }
  // deterministic per-block virial partial
  __shared__ double red[kBlock];
  red[threadIdx.x] = vir;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
}

// Packed copies for the pair gathers: (x, y, z, q) and (type, id).
__global__ void k_pack(int n, long cap_total, const double *x, const double *y, const double *z,
                       const double *q, const int *type, const int *id, const int *flags,
                       double4 *xq, int2 *tid)
{
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= (long)n + flags[F_NGHOST] || a >= cap_total) return;
  xq[a] = make_double4(x[a], y[a], z[a], q[a]);
  tid[a] = make_int2(type[a], id[a]);
}

__global__ void k_reduce_part(int count, const double *part, double *out, int add)
{
  __shared__ double red[kBlock];
  double s = 0;
  for (int i = threadIdx.x; i < count; i += kBlock) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = add ? *out + red[0] : red[0];
}

int reduce_partials(cav_ctx *c, int count, int slot, bool add)
{
  k_reduce_part<<<1, kBlock, 0, c->stream>>>(count, c->part, c->scal + slot, add ? 1 : 0);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

static PairArgs pair_args(cav_ctx *c)
{
  PairArgs a;
  a.rep = nullptr;
  a.accumulate = 0;
  a.n = c->n;
  a.nt = c->nt;
  a.cap = c->nl_cap;
  a.x = c->x;
  a.y = c->y;
  a.z = c->z;
  a.q = c->q;
  a.minv = c->d_minv;
  a.type = c->type;
  a.id = c->id;
  a.mol = c->mol;
  a.nl = c->nl;
  a.nn = c->nn;
  a.order = c->order;
  a.xq = c->xq;
  a.tid = c->tid;
  const int nt2 = c->nt * c->nt;
  a.c12 = c->d_eps;           // laid out as [c12 | c6 | cshift] blocks, see upload
  a.c6 = c->d_eps + nt2;
  a.cshift = c->d_eps + 2 * nt2;
  a.lj_rc2 = c->lj_rc * c->lj_rc;
  a.alpha = c->er_alpha;
  a.a2 = c->er_alpha * c->er_alpha;
  a.ke = c->er_ke;
  a.two_a_rpi = 2 * c->er_alpha * CAV_PIS_INV;
  a.lambda = c->d_lambda;
  a.ax = c->ax;
  a.ay = c->ay;
  a.az = c->az;
  a.part = c->part;
  return a;
}

static int pair_grid(const cav_ctx *c) { return (grid_for((long)c->n * kLPA) + 7) / 8 * 8; }

template <typename T>
static void launch_pairs_t(cav_ctx *c, const PairArgs &a, bool lj, bool er, bool lam)
{
  const int g = pair_grid(c);
  if (lj && er && lam)
    k_pairs<T, true, true, true><<<g, kBlock, 0, c->stream>>>(a);
  else if (lj && er)
    k_pairs<T, true, true, false><<<g, kBlock, 0, c->stream>>>(a);
  else if (lj)
    k_pairs<T, true, false, false><<<g, kBlock, 0, c->stream>>>(a);
  else if (er && lam)
    k_pairs<T, false, true, true><<<g, kBlock, 0, c->stream>>>(a);
  else
    k_pairs<T, false, true, false><<<g, kBlock, 0, c->stream>>>(a);
}

int launch_pairs(cav_ctx *c, bool want_lj, bool want_er, bool accumulate)
{
  const bool lj = want_lj && c->lj_on && !c->lj_cellwalk, er = want_er && c->er_on;
  if (!lj && !er) return CAV_OK;
  PairArgs a = pair_args(c);
  a.accumulate = accumulate ? 1 : 0;
  const long tot = (long)c->n + c->gcap;
  k_pack<<<grid_for(tot), kBlock, 0, c->stream>>>(c->n, tot, c->x, c->y, c->z, c->q, c->type,
                                                 c->id, c->flags, c->xq, c->tid);
  timer_begin(c, c->t_pair);
  if (c->precision == CAV_FP64)
    launch_pairs_t<double>(c, a, lj, er, c->lambda_set);
  else
    launch_pairs_t<float>(c, a, lj, er, c->lambda_set);
  timer_end(c, c->t_pair);
  CAV_HIP(hipGetLastError());
  return reduce_partials(c, pair_grid(c), 0, accumulate);
}

__global__ void k_zero_acc(int n, double *ax, double *ay, double *az)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) ax[i] = ay[i] = az[i] = 0.0;
}

__global__ void k_zero_scalar(double *p) { *p = 0.0; }

__global__ void k_fill_d(double *p, long n, double v)
{
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void kb_zero_own(const RepDev *rep, double *ax, double *ay, double *az)
{
  const RepDev &r = rep[blockIdx.y];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= r.n) return;
  ax[r.obase + i] = ay[r.obase + i] = az[r.obase + i] = 0.0;
}

int launch_zero_acc(cav_ctx *c)
{
  k_zero_acc<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->ax, c->ay, c->az);
  k_zero_scalar<<<1, 1, 0, c->stream>>>(c->scal);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// --------------------------------------------------------------- cell walks --

struct CellArgs {
  int n;
  const double *x, *y, *z, *q;
  const int *type, *bin_of, *st_off, *st, *start, *items;
  double lox, loy, loz, cinv;
  int nbx, nby, nbz;
};

__device__ __forceinline__ int cell_of(const CellArgs &c, double px, double py, double pz)
{
  int ix = (int)floor(__dmul_rn(__dsub_rn(px, c.lox), c.cinv));
  int iy = (int)floor(__dmul_rn(__dsub_rn(py, c.loy), c.cinv));
  int iz = (int)floor(__dmul_rn(__dsub_rn(pz, c.loz), c.cinv));
  ix = ix < 0 ? 0 : (ix >= c.nbx ? c.nbx - 1 : ix);
  iy = iy < 0 ? 0 : (iy >= c.nby ? c.nby - 1 : iy);
  iz = iz < 0 ? 0 : (iz >= c.nbz ? c.nbz - 1 : iz);
  return ix + c.nbx * iy + c.nbx * c.nby * iz;
}

static CellArgs cell_args(cav_ctx *c)
{
  CellArgs a;
  a.n = c->n;
  a.x = c->x;
  a.y = c->y;
  a.z = c->z;
  a.q = c->q;
  a.type = c->type;
  a.bin_of = c->bin_of;
  a.st_off = c->d_st_off;
  a.st = c->d_st;
  a.start = c->bin_start;
  a.items = c->bin_items;
  a.lox = c->lo[0];
  a.loy = c->lo[1];
  a.loz = c->lo[2];
  a.cinv = c->cinv;
  a.nbx = c->nb[0];
  a.nby = c->nb[1];
  a.nbz = c->nb[2];
  return a;
}

// Lj_cell_list::calculate_acceleration (lj_cell_list.cpp:85-175) as a gather:
// every owned j != i and every ghost in the 27 cells, skip dr^2 > rc^2.  Adds
// into the acceleration (the pair kernel, if any, has written it first).
__global__ void k_lj_cellwalk(CellArgs c, int nt, const double *c12, const double *c6,
                              const double *cshift, double rc2, const double *minv, double *ax,
                              double *ay, double *az)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= c.n) return;
  const double xi = c.x[i], yi = c.y[i], zi = c.z[i];
  const int ti = c.type[i];
  const int b = c.bin_of[i];
  double fx = 0, fy = 0, fz = 0;
  for (int s = c.st_off[b]; s < c.st_off[b + 1]; ++s) {
    const int cell = c.st[s];
    for (int t = c.start[cell]; t < c.start[cell + 1]; ++t) {
      const int j = c.items[t];
      if (j == i) continue;
      const double dx = __dsub_rn(c.x[j], xi), dy = __dsub_rn(c.y[j], yi), dz = __dsub_rn(c.z[j], zi);
      const double r2 = dot_rn_p(dx, dy, dz);
      if (r2 > rc2) continue;
      const int tj = c.type[j];
      const int p = (j < c.n && j < i) ? tj * nt + ti : ti * nt + tj;
      const double ri2 = 1.0 / r2, ri4 = ri2 * ri2, ri8 = ri4 * ri4;
      const double coef = -c12[p] * (ri8 * ri4 * ri2) + c6[p] * ri8 + cshift[p];
      fx += coef * dx;
      fy += coef * dy;
      fz += coef * dz;
    }
  }
  const double mi = minv[ti];
  ax[i] += fx * mi;
  ay[i] += fy * mi;
  az[i] += fz * mi;
}

int launch_lj_cellwalk(cav_ctx *c)
{
  const int nt2 = c->nt * c->nt;
  k_lj_cellwalk<<<grid_for(c->n), kBlock, 0, c->stream>>>(
      cell_args(c), c->nt, c->d_eps, c->d_eps + nt2, c->d_eps + 2 * nt2, c->lj_rc * c->lj_rc,
      c->d_minv, c->ax, c->ay, c->az);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// Electrostatic_ewald_r::potential/field at a point (electrostatic_ewald_r/
// potential.cpp:30-86, field.cpp:31-96): every owned or ghost atom in the 27
// cells around the point's bin, no cutoff, exact-zero separation skipped.
__device__ void point_r_eval(const CellArgs &c, double px, double py, double pz, double alpha,
                             double *field, double *pot)
{
  const int b = cell_of(c, px, py, pz);
  const double a2 = alpha * alpha;
  double phi = 0, ex = 0, ey = 0, ez = 0;
  for (int s = c.st_off[b]; s < c.st_off[b + 1]; ++s) {
    const int cell = c.st[s];
    for (int t = c.start[cell]; t < c.start[cell + 1]; ++t) {
      const int j = c.items[t];
      const double dx = px - c.x[j], dy = py - c.y[j], dz = pz - c.z[j];
      if (dx == 0 && dy == 0 && dz == 0) continue;
      const double d2 = dx * dx + dy * dy + dz * dz, dn = sqrt(d2);
      const double ec = erfc(alpha * dn) / dn;
      const double qj = c.q[j];
      phi += qj * ec;
      const double sr = (2 * alpha * CAV_PIS_INV * exp(-a2 * d2) + ec) / d2;
      ex += qj * sr * dx;
      ey += qj * sr * dy;
      ez += qj * sr * dz;
    }
  }
  *pot = phi;
  field[0] = ex;
  field[1] = ey;
  field[2] = ez;
}

__global__ void k_energy_r(CellArgs c, double alpha, double *part)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double e = 0;
  if (i < c.n) {
    double f[3], phi;
    point_r_eval(c, c.x[i], c.y[i], c.z[i], alpha, f, &phi);
    e = c.q[i] * phi;
  }
  __shared__ double red[kBlock];
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

int launch_energy_r(cav_ctx *c, double *e)
{
  k_energy_r<<<grid_for(c->n), kBlock, 0, c->stream>>>(cell_args(c), c->er_alpha, c->part);
  CAV_TRY(reduce_partials(c, grid_for(c->n), 2));
  double v;
  CAV_HIP(hipMemcpyAsync(&v, c->scal + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  *e = 0.5 * v * c->er_ke;  // energy.cpp:50 and potential.cpp:85
  return CAV_OK;
}

__global__ void k_point_r(CellArgs c, int np, const double *pts, double alpha, double ke,
                          double *field, double *pot)
{
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  double f[3], phi;
  point_r_eval(c, pts[3 * p], pts[3 * p + 1], pts[3 * p + 2], alpha, f, &phi);
  pot[p] = phi * ke;
  field[3 * p] = f[0] * ke;
  field[3 * p + 1] = f[1] * ke;
  field[3 * p + 2] = f[2] * ke;
}

int point_field_r(cav_ctx *c, int np, const double *pts, double *field, double *pot)
{
  double *d;
  CAV_HIP(hipMalloc(&d, sizeof(double) * 7 * (size_t)(np > 0 ? np : 1)));
  CAV_HIP(hipMemcpyAsync(d, pts, sizeof(double) * 3 * np, hipMemcpyHostToDevice, c->stream));
  k_point_r<<<grid_for(np), kBlock, 0, c->stream>>>(cell_args(c), np, d, c->er_alpha, c->er_ke,
                                                    d + 3 * np, d + 6 * np);
  CAV_HIP(hipGetLastError());
  if (field)
    CAV_HIP(hipMemcpyAsync(field, d + 3 * np, sizeof(double) * 3 * np, hipMemcpyDeviceToHost,
                           c->stream));
  if (pot)
    CAV_HIP(hipMemcpyAsync(pot, d + 6 * np, sizeof(double) * np, hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  CAV_HIP(hipFree(d));
  return CAV_OK;
}

// Shifted-force LJ energy over the full rows: each (i,j) entry carries half of
// U(r) = 4e[(s/r)^12 - (s/r)^6] + r^2 cshift/2 - (same at rc); every pair is
// seen from both ends (owned pairs twice in the full rows, periodic pairs once
// from each owned end through its ghost image).
__global__ void k_energy_lj(PairArgs a, const double *u12, const double *u6, const double *uc,
                            double *part)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double e = 0;
  if (i < a.n) {
    const double xi = a.x[i], yi = a.y[i], zi = a.z[i];
    const int ti = a.type[i];
    for (int k = 0; k < a.nn[i]; ++k) {
      const int j = a.nl[(long)i * a.cap + k];
      const double dx = a.x[j] - xi, dy = a.y[j] - yi, dz = a.z[j] - zi;
      const double r2 = dot_rn_p(dx, dy, dz);
      if (r2 > a.lj_rc2) continue;
      const int tj = a.type[j];
      const int p = (j < a.n && j < i) ? tj * a.nt + ti : ti * a.nt + tj;
      const double ri2 = 1.0 / r2, ri6 = ri2 * ri2 * ri2;
      e += 0.5 * (u12[p] * ri6 * ri6 - u6[p] * ri6 + 0.5 * a.cshift[p] * r2 - uc[p]);
    }
  }
  __shared__ double red[kBlock];
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

int launch_energy_lj(cav_ctx *c, double *e)
{
  const int nt2 = c->nt * c->nt;
  PairArgs a = pair_args(c);
  k_energy_lj<<<grid_for(c->n), kBlock, 0, c->stream>>>(a, c->d_eps + 3 * nt2, c->d_eps + 4 * nt2,
                                                        c->d_eps + 5 * nt2, c->part);
  CAV_TRY(reduce_partials(c, grid_for(c->n), 1));
  CAV_HIP(hipMemcpyAsync(e, c->scal + 1, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  return CAV_OK;
}


// ---------------------------------------------------------- batched replicas --

__global__ void kb_pack(const RepDev *rep, const int *flags, const double *x, const double *y,
                        const double *z, const double *q, const int *type, const int *id,
                        double4 *xq, int2 *tid)
{
  const RepDev &r = rep[blockIdx.y];
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= (long)r.n + flags[16 * blockIdx.y + F_NGHOST]) return;
  const long s = r.base + a;
  xq[s] = make_double4(x[s], y[s], z[s], q[s]);
  tid[s] = make_int2(type[s], id[s]);
}

// per replica: sum of its per-block partials in a fixed order
__global__ void kb_reduce_rows(int nper, const double *part, double *out)
{
  __shared__ double red[kBlock];
  const double *p = part + (long)blockIdx.x * nper;
  double s = 0;
  for (int i = threadIdx.x; i < nper; i += kBlock) s += p[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

template <typename T>
static void batch_pairs_t(cav_batch *b, const PairArgs &a, dim3 g)
{
  if (b->lj_on && b->er_on)
    k_pairs<T, true, true, false><<<g, kBlock, 0, b->stream>>>(a);
  else if (b->lj_on)
    k_pairs<T, true, false, false><<<g, kBlock, 0, b->stream>>>(a);
  else
    k_pairs<T, false, true, false><<<g, kBlock, 0, b->stream>>>(a);
}

// LJ + Ewald_r accelerations of every replica (overwriting), virial per replica.
int batch_pairs(cav_batch *b)
{
  int maxslots = 0;
  for (const RepDev &r : b->reps) maxslots = std::max(maxslots, r.n + r.gcap);
  kb_pack<<<dim3(grid_for(maxslots), b->B), kBlock, 0, b->stream>>>(
      b->d_rep, b->flags, b->x, b->y, b->z, b->q, b->type, b->id, b->xq, b->tid);
  const dim3 g(b->pair_blocks, b->B);
  if (!b->lj_on && !b->er_on) {
    kb_zero_own<<<dim3(grid_for(b->maxn), b->B), kBlock, 0, b->stream>>>(b->d_rep, b->ax, b->ay,
                                                                        b->az);
    k_fill_d<<<grid_for((long)b->B * b->pair_blocks), kBlock, 0, b->stream>>>(
        b->part, (long)b->B * b->pair_blocks, 0.0);
  } else {
    PairArgs a;
    a.n = 0;
    a.nt = b->nt;
    a.cap = b->cap;
    a.x = a.y = a.z = a.q = nullptr;
    a.minv = b->d_minv;
    a.type = a.id = a.mol = nullptr;
    a.nl = b->nl;
    a.nn = b->nn;
    a.order = nullptr;
    a.xq = b->xq;
    a.tid = b->tid;
    const int nt2 = b->nt * b->nt;
    a.c12 = b->d_lj;
    a.c6 = b->d_lj + nt2;
    a.cshift = b->d_lj + 2 * nt2;
    a.lj_rc2 = b->lj_rc * b->lj_rc;
    a.alpha = b->er_alpha;
    a.a2 = b->er_alpha * b->er_alpha;
    a.ke = b->er_ke;
    a.two_a_rpi = 2 * b->er_alpha * CAV_PIS_INV;
    a.lambda = nullptr;
    a.ax = b->ax;
    a.ay = b->ay;
    a.az = b->az;
    a.part = b->part;
    a.accumulate = 0;
    a.rep = b->d_rep;
    if (b->precision == CAV_FP64)
      batch_pairs_t<double>(b, a, g);
    else
      batch_pairs_t<float>(b, a, g);
  }
  kb_reduce_rows<<<b->B, kBlock, 0, b->stream>>>(b->pair_blocks, b->part, b->vir);
  BAT_HIP(hipGetLastError());
  return CAV_OK;
}

}  // namespace cav
