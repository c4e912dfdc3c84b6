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
  // (a*a + b*b) + c*c with every product and sum rounded on its own, as the
  // reference's Vector::operator* (hipcc contracts to FMA by default)
#pragma clang fp contract(off)
  return (a * a + b * b) + c * c;
}

// erfc(x) = t * P(t) * exp(-x^2), t = 1/(1 + p x): the Hastings form of
// Abramowitz & Stegun 7.1.26 with a degree-8 polynomial whose coefficients
// were fitted here by weighted least squares on [0, 6] (relative error 1.3e-8
// in exact arithmetic, 2.6e-7 evaluated in FP32).  exp(-x^2) is shared with the
// Gaussian term of the Ewald force.
__device__ __forceinline__ float erfc_hastings(float x, float expmx2)
{
  const float t = __builtin_amdgcn_rcpf(fmaf(0.47047f, x, 1.0f));  // v_rcp_f32, 1 ulp
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
  // rows as bin-order slots (or nullptr): neighbours are then read from the
  // bin-ordered copies, where the neighbours of cell-ordered atoms lie close
  // together, with the atom index from the list build's bin order
  const int *nlb;
  const double4 *xqs;
  const int2 *tids;
  const int *litems;
};

// Neighbours whose loads a lane issues together in the row-gather kernel
// (measured at N=262144 with gathers by atom index: FP32 2/4/6/8 ->
// 0.187/0.166/0.157/0.164 ms, FP64 0.204/0.220/0.208/0.254 ms, the FP64 pair
// math needs the registers; from the bin-ordered copies, FP32 4/6/8/10 ->
// 0.111/0.108/0.115/0.126 ms).
template <typename T>
constexpr int pair_u() { return sizeof(T) == 8 ? 2 : 6; }

// LJ per type pair (c12, c6, cshift) staged in LDS as T at kernel start:
// 3 nt^2 values, converted once instead of per pair.
template <typename T>
__device__ __forceinline__ void stage_lj_tables(const PairArgs &a, int nt2, T *tab)
{
  for (int t = threadIdx.x; t < 3 * nt2; t += blockDim.x)
    tab[t] = (T)(t < nt2 ? a.c12[t] : (t < 2 * nt2 ? a.c6[t - nt2] : a.cshift[t - 2 * nt2]));
}

template <typename T>
__device__ __forceinline__ T recip(T x)
{
  if constexpr (sizeof(T) == 8)
    return 1.0 / x;
  else
    return __builtin_amdgcn_rcpf(x);  // mixed form: v_rcp_f32, 1 ulp
}

// Force coefficient of one listed pair (force on i = coef * (r_j - r_i)), in T =
// double (the reference's arithmetic) or float (mixed form); adds the pair's
// Ewald_r virial term for the reference's list entries.
//   LJ: lj_cell_list.cpp:141-142 skips dr^2 > rc^2; shifted-force coefficients.
//   ER: electrostatic_ewald_r.cpp:162-211, every listed pair, exact zero skipped,
//       lambda of same-molecule pairs (0 skips the pair).
template <typename T, bool LJ, bool ER, bool LAM>
__device__ __forceinline__ T pair_coef(const PairArgs &a, const T *tab, int nt2, int i, int j,
                                       int ti, int idi, int moli, const int2 &tj, T qi, T keqi,
                                       const double4 &pi4, const double4 &pj, T &dx, T &dy, T &dz,
                                       double &vir)
{
  const double dxd = pj.x - pi4.x, dyd = pj.y - pi4.y, dzd = pj.z - pi4.z;
  dx = (T)dxd;
  dy = (T)dyd;
  dz = (T)dzd;
  T r2;
  if constexpr (sizeof(T) == 8)
    r2 = dot_rn_p(dxd, dyd, dzd);  // the exact rounding order of the reference test
  else
    r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
  T coef = 0;
  // type-pair entry as the reference indexes it: [lower owned index][partner]
  const int p = (j < a.n && j < i) ? tj.x * a.nt + ti : ti * a.nt + tj.x;
  if constexpr (LJ) {
    if (!(r2 > (T)a.lj_rc2)) {
      const T ri2 = recip(r2);
      const T ri4 = ri2 * ri2, ri8 = ri4 * ri4;
      coef += -tab[p] * (ri8 * ri4 * ri2) + tab[nt2 + p] * ri8 + tab[2 * nt2 + p];
    }
  }
  if constexpr (ER) {
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
  return coef;
}

// Deterministic per-block sum of the virial: fixed shuffle tree per wave, then
// the waves in order.
__device__ __forceinline__ void block_virial(double vir, double *part)
{
  __shared__ double s_red[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) vir += __shfl_xor(vir, o, 64);
  if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = vir;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) v += s_red[w];
    *part = v;
  }
}

// Row-gather pair kernel: kLPA lanes share one atom's row (coalesced row
// reads, balanced work); each lane fetches pair_u<T>() indices, then their (x, y,
// z, q) and (type, id), before computing, so the gathers overlap.  The
// partial forces of an atom's lanes are summed with a fixed shuffle tree, so
// FP64 results stay reproducible.
template <typename T, bool LJ, bool ER, bool LAM, bool BINS>
__global__ __launch_bounds__(kBlock) void k_pairs(PairArgs a)
{
  extern __shared__ double dyn_pair[];
  T *tab = reinterpret_cast<T *>(dyn_pair);
  const int nt2 = LJ ? a.nt * a.nt : 0;
  stage_lj_tables(a, nt2, tab);
  __syncthreads();
  if (a.rep) {  // one replica per grid row: shift every array to it
    const RepDev &r = a.rep[blockIdx.y];
    const int base = r.base, obase = r.obase;
    a.n = r.n;
    a.xq += base;
    a.tid += base;
    a.nl += (long)obase * a.cap;
    a.nn += obase;
    if (a.order) a.order += obase;
    if constexpr (BINS) {
      a.nlb += (long)obase * a.cap;
      a.xqs += base;
      a.tids += base;
      a.litems += base;
    }
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
    const int moli = LAM ? a.mol[i] : 0;
    const T qi = (T)pi4.w;
    const T keqi = (T)(a.ke) * qi;
    const int cnt = a.nn[i];
    const int *row = (BINS ? a.nlb : a.nl) + (long)i * a.cap;
    constexpr int U = pair_u<T>();
    for (int k0 = sub; k0 < cnt; k0 += U * kLPA) {
      int jj[U];
      double4 pj[U];
      int2 tj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u * kLPA;
        jj[u] = k < cnt ? row[k] : -1;
      }
      if constexpr (BINS) {
        // entries past the row read slot 0 and are skipped below
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int pp = jj[u] < 0 ? 0 : jj[u];
          pj[u] = a.xqs[pp];
          tj[u] = a.tids[pp];
          jj[u] = jj[u] < 0 ? -1 : a.litems[pp];
        }
      } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int j = jj[u] < 0 ? i : jj[u];
          pj[u] = a.xq[j];
          tj[u] = a.tid[j];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (jj[u] < 0) break;
        T dx, dy, dz;
        const T coef = pair_coef<T, LJ, ER, LAM>(a, tab, nt2, i, jj[u], ti, tii.y, moli, tj[u], qi,
                                                 keqi, pi4, pj[u], dx, dy, dz, vir);
        fx = fma(coef, dx, fx);
        fy = fma(coef, dy, fy);
        fz = fma(coef, dz, fz);
      }
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
  block_virial(vir, a.part + blockIdx.x);
}

// ------------------------------------- force_field lj and short-range kernel --
//
// The Verlet-list force fields that are not fused into k_pairs: the reference's
// force_field lj (lj.cpp:341-483) and Electrostatic_short_range
// (electrostatic_short_range.cpp:133-258).  Each adds its own pass over the
// same full rows (the reference runs each field as its own loop), with the
// row-gather layout of k_pairs: kLPA lanes per atom, partial forces summed by
// a fixed shuffle tree, one virial partial per block.
struct ListArgs {
  // force_field lj per type pair ([row type][partner type]): coef(r) =
  // -c12 r^-14 + c6 r^-8 + cshift, pair skipped when r^2 > rc2 (rc2 < 0 for
  // eps = 0, which the reference skips)
  const double *ljl;  // [c12 | c6 | cshift | rc2] blocks of nt^2
  int ljl_ignore_intra, ljl_jump;
  double ljl_jump_tol;
  // Electrostatic_short_range
  double sr_rc, sr_rc2, sr_ke, sr_beta1, sr_ext[3];
  const double *sr_lambda;  // nt^2, or nullptr
};

// Force coefficient of one listed pair for the list-based fields.  The
// virial of each field counts the reference's list entries (owned j > i or a
// ghost) with id_i < id_j, from the field's own dr^2 (lj clamps it with
// jump_tol).
template <typename T, bool LJL, bool SR>
__device__ __forceinline__ T list_coef(const PairArgs &a, const ListArgs &b, int nt2, int i, int j,
                                       int ti, int idi, int moli, int molj, const int2 &tj, T qi,
                                       const double4 &pi4, const double4 &pj, T &dx, T &dy, T &dz,
                                       double &vir)
{
  const double dxd = pj.x - pi4.x, dyd = pj.y - pi4.y, dzd = pj.z - pi4.z;
  dx = (T)dxd;
  dy = (T)dyd;
  dz = (T)dzd;
  T r2;
  if constexpr (sizeof(T) == 8)
    r2 = dot_rn_p(dxd, dyd, dzd);
  else
    r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
  const int p = (j < a.n && j < i) ? tj.x * a.nt + ti : ti * a.nt + tj.x;
  const bool counted = (j >= a.n || j > i) && idi < tj.y;
  T coef = 0;
  if constexpr (LJL) {
    if (!(b.ljl_ignore_intra && moli == molj) && !(r2 > (T)b.ljl[3 * nt2 + p])) {
      T d2 = r2;
      if (b.ljl_jump && d2 < (T)b.ljl_jump_tol) d2 = (T)b.ljl_jump_tol;
      const T ri2 = recip(d2);
      const T ri4 = ri2 * ri2, ri8 = ri4 * ri4;
      const T cl = -(T)b.ljl[p] * (ri8 * ri4 * ri2) + (T)b.ljl[nt2 + p] * ri8 + (T)b.ljl[2 * nt2 + p];
      coef += cl;
      if (counted) vir += -(double)cl * (double)d2;
    }
  }
  if constexpr (SR) {
    T lam = 1;
    bool skip = false;
    if (b.sr_lambda && moli == molj) {
      const T l = (T)b.sr_lambda[p];
      if (l == 0) skip = true;
      lam = l;
    }
    if (!skip && r2 < (T)b.sr_rc2) {
      const T qj = (T)pj.w;
      T cs;
      if constexpr (sizeof(T) == 8) {
        const double rn = sqrt(r2);
        cs = -lam * b.sr_ke * (double)qi * qj / (r2 * rn) * (1.0 - pow(rn / b.sr_rc, b.sr_beta1));
      } else {
        const float rinv = __frsqrt_rn(r2);
        const float rn = r2 * rinv;
        const float sh = 1.0f - powf(rn * (float)(1.0 / b.sr_rc), (float)b.sr_beta1);
        cs = -lam * (float)b.sr_ke * qi * qj * (rinv * rinv * rinv) * sh;
      }
      coef += cs;
      if (counted) vir += -(double)cs * (double)r2;
    }
  }
  return coef;
}

template <typename T, bool LJL, bool SR>
__global__ __launch_bounds__(kBlock) void k_pairs_list(PairArgs a, ListArgs b)
{
  const int nt2 = a.nt * a.nt;
  const long per = gridDim.x >> 3;  // XCD-aware order, as k_pairs
  const long blk = (long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const long gid = blk * kBlock + threadIdx.x;
  const int t = (int)(gid / kLPA), sub = (int)(gid % kLPA);
  double vir = 0.0;
  T fx = 0, fy = 0, fz = 0;
  int i = 0, ti = 0;
  T qi = 0;
  if (t < a.n) {
    i = a.order ? a.order[t] : t;
    const double4 pi4 = a.xq[i];
    const int2 tii = a.tid[i];
    ti = tii.x;
    const int moli = a.mol[i];
    qi = (T)pi4.w;
    const int cnt = a.nn[i];
    const int *row = a.nl + (long)i * a.cap;
    constexpr int U = 2;
    for (int k0 = sub; k0 < cnt; k0 += U * kLPA) {
      int jj[U];
      double4 pj[U];
      int2 tj[U];
      int mj[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int k = k0 + u * kLPA;
        jj[u] = k < cnt ? row[k] : -1;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = jj[u] < 0 ? i : jj[u];
        pj[u] = a.xq[j];
        tj[u] = a.tid[j];
        mj[u] = a.mol[j];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (jj[u] < 0) break;
        T dx, dy, dz;
        const T coef = list_coef<T, LJL, SR>(a, b, nt2, i, jj[u], ti, tii.y, moli, mj[u], tj[u], qi,
                                             pi4, pj[u], dx, dy, dz, vir);
        fx = fma(coef, dx, fx);
        fy = fma(coef, dy, fy);
        fz = fma(coef, dz, fz);
      }
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
    double ex = 0, ey = 0, ez = 0;
    if constexpr (SR) {  // the external field, once per atom after its pairs
      ex = b.sr_ext[0] * (double)qi * mi;
      ey = b.sr_ext[1] * (double)qi * mi;
      ez = b.sr_ext[2] * (double)qi * mi;
    }
    if (a.accumulate) {
      a.ax[i] += (double)fx * mi + ex;
      a.ay[i] += (double)fy * mi + ey;
      a.az[i] += (double)fz * mi + ez;
    } else {
      a.ax[i] = (double)fx * mi + ex;
      a.ay[i] = (double)fy * mi + ey;
      a.az[i] = (double)fz * mi + ez;
    }
  }
  block_virial(vir, a.part + blockIdx.x);
}

// ----------------------------------------------------- cell-staged pair kernel --
//
// One workgroup per cell: the cell's de-duplicated 27-stencil is staged in LDS
// in the order the list build walked it (stencil cells in order, each in bin
// order), from the bin-ordered copies (x, y, z, q) and (type, id) refreshed
// every step; each Verlet entry carries its position in that order, so every
// j of the cell's owned atoms is an LDS read.  Lanes per atom adapt to the
// cell's owned count; the partial forces of an atom's lanes are summed with a
// fixed shuffle tree.  A cell whose stencil exceeds the LDS tile falls back to
// gathering its rows from the global copies.  The tile is sized per launch
// from the largest stencil met so far (a flag word the kernel raises).
constexpr int kPairCand = 1024;
constexpr int kPairLds = sizeof(double4) + sizeof(int2) + sizeof(int);  // bytes per staged atom

struct CellPairArgs {
  int nbins, cap, seen;
  const int *start, *st_off, *st, *items;
  const double4 *xqs;
  const int2 *tids;
  const unsigned short *nll;
  int *flags;
};

template <typename T, bool LJ, bool ER>
__global__ __launch_bounds__(kBlock) void k_pairs_cells(PairArgs a, CellPairArgs ca)
{
  // staged stencil: (x, y, z, q)[cap], then (type, id)[cap], then slot index[cap],
  // then the LJ tables (3 nt^2 values of T)
  extern __shared__ double4 lds[];
  double4 *lq = lds;
  int2 *lt = reinterpret_cast<int2 *>(lq + ca.cap);
  int *lj = reinterpret_cast<int *>(lt + ca.cap);
  T *tab = reinterpret_cast<T *>(lj + ca.cap);
  const int nt2 = LJ ? a.nt * a.nt : 0;
  stage_lj_tables(a, nt2, tab);
  __shared__ int s_base[32], s_off[33];
  __shared__ int sh_nc, sh_self;
  __shared__ T s_f[3][kBlock];  // partial forces of the atoms' lanes
  const long per = gridDim.x >> 3;  // XCD-aware cell order, as k_pairs
  const int b = (int)((long)(blockIdx.x & 7) * per + (blockIdx.x >> 3));
  double vir = 0.0;
  if (b < ca.nbins) {
    const int s0 = ca.start[b], e0 = ca.start[b + 1];
    const int sb = ca.st_off[b], ns = ca.st_off[b + 1] - sb;  // <= 27 stencil cells
    if (threadIdx.x < 64) {
      // the stencil's cell ranges and their running offsets in staging order
      const int lane = threadIdx.x;
      if (lane == 0) sh_self = -1;
      int cell = -1, cb = 0, cn = 0;
      if (lane < ns) {
        cell = ca.st[sb + lane];
        cb = ca.start[cell];
        cn = ca.start[cell + 1] - cb;
      }
      int incl = cn;
      for (int o = 1; o < 32; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
      }
      if (lane < ns) {
        s_base[lane] = cb;
        s_off[lane + 1] = incl;
        if (cell == b) sh_self = incl - cn;  // this cell's own atoms in the staging order
      }
      if (lane == 0) s_off[0] = 0;
      if (lane == (ns > 0 ? ns - 1 : 0)) {
        sh_nc = ns > 0 ? incl : 0;
        if (ns > 0 && incl > ca.seen) atomicMax(ca.flags + F_MAXST, incl);
      }
    }
    // owned atoms come first in a sorted bin
    int no = 0;
    for (int t0 = s0; t0 < e0; t0 += kBlock) {
      const int t = t0 + (int)threadIdx.x;
      no += __syncthreads_count(t < e0 && ca.items[t] < a.n);
    }
    __syncthreads();
    const int nc = sh_nc, self = sh_self;
    const bool staged = nc <= ca.cap && self >= 0;
    if (no > 0) {
      // lanes per atom: as many as fill the block (any count; the partial
      // forces are summed through LDS in lane order)
      const int L = no >= kBlock ? 1 : (kBlock / no < 64 ? kBlock / no : 64);
      const int per_pass = kBlock / L;
      const int slot = threadIdx.x / L, sub = threadIdx.x - slot * L;
      constexpr int U = 4;  // neighbour indices fetched together
      for (int a0 = 0; a0 < no; a0 += per_pass) {
        const int ai = a0 + slot;
        const bool act = slot < per_pass && ai < no;
        // the row's first indices are fetched before the stencil is staged:
        // rows hold cap entries, so reads below cap stay in the row
        const int i = act ? ca.items[s0 + ai] : 0;
        const long row = (long)i * a.cap;
        const int cnt = act ? a.nn[i] : 0;
        int loc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int k = sub + u * L;
          loc[u] = (act && k < a.cap) ? (staged ? (int)ca.nll[row + k] : a.nl[row + k]) : 0;
        }
        if (a0 == 0) {
          if (staged) {
            for (int t = threadIdx.x; t < nc; t += kBlock) {
              int c = 0;
              while (s_off[c + 1] <= t) ++c;
              const int src = s_base[c] + t - s_off[c];
              lq[t] = ca.xqs[src];
              lt[t] = ca.tids[src];
              lj[t] = ca.items[src];
            }
          }
          __syncthreads();
        }
        T fx = 0, fy = 0, fz = 0;
        int ti = 0;
        if (act) {
          double4 pi4;
          int2 tii;
          if (staged) {
            pi4 = lq[self + ai];
            tii = lt[self + ai];
          } else {
            pi4 = a.xq[i];
            tii = a.tid[i];
          }
          ti = tii.x;
          const T qi = (T)pi4.w, keqi = (T)a.ke * qi;
          for (int k0 = sub; k0 < cnt; k0 += U * L) {
            if (k0 != sub) {
#pragma unroll
              for (int u = 0; u < U; ++u) {
                const int k = k0 + u * L;
                loc[u] = k < cnt ? (staged ? (int)ca.nll[row + k] : a.nl[row + k]) : 0;
              }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
              if (k0 + u * L >= cnt) break;
              int j;
              double4 pj;
              int2 tj;
              if (staged) {
                pj = lq[loc[u]];
                tj = lt[loc[u]];
                j = lj[loc[u]];
              } else {
                j = loc[u];
                pj = a.xq[j];
                tj = a.tid[j];
              }
              T dx, dy, dz;
              const T coef = pair_coef<T, LJ, ER, false>(a, tab, nt2, i, j, ti, tii.y, 0, tj, qi,
                                                         keqi, pi4, pj, dx, dy, dz, vir);
              fx = fma(coef, dx, fx);
              fy = fma(coef, dy, fy);
              fz = fma(coef, dz, fz);
            }
          }
        }
        s_f[0][threadIdx.x] = fx;
        s_f[1][threadIdx.x] = fy;
        s_f[2][threadIdx.x] = fz;
        __syncthreads();
        if (act && sub == 0) {
          fx = fy = fz = 0;
          for (int l = 0; l < L; ++l) {
            fx += s_f[0][threadIdx.x + l];
            fy += s_f[1][threadIdx.x + l];
            fz += s_f[2][threadIdx.x + l];
          }
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
        __syncthreads();
      }
    }
  }
  block_virial(vir, a.part + blockIdx.x);
}

// Packed copies for the pair gathers: (x, y, z, q) and (type, id).
__global__ void k_pack(int n, long cap_total, const double *x, const double *y, const double *z,
                       const double *q, const int *type, const int *id, const int *flags,
                       double4 *xq, int2 *tid, const int *binpos, double4 *xqs, int2 *tids)
{
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= (long)n + flags[F_NGHOST] || a >= cap_total) return;
  const double4 v = make_double4(x[a], y[a], z[a], q[a]);
  const int2 w = make_int2(type[a], id[a]);
  xq[a] = v;
  tid[a] = w;
  // the same in bin order, for the cell-staged pair kernel: only slots that
  // existed at the last list build have a bin position
  if (binpos && a < (long)n + flags[F_NGHOST_OLD]) {
    const int t = binpos[a];
    if (t >= 0 && t < cap_total) {
      xqs[t] = v;
      tids[t] = w;
    }
  }
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
  if (count > c->npart) return fail(c, CAV_ERR_STATE, "partial-sum buffer smaller than the grid");
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
  a.nlb = nullptr;
  a.xqs = nullptr;
  a.tids = nullptr;
  a.litems = nullptr;
  return a;
}

static int pair_grid(const cav_ctx *c) { return (grid_for((long)c->n * kLPA) + 7) / 8 * 8; }

template <typename T, bool BINS>
static void launch_pairs_tb(cav_ctx *c, const PairArgs &a, bool lj, bool er, bool lam)
{
  const int g = pair_grid(c);
  const size_t lds = 3 * sizeof(T) * a.nt * a.nt;  // LJ tables
  if (lj && er && lam)
    k_pairs<T, true, true, true, BINS><<<g, kBlock, lds, c->stream>>>(a);
  else if (lj && er)
    k_pairs<T, true, true, false, BINS><<<g, kBlock, lds, c->stream>>>(a);
  else if (lj)
    k_pairs<T, true, false, false, BINS><<<g, kBlock, lds, c->stream>>>(a);
  else if (er && lam)
    k_pairs<T, false, true, true, BINS><<<g, kBlock, 0, c->stream>>>(a);
  else
    k_pairs<T, false, true, false, BINS><<<g, kBlock, 0, c->stream>>>(a);
}

// rows as bin-order slots (a.nlb) or as atom indices, each its own kernel
template <typename T>
static void launch_pairs_t(cav_ctx *c, const PairArgs &a, bool lj, bool er, bool lam)
{
  if (a.nlb)
    launch_pairs_tb<T, true>(c, a, lj, er, lam);
  else
    launch_pairs_tb<T, false>(c, a, lj, er, lam);
}

// The cell-staged kernel runs on request (CAV_PAIRS_CELLS) where it applies:
// the cell-built list (its entries' staging positions), no per-molecule
// lambda.  Otherwise, and by default, the row-gather kernel runs (faster on
// the ion workloads measured at N=262144, rho=0.3: 0.165 vs 0.232 ms mixed).
// The row kernel reads neighbours through the list build's bin-order slots
// when the current list carries them (CAV_PAIRS_AUTO; lists built from cells).
bool sorted_rows(const cav_ctx *c)
{
  return c->pair_kind == CAV_PAIRS_AUTO && !c->all_atom_test && c->nlb && c->nlb_built;
}

bool cell_staged(const cav_ctx *c)
{
  return c->pair_kind == CAV_PAIRS_CELLS && c->from_cells && c->build_cells && !c->all_atom_test &&
         !c->lambda_set && c->nll && c->nll_built;
}

static int cell_grid(const cav_ctx *c) { return (c->nbins + 7) / 8 * 8; }

template <typename T>
static void launch_pairs_cells_t(cav_ctx *c, const PairArgs &a, bool lj, bool er)
{
  CellPairArgs ca;
  ca.nbins = c->nbins;
  ca.start = c->lstart;
  ca.st_off = c->d_st_off;
  ca.st = c->d_st;
  ca.items = c->litems;
  ca.xqs = c->xqs;
  ca.tids = c->tids;
  ca.nll = c->nll;
  ca.flags = c->flags;
  // LDS tile: the largest stencil met so far (last synchronised flags), in
  // steps of 64 atoms; cells beyond it gather from global memory and raise it
  const int seen = c->h_flags ? c->h_flags[F_MAXST] : 0;
  ca.seen = seen;
  ca.cap = seen > 0 ? std::min(kPairCand, (seen + 63) / 64 * 64) : kPairCand;
  const size_t lds = (size_t)ca.cap * kPairLds + 3 * sizeof(T) * a.nt * a.nt;
  const int g = cell_grid(c);
  if (lj && er)
    k_pairs_cells<T, true, true><<<g, kBlock, lds, c->stream>>>(a, ca);
  else if (lj)
    k_pairs_cells<T, true, false><<<g, kBlock, lds, c->stream>>>(a, ca);
  else
    k_pairs_cells<T, false, true><<<g, kBlock, lds, c->stream>>>(a, ca);
}

// The list-based fields (force_field lj, Electrostatic_short_range) on the
// current list, one pass for the fields in use; `accumulate` false overwrites
// the accelerations and the virial (this pass is then the first force of the
// step); `pack` first refreshes the packed copies (when no fused pair pass has
// packed them this step).
int launch_pairs_list(cav_ctx *c, bool want_ljl, bool want_sr, bool accumulate, bool pack)
{
  const bool ljl = want_ljl && c->ljl_on, sr = want_sr && c->sr_on;
  if (!ljl && !sr) return CAV_OK;
  if (pack) CAV_TRY(launch_pack(c));
  CAV_TRY(launch_nl_fill(c));
  PairArgs a = pair_args(c);
  a.accumulate = accumulate ? 1 : 0;
  ListArgs b;
  b.ljl = c->d_ljl;
  b.ljl_ignore_intra = c->ljl_ignore_intra;
  b.ljl_jump = c->ljl_jump;
  b.ljl_jump_tol = c->ljl_jump_tol;
  b.sr_rc = c->sr_rc;
  b.sr_rc2 = c->sr_rc * c->sr_rc;
  b.sr_ke = c->sr_ke;
  b.sr_beta1 = c->sr_beta + 1.0;
  for (int d = 0; d < 3; ++d) b.sr_ext[d] = c->sr_ext[d];
  b.sr_lambda = c->sr_lambda_set ? c->d_sr_lambda : nullptr;
  const int g = pair_grid(c);
  if (g > c->npart) return fail(c, CAV_ERR_STATE, "partial-sum buffer smaller than the pair grid");
  timer_begin(c, c->t_pair);
  if (c->precision == CAV_FP64) {
    if (ljl && sr)
      k_pairs_list<double, true, true><<<g, kBlock, 0, c->stream>>>(a, b);
    else if (ljl)
      k_pairs_list<double, true, false><<<g, kBlock, 0, c->stream>>>(a, b);
    else
      k_pairs_list<double, false, true><<<g, kBlock, 0, c->stream>>>(a, b);
  } else {
    if (ljl && sr)
      k_pairs_list<float, true, true><<<g, kBlock, 0, c->stream>>>(a, b);
    else if (ljl)
      k_pairs_list<float, true, false><<<g, kBlock, 0, c->stream>>>(a, b);
    else
      k_pairs_list<float, false, true><<<g, kBlock, 0, c->stream>>>(a, b);
  }
  timer_end(c, c->t_pair);
  CAV_HIP(hipGetLastError());
  return reduce_partials(c, g, 0, accumulate);
}

// Repacks (x, y, z, q) and (type, id) of the current owned atoms and ghosts
// for readers of the packed copies other than the pair kernels (the adjoint).
int launch_pack(cav_ctx *c)
{
  const long tot = (long)c->n + c->gcap;
  k_pack<<<grid_for(tot), kBlock, 0, c->stream>>>(c->n, tot, c->x, c->y, c->z, c->q, c->type,
                                                 c->id, c->flags, c->xq, c->tid, nullptr, nullptr,
                                                 nullptr);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

int launch_pairs(cav_ctx *c, bool want_lj, bool want_er, bool accumulate)
{
  const bool lj = want_lj && c->lj_on && !c->lj_cellwalk, er = want_er && c->er_on;
  if (!lj && !er) return CAV_OK;
  PairArgs a = pair_args(c);
  a.accumulate = accumulate ? 1 : 0;
  const long tot = (long)c->n + c->gcap;
  const bool cells = cell_staged(c);
  // every block writes one virial partial: the buffer must cover the grid
  if ((cells ? cell_grid(c) : pair_grid(c)) > c->npart)
    return fail(c, CAV_ERR_STATE, "partial-sum buffer smaller than the pair grid");
  const bool bins = !cells && sorted_rows(c);
  if (bins) {
    a.nlb = c->nlb;
    a.xqs = c->xqs;
    a.tids = c->tids;
    a.litems = c->litems;
  } else {
    CAV_TRY(launch_nl_fill(c));
  }
  k_pack<<<grid_for(tot), kBlock, 0, c->stream>>>(c->n, tot, c->x, c->y, c->z, c->q, c->type,
                                                 c->id, c->flags, c->xq, c->tid,
                                                 cells || bins ? c->binpos : nullptr, c->xqs,
                                                 c->tids);
  timer_begin(c, c->t_pair);
  if (cells) {
    if (c->precision == CAV_FP64)
      launch_pairs_cells_t<double>(c, a, lj, er);
    else
      launch_pairs_cells_t<float>(c, a, lj, er);
  } else if (c->precision == CAV_FP64) {
    launch_pairs_t<double>(c, a, lj, er, c->lambda_set);
  } else {
    launch_pairs_t<float>(c, a, lj, er, c->lambda_set);
  }
  timer_end(c, c->t_pair);
  CAV_HIP(hipGetLastError());
  return reduce_partials(c, cells ? cell_grid(c) : pair_grid(c), 0, accumulate);
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
  int ix = (int)floor(xmul(xsub(px, c.lox), c.cinv));
  int iy = (int)floor(xmul(xsub(py, c.loy), c.cinv));
  int iz = (int)floor(xmul(xsub(pz, c.loz), c.cinv));
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
      const double dx = xsub(c.x[j], xi), dy = xsub(c.y[j], yi), dz = xsub(c.z[j], zi);
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
  CAV_TRY(launch_nl_fill(c));
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

// Packed copies of every replica's slots, by index and in the last list
// build's bin order (binpos: slot -> replica-local bin position; a changed
// ghost count forces a rebuild first, so every slot has one).
__global__ void kb_pack(const RepDev *rep, const int *flags, const double *x, const double *y,
                        const double *z, const double *q, const int *type, const int *id,
                        double4 *xq, int2 *tid, const int *binpos, double4 *xqs, int2 *tids)
{
  const RepDev &r = rep[blockIdx.y];
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= (long)r.n + flags[16 * blockIdx.y + F_NGHOST]) return;
  const long s = r.base + a;
  const double4 v = make_double4(x[s], y[s], z[s], q[s]);
  const int2 w = make_int2(type[s], id[s]);
  xq[s] = v;
  tid[s] = w;
  if (binpos) {
    const long t = r.base + binpos[s];
    xqs[t] = v;
    tids[t] = w;
  }
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

// every replica's rows are written both as indices and as bin-order slots:
// the pair kernel reads the slots, the downloads and the adjoint the indices
template <typename T>
static void batch_pairs_t(cav_batch *b, const PairArgs &a, dim3 g)
{
  const size_t lds = 3 * sizeof(T) * a.nt * a.nt;  // LJ tables
  if (b->lj_on && b->er_on)
    k_pairs<T, true, true, false, true><<<g, kBlock, lds, b->stream>>>(a);
  else if (b->lj_on)
    k_pairs<T, true, false, false, true><<<g, kBlock, lds, b->stream>>>(a);
  else
    k_pairs<T, false, true, false, true><<<g, kBlock, 0, b->stream>>>(a);
}

// LJ + Ewald_r accelerations of every replica (overwriting), virial per replica.
int batch_pairs(cav_batch *b)
{
  int maxslots = 0;
  for (const RepDev &r : b->reps) maxslots = std::max(maxslots, r.n + r.gcap);
  kb_pack<<<dim3(grid_for(maxslots), b->B), kBlock, 0, b->stream>>>(
      b->d_rep, b->flags, b->x, b->y, b->z, b->q, b->type, b->id, b->xq, b->tid,
      b->nl_on ? b->binpos : nullptr, b->xqs, b->tids);
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
    a.order = b->order;
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
    a.nlb = b->nlb;
    a.xqs = b->xqs;
    a.tids = b->tids;
    a.litems = b->litems;
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
