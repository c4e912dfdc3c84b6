// Ghost images, cell binning, rebuild test and Verlet-list construction on
// gfx950.  Every kernel that depends on the per-step rebuild decision reads the
// device flag and returns early, so a step never waits for the host.
#include <algorithm>

#include "cav_internal.hpp"

namespace cav {

// ------------------------------------------------------------------ scans --

// Exclusive scan of one value per thread over a 256-thread block: wave64
// prefix sums with DPP-backed shuffles, then the 4 wave totals through LDS.
__device__ __forceinline__ int block_exclusive_scan(int v, int *total)
{
  __shared__ int wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int t = __shfl_up(inc, o, 64);
    if (lane >= o) inc += t;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wave) base += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  if (total) *total = tot;
  return base + inc - v;
}

// Per-block sums of `count(i)`; the functor recomputes the per-item count.
template <class F>
__global__ void k_block_sums(long nitems, F count, int *blk, const int *cond)
{
  if (cond && !*cond) return;
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  int v = i < nitems ? count(i) : 0;
  int tot;
  block_exclusive_scan(v, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// Exclusive scan of the block sums in one block; writes the grand total.  Each
// thread sums a contiguous run of ceil(nblk / 256) entries, one block scan
// combines the runs, and each thread then writes its run's prefixes: one
// pass with two barriers instead of one block scan per 256 entries.
__global__ void k_scan_blocks(int nblk, int *blk, int *total, const int *cond)
{
  if (cond && !*cond) return;
  const int per = (nblk + kBlock - 1) / kBlock;
  const int b0 = threadIdx.x * per, b1 = min(nblk, b0 + per);
  int s = 0;
  for (int i = b0; i < b1; ++i) s += blk[i];
  int tot;
  int ex = block_exclusive_scan(s, &tot);
  for (int i = b0; i < b1; ++i) {
    const int v = blk[i];
    blk[i] = ex;
    ex += v;
  }
  if (threadIdx.x == 0) *total = tot;
}

// ----------------------------------------------------------------- ghosts --

struct GhostRule {
  const double *x, *y, *z;
  double xl, xu, yl, yu, zl, zu;
  int bx, by, bz;
  __device__ __forceinline__ void sides(long i, int &sx, int &sy, int &sz) const
  {
    // exchange_ghost_single_md_domain.cpp:70-101: strict tests, times the bc flag
    sx = (x[i] < xl ? -1 : (x[i] > xu ? 1 : 0)) * bx;
    sy = (y[i] < yl ? -1 : (y[i] > yu ? 1 : 0)) * by;
    sz = (z[i] < zl ? -1 : (z[i] > zu ? 1 : 0)) * bz;
  }
  __device__ __forceinline__ int operator()(long i) const
  {
    int sx, sy, sz;
    sides(i, sx, sy, sz);
    const int a = sx != 0, b = sy != 0, c = sz != 0;
    return a + b + c + (a & b) + (a & c) + (b & c) + (a & b & c);
  }
};

// Writes the images of owned atom i at ghost slots [off, off+count): order x,
// y, z, xy, xz, yz, xyz (exchange_ghost_single_md_domain.cpp:167-215); ghost g
// lives at index n+g of every per-atom array.
__global__ void k_ghost_scatter(int n, GhostRule rule, double Lx, double Ly, double Lz,
                                const int *blk, double *x, double *y, double *z, int *type,
                                int *id, int *mol, double *q, int *gsrc)
{
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  const int cnt = i < n ? rule(i) : 0;
  const int off = block_exclusive_scan(cnt, nullptr) + blk[blockIdx.x];
  if (i >= n || cnt == 0) return;
  int sx, sy, sz;
  rule.sides(i, sx, sy, sz);
  const double px = x[i], py = y[i], pz = z[i];
  const double gx = px - sx * Lx, gy = py - sy * Ly, gz = pz - sz * Lz;
  const int t = type[i], d = id[i], m = mol[i];
  const double qq = q[i];
  int g = off;
  auto put = [&](double a, double b, double c) {
    const long s = (long)n + g;
    x[s] = a;
    y[s] = b;
    z[s] = c;
    type[s] = t;
    id[s] = d;
    mol[s] = m;
    q[s] = qq;
    gsrc[g] = (int)i;
    ++g;
  };
  if (sx) put(gx, py, pz);
  if (sy) put(px, gy, pz);
  if (sz) put(px, py, gz);
  if (sx && sy) put(gx, gy, pz);
  if (sx && sz) put(gx, py, gz);
  if (sy && sz) put(px, gy, gz);
  if (sx && sy && sz) put(gx, gy, gz);
}

int launch_ghosts(cav_ctx *c)
{
  GhostRule r;
  r.x = c->x;
  r.y = c->y;
  r.z = c->z;
  r.xl = c->lo[0] + c->ghost_cutoff;
  r.xu = c->hi[0] - c->ghost_cutoff;
  r.yl = c->lo[1] + c->ghost_cutoff;
  r.yu = c->hi[1] - c->ghost_cutoff;
  r.zl = c->lo[2] + c->ghost_cutoff;
  r.zu = c->hi[2] - c->ghost_cutoff;
  r.bx = c->bc[0];
  r.by = c->bc[1];
  r.bz = c->bc[2];
  const int nb = grid_for(c->n);
  k_block_sums<<<nb, kBlock, 0, c->stream>>>(c->n, r, c->blk, nullptr);
  k_scan_blocks<<<1, kBlock, 0, c->stream>>>(nb, c->blk, c->flags + F_NGHOST, nullptr);
  k_ghost_scatter<<<nb, kBlock, 0, c->stream>>>(c->n, r, c->L[0], c->L[1], c->L[2], c->blk, c->x,
                                                c->y, c->z, c->type, c->id, c->mol, c->q,
                                                c->gsrc);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// ------------------------------------------------------------ rebuild test --

// Lane `src`'s double, as a scalar.
__device__ __forceinline__ double lane_d(double v, int src)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// The w word of a bin-ordered candidate (cpos): the atom index in the low 32
// bits and its bin-order slot in the high 32 bits, so the list build can
// record either.
__device__ __forceinline__ double cpos_tag(int j, long t)
{
  return __longlong_as_double((long long)((unsigned long long)(unsigned)t << 32 | (unsigned)j));
}
__device__ __forceinline__ int cpos_index(double w) { return (int)(__double_as_longlong(w) & 0xffffffffll); }
__device__ __forceinline__ int cpos_slot(double w) { return (int)(__double_as_longlong(w) >> 32); }

// A double that is the same in every lane, moved to scalar registers.
__device__ __forceinline__ double uniform_d(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double dot_rn(double a, double b, double c)
{
  // Vector::operator* as the reference evaluates it: (a*a + b*b) + c*c, every
  // product and sum rounded on its own (hipcc contracts to FMA by default).
#pragma clang fp contract(off)
  return (a * a + b * b) + c * c;
}

// Neighborlist::update_is_needed (neighborlist.cpp:91-139): size change, or any
// owned/ghost displacement^2 (ghosts compared by index) above the threshold.
__global__ void k_rebuild_check(int n, long cap_total, const double *x, const double *y,
                                const double *z, const double *xo, const double *yo,
                                const double *zo, double thr_sq, int rebuild_test, int *flags)
{
  if (flags[F_REBUILD]) return;
  const int G = flags[F_NGHOST], Gold = flags[F_NGHOST_OLD];
  if (rebuild_test || G != Gold) {
    if (blockIdx.x == 0 && threadIdx.x == 0) flags[F_REBUILD] = 1;
    return;
  }
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= (long)n + G || a >= cap_total) return;
  const double dx = xsub(x[a], xo[a]), dy = xsub(y[a], yo[a]), dz = xsub(z[a], zo[a]);
  if (dot_rn(dx, dy, dz) > thr_sq) flags[F_REBUILD] = 1;
}

int launch_rebuild_check(cav_ctx *c)
{
  const long tot = (long)c->n + c->gcap;
  k_rebuild_check<<<grid_for(tot), kBlock, 0, c->stream>>>(c->n, tot, c->x, c->y, c->z, c->xo,
                                                          c->yo, c->zo, c->thr_sq,
                                                          c->rebuild_test, c->flags);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// ------------------------------------------------------------------- cells --

struct BinRule {
  double lox, loy, loz, cinv;
  int nbx, nby, nbz;
  int shift;  // 0: the reference grid; 1: the grid with a ghost layer past every face
  // binlist_index + neigh_bin_index (neighborlist.cpp:408-436): floor, clamp,
  // x-fastest linear index.  With the ghost layer the index is the
  // reference's plus one and the grid has two more cells per axis: ghosts
  // then lie in their own cells instead of clamped into the edge cells, and
  // those farther than one cell from the box, which no owned atom reaches,
  // share the outermost layer.
  __device__ __forceinline__ int operator()(double px, double py, double pz) const
  {
    int ix = (int)floor(xmul(xsub(px, lox), cinv)) + shift;
    int iy = (int)floor(xmul(xsub(py, loy), cinv)) + shift;
    int iz = (int)floor(xmul(xsub(pz, loz), cinv)) + shift;
    ix = ix < 0 ? 0 : (ix >= nbx ? nbx - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= nby ? nby - 1 : iy);
    iz = iz < 0 ? 0 : (iz >= nbz ? nbz - 1 : iz);
    return ix + nbx * iy + nbx * nby * iz;
  }
};

static BinRule bin_rule(const cav_ctx *c, bool ghost_layer)
{
  BinRule b;
  b.lox = c->lo[0];
  b.loy = c->lo[1];
  b.loz = c->lo[2];
  b.cinv = c->cinv;
  b.shift = ghost_layer ? 1 : 0;
  b.nbx = c->nb[0] + 2 * b.shift;
  b.nby = c->nb[1] + 2 * b.shift;
  b.nbz = c->nb[2] + 2 * b.shift;
  return b;
}

__global__ void k_zero_int(int *p, long n, const int *cond)
{
  if (cond && !*cond) return;
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Atoms are binned in the bin order of the last list build (`border`, a
// permutation of its n + G_old slots; atoms past them in index order; entries
// of atoms that no longer exist skipped), so lanes next to each other mostly
// fall into one bin and a wave adds one count per run of equal bins instead
// of one per atom.  The bins' contents are the same in any order: the fill
// below uses the same order and every bin is sorted by index afterwards.
__device__ __forceinline__ long bin_order_atom(long t, long tot, long told, const int *border)
{
  const long a = t < told ? (long)border[t] : t;
  return a >= 0 && a < tot ? a : -1;
}

// Runs of equal bins (b >= 0) among consecutive lanes: the length of the run
// this lane heads (0 if it heads none); `hm` returns the wave's run heads.
__device__ __forceinline__ int bin_run(int b, unsigned long long &hm)
{
  const int lane = threadIdx.x & 63;
  const int pb = __shfl_up(b, 1, 64);
  const bool brk = lane == 0 || b != pb;
  const unsigned long long bm = __ballot(brk);
  hm = __ballot(brk && b >= 0);
  if (!(brk && b >= 0)) return 0;
  const unsigned long long above = bm & ~((2ull << lane) - 1ull);
  return (above ? __ffsll((long long)above) - 1 : 64) - lane;
}

__global__ void k_bin_count(int n, long cap_total, BinRule rule, const double *x, const double *y,
                            const double *z, int *bin_of, int *bin_cnt, const int *flags,
                            const int *cond, const int *border)
{
  if (cond && !*cond) return;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  const long tot = (long)n + flags[F_NGHOST];
  const long told = flags[F_NGHOST_OLD] < 0 ? 0 : (long)n + flags[F_NGHOST_OLD];
  const long a = t < cap_total && t < (tot > told ? tot : told) ? bin_order_atom(t, tot, told, border) : -1;
  const int b = a >= 0 ? rule(x[a], y[a], z[a]) : -1;
  if (a >= 0) bin_of[a] = b;
  unsigned long long hm;
  const int len = bin_run(b, hm);
  if (len) atomicAdd(bin_cnt + b, len);
}

struct CountOf {
  const int *cnt;
  __device__ __forceinline__ int operator()(long i) const { return cnt[i]; }
};

__global__ void k_bin_start(int nbins, const int *cnt, const int *blk, int *start, const int *cond)
{
  if (cond && !*cond) return;
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  const int v = i < nbins ? cnt[i] : 0;
  const int off = block_exclusive_scan(v, nullptr) + blk[blockIdx.x];
  if (i < nbins) start[i] = off;
  if (i == nbins - 1) start[nbins] = off + v;
}

__global__ void k_bin_fill(int n, long cap_total, const int *bin_of, const int *start, int *cur,
                           int *items, const int *flags, const int *cond, const int *border)
{
  if (cond && !*cond) return;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  const long tot = (long)n + flags[F_NGHOST];
  const long told = flags[F_NGHOST_OLD] < 0 ? 0 : (long)n + flags[F_NGHOST_OLD];
  const long a = t < cap_total && t < (tot > told ? tot : told) ? bin_order_atom(t, tot, told, border) : -1;
  const int b = a >= 0 ? bin_of[a] : -1;
  unsigned long long hm;
  const int len = bin_run(b, hm);
  int base = 0;
  if (len) base = atomicAdd(cur + b, len);
  // every lane of a run takes the next slot after its head's
  const int lane = threadIdx.x & 63;
  const unsigned long long mine = hm & ((2ull << lane) - 1ull);
  const int head = mine ? 63 - __clzll((long long)mine) : lane;
  base = __shfl(base, head, 64);
  if (b >= 0) items[start[b] + base + (lane - head)] = (int)a;
}

// Insertion order of update_cell_list (owned ascending, then ghosts n+g
// ascending, neighborlist.cpp:457-471) is ascending index order: every bin is
// sorted after the (atomic, unordered) fill.
// One bin per wave (kSortWaves bins per workgroup, no block barrier): a bin of
// at most 64 atoms is sorted by rank (each lane counts the smaller indices of
// the others through shuffles; indices are distinct), a larger one with a
// bitonic network in the wave's LDS, and beyond kSortWaveMax atoms by lane 0.
constexpr int kSortWaves = 4, kSortWaveMax = 2048;

__device__ void sort_bin_wave(int s, int e, int *items, int *v)
{
  const int lane = threadIdx.x & 63;
  const int m = e - s;
  if (m <= 1) return;
  if (m <= 64) {
    const int x = lane < m ? items[s + lane] : 0x7fffffff;
    int rank = 0;
    for (int k = 0; k < m; ++k) rank += __shfl(x, k, 64) < x ? 1 : 0;
    if (lane < m) items[s + rank] = x;
    return;
  }
  if (m <= kSortWaveMax) {
    int p2 = 128;
    while (p2 < m) p2 <<= 1;
    for (int t = lane; t < p2; t += 64) v[t] = t < m ? items[s + t] : 0x7fffffff;
    __syncwarp();
    const int half = p2 >> 1;
    for (int k = 2; k <= p2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int t = lane; t < half; t += 64) {
          const int lo = ((t & ~(j - 1)) << 1) | (t & (j - 1));
          const int hi = lo | j;
          const int a = v[lo], c = v[hi];
          if ((a > c) == ((lo & k) == 0)) {
            v[lo] = c;
            v[hi] = a;
          }
        }
        __syncwarp();
      }
    }
    for (int t = lane; t < m; t += 64) items[s + t] = v[t];
    return;
  }
  if (lane != 0) return;
  for (int i = s + 1; i < e; ++i) {
    const int val = items[i];
    int j = i - 1;
    while (j >= s && items[j] > val) {
      items[j + 1] = items[j];
      --j;
    }
    items[j + 1] = val;
  }
}

__global__ __launch_bounds__(64 * kSortWaves) void k_bin_sort(int nbins, const int *start, int *items,
                                                              const int *cond)
{
  __shared__ int v_[kSortWaves][kSortWaveMax];
  if (cond && !*cond) return;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = blockIdx.x * kSortWaves + wv;
  if (b >= nbins) return;
  sort_bin_wave(start[b], start[b + 1], items, v_[wv]);
}

// (x, y, z, index and slot) in bin order, for contiguous candidate staging in
// the list build; on the list build's own binning (list_cond set) also the
// bin order the list keeps: slot -> atom (litems) and atom -> slot (binpos)
__global__ void k_cell_pack(long cap_total, int n, const int *items, const double *x, const double *y,
                            const double *z, double4 *cpos, const int *flags, const int *cond,
                            int *litems, int *binpos, const int *list_cond)
{
  if (cond && !*cond) return;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long)n + flags[F_NGHOST] || t >= cap_total) return;
  const int j = items[t];
  cpos[t] = make_double4(x[j], y[j], z[j], cpos_tag(j, t));
  if (litems && *list_cond) {
    litems[t] = j;
    binpos[j] = (int)t;
  }
}

// Owned atoms in cell order: stable compaction of bin_items (ghost entries
// dropped).  Consecutive lanes then walk overlapping stencils.
__global__ void k_order_scatter(long cap_total, int n, const int *items, const int *blk, int *order,
                                const int *flags, const int *cond)
{
  if (cond && !*cond) return;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  const long tot = (long)n + flags[F_NGHOST];
  const int a = (t < tot && t < cap_total) ? items[t] : n;
  const int f = a < n ? 1 : 0;
  const int off = block_exclusive_scan(f, nullptr) + blk[blockIdx.x];
  if (f) order[off] = a;
}

struct OwnedFlagBounded {
  const int *items;
  const int *flags;
  int n;
  __device__ __forceinline__ int operator()(long t) const
  {
    return (t < (long)n + flags[F_NGHOST] && items[t] < n) ? 1 : 0;
  }
};

int launch_bins(cav_ctx *c, bool conditional, bool for_list)
{
  const int *cond = conditional ? c->flags + F_REBUILD : nullptr;
  // every list build keeps its bin order (the bin-order pair gathers and the
  // cell-staged kernel read it, and the next binning visits atoms in it)
  const bool keep = for_list;
  const long tot = (long)c->n + c->gcap;
  const bool xg = for_list && list_on_ghost_grid(c);
  const BinRule rule = bin_rule(c, xg);
  const int nbins = xg ? c->xnbins : c->nbins;
  k_zero_int<<<grid_for(nbins), kBlock, 0, c->stream>>>(c->bin_cnt, nbins, cond);
  k_bin_count<<<grid_for(tot), kBlock, 0, c->stream>>>(c->n, tot, rule, c->x, c->y, c->z,
                                                      c->bin_of, c->bin_cnt, c->flags, cond,
                                                      c->litems);
  const int nb = grid_for(nbins);
  CountOf cnt{c->bin_cnt};
  k_block_sums<<<nb, kBlock, 0, c->stream>>>(nbins, cnt, c->scan_tmp, cond);
  k_scan_blocks<<<1, kBlock, 0, c->stream>>>(nb, c->scan_tmp, c->flags + F_COUNT - 1, cond);
  k_bin_start<<<nb, kBlock, 0, c->stream>>>(nbins, c->bin_cnt, c->scan_tmp, c->bin_start, cond);
  k_zero_int<<<grid_for(nbins), kBlock, 0, c->stream>>>(c->bin_cnt, nbins, cond);
  k_bin_fill<<<grid_for(tot), kBlock, 0, c->stream>>>(c->n, tot, c->bin_of, c->bin_start,
                                                     c->bin_cnt, c->bin_items, c->flags, cond,
                                                     c->litems);
  k_bin_sort<<<(nbins + kSortWaves - 1) / kSortWaves, 64 * kSortWaves, 0, c->stream>>>(
      nbins, c->bin_start, c->bin_items, cond);
  k_cell_pack<<<grid_for(tot), kBlock, 0, c->stream>>>(tot, c->n, c->bin_items, c->x, c->y, c->z,
                                                      c->cpos, c->flags, cond,
                                                      keep ? c->litems : nullptr,
                                                      keep ? c->binpos : nullptr,
                                                      c->flags + F_REBUILD);
  const int nbo = grid_for(tot);
  OwnedFlagBounded of{c->bin_items, c->flags, c->n};
  k_block_sums<<<nbo, kBlock, 0, c->stream>>>(tot, of, c->ord_blk, cond);
  k_scan_blocks<<<1, kBlock, 0, c->stream>>>(nbo, c->ord_blk, c->flags + F_COUNT - 2, cond);
  k_order_scatter<<<nbo, kBlock, 0, c->stream>>>(tot, c->n, c->bin_items, c->ord_blk, c->order,
                                                 c->flags, cond);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// ------------------------------------------------------------- Verlet list --

// Stencil candidates staged in LDS per pass (8 KB of a wave's LDS at 256).
// Measured at N=262144, rho=0.3 (list build, ms): 128 -> 0.218, 192 -> 0.208,
// 256 -> 0.204, 512 -> 0.257, 1024 -> 0.470: fewer candidates per pass, more
// waves per CU; the 512 batched replicas step 4.72 ms at 256 against 5.27 at 512.
#ifndef CAV_LIST_CAND
#define CAV_LIST_CAND 256
#endif
constexpr int kCand = CAV_LIST_CAND;

// Full Verlet rows: for owned i every owned j != i and every ghost within
// cutoff_extra (strict, dr = pos_i - pos_j rounded as the reference rounds it,
// neighborlist.cpp:531-566).  The reference's half list is the subset with
// owned j > i plus every ghost; both directions are kept so that forces are
// gathered per atom without atomics.  Rows are row-major: nl[i*cap + k].
// With nlb the entries are (also) written as bin-order slots; with nl null
// only the slots are written and F_NL_STALE tells the readers of the index
// rows to fill them first (launch_nl_fill).
//
// One wave per cell: the candidates of the cell's de-duplicated stencil are
// staged in LDS (in stencil order, each cell in bin order), then the cell's
// owned atoms are tested kAPI at a time against 64 candidates per iteration
// (each staged candidate is read once for kAPI independent tests).  Candidates
// come from the cell-ordered packed copy (x, y, z, index) written after the
// bin sort, so staging reads are contiguous per stencil cell; passing
// lanes append in lane order through a ballot and a prefix popcount, so rows
// are written coalesced and in a deterministic order.
#ifndef CAV_LIST_API
#define CAV_LIST_API 4
#endif
constexpr int kAPI = CAV_LIST_API;  // owned atoms tested per staged candidate

// Cells per workgroup, one wave each (measured at N=262144: 1 -> 0.307 ms,
// 2 -> 0.318, 4 -> 0.323, 8 -> 0.374; the build is not occupancy bound).
// Cells go to workgroups in plain order: giving each XCD a contiguous run of
// cells (as the pair kernel does with atoms) cut the fetched bytes from 90 to
// 34 MB per build but made it slower, 0.251 -> 0.288 ms.
#ifndef CAV_LIST_WPB
#define CAV_LIST_WPB 1
#endif
constexpr int kListWaves = CAV_LIST_WPB;

static int list_grid(long cells) { return (int)((cells + kListWaves - 1) / kListWaves); }

// The row outputs a build writes (index rows nl, staging positions nll,
// bin-order slots nlb) are template flags, so the append path of the inner
// loop carries only the stores the launch needs.
template <bool NL, bool NLL, bool NLB>
__global__ __launch_bounds__(64 * kListWaves) void k_list_wave(int n, int nbins, int cap, double rce2,
                                                  const double4 *cpos, const int *st_off,
                                                  const int *st, const int *start,
                                                  const int *items, int *nl, int *nn, int *flags,
                                                  const int *cond, const RepDev *rep,
                                                  const int *cell_rep, unsigned short *nll,
                                                  int *nlb)
{
  // Each wave builds the rows of one cell on its own (wave-local LDS, no
  // block barrier).  The wave index is read as a scalar so that everything
  // derived from the cell stays in scalar registers.
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = (int)((long)blockIdx.x * kListWaves + wv);
  if (b >= nbins) return;
  // batched replicas: this cell's replica decides the owned range, the row base
  // and the flags; entries are stored replica-local (owned i, ghost n+g)
  int base = 0, obase = 0;
  if (rep) {
    const int r = cell_rep[b];
    base = rep[r].base;
    n = rep[r].n;
    obase = rep[r].obase;
    flags += 16 * r;
    cond = flags + F_REBUILD;
    start += r;  // every replica's cell starts carry one closing entry
  }
  if (cond && !*cond) return;
  if (!rep && b == 0 && threadIdx.x == 0) flags[F_NL_STALE] = NL ? 0 : 1;
  // staged candidates (x, y, z, index and slot) in the cell's stencil order; a stage is
  // padded to a multiple of 64 with far-away entries so every lane tests one
  __shared__ double4 sc_[kListWaves][kCand];
  double4 *sc = sc_[wv];
  const int lane = threadIdx.x & 63;
  const int s0 = start[b], e0 = start[b + 1];
  // owned atoms come first in a sorted bin
  int no = 0;
  for (int t = s0; t < e0; t += 64) {
    const int v = (t + lane < e0) ? items[t + lane] - base : n;
    no += __popcll(__ballot(v < n));
  }
  if (no == 0) return;
  const int sb = st_off[b], ns = st_off[b + 1] - sb;  // <= 27 stencil cells
  // padding and empty owned slots sit at +-1e300: their squared distance to
  // anything is +inf, so they never pass the cutoff test
  const double4 far_cand = make_double4(1e300, 1e300, 1e300, cpos_tag(-1, -1));
  // the stencil's cell ranges, loaded by lanes 0..ns-1 at once, and their
  // running offsets in staging order (stencil cells in order, each in bin order)
  __shared__ int s_base_[kListWaves][32], s_off_[kListWaves][33];
  int *s_base = s_base_[wv], *s_off = s_off_[wv];
  int total;
  {
    int cb = 0, cn = 0;
    if (lane < ns) {
      const int cell = st[sb + lane];
      cb = start[cell];
      cn = start[cell + 1] - cb;
    }
    int incl = cn;
    for (int o = 1; o < 32; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    if (lane < ns) {
      s_base[lane] = cb;
      s_off[lane + 1] = incl;
    }
    if (lane == 0) s_off[0] = 0;
    total = __shfl(incl, ns > 0 ? ns - 1 : 0, 64);
    if (ns == 0) total = 0;
    __syncwarp();
  }
  for (int og = 0; og < no; og += 64) {
    const int gsize = (no - og) < 64 ? (no - og) : 64;
    // lane t holds owned atom og+t (position, slot) and its row count; the
    // kAPI atoms of a group are read from those lanes, never from memory
    const double4 mine = lane < gsize ? cpos[s0 + og + lane]
                                      : make_double4(-1e300, -1e300, -1e300, cpos_tag(-1, -1));
    int cnt = 0;
    for (int mbase = 0; mbase < total; mbase += kCand) {
      // stage candidates mbase .. mbase+m-1 of the flattened stencil: each
      // lane finds its cells by walking the offsets forward, all loads at once
      const int m = (total - mbase) < kCand ? (total - mbase) : kCand;
      int c = 0;
      for (int t = lane; t < m; t += 64) {
        const int g = mbase + t;
        while (s_off[c + 1] <= g) ++c;
        sc[t] = cpos[s_base[c] + g - s_off[c]];
      }
      const int mpad = (m + 63) & ~63;
      if (m + lane < mpad) sc[m + lane] = far_cand;
      __syncwarp();
      for (int o0 = 0; o0 < gsize; o0 += kAPI) {
        // the kAPI owned atoms are the same for every lane: positions, slots,
        // row offsets and row counters in scalar registers (empty slots of
        // the group carry the far sentinel and never pass)
        int ia[kAPI], c0[kAPI];
        long row[kAPI];
        double px[kAPI], py[kAPI], pz[kAPI];
#pragma unroll
        for (int u = 0; u < kAPI; ++u) {
          const int src = o0 + u;  // < 64: o0 is a multiple of kAPI below 64
          ia[u] = __builtin_amdgcn_readlane(cpos_index(mine.w), src);
          px[u] = lane_d(mine.x, src);
          py[u] = lane_d(mine.y, src);
          pz[u] = lane_d(mine.z, src);
          c0[u] = __builtin_amdgcn_readlane(cnt, src);
          row[u] = (long)(obase + ia[u] - base) * cap;
        }
        for (int cc = 0; cc < mpad; cc += 64) {
          const int idx = cc + lane;
          const double4 q = sc[idx];
          const int j = cpos_index(q.w);
#pragma unroll
          for (int u = 0; u < kAPI; ++u) {
            const double dx = xsub(px[u], q.x), dy = xsub(py[u], q.y), dz = xsub(pz[u], q.z);
            const bool pass = j != ia[u] && dot_rn(dx, dy, dz) < rce2;
            // the ballot of the compare itself (a scalar mask), the lane's
            // rank among the passing lanes by mbcnt, and a scalar row counter
            const unsigned long long mask = __builtin_amdgcn_ballot_w64(pass);
            if (mask) {
              const int cu = c0[u];
              const unsigned pre = __builtin_amdgcn_mbcnt_hi(
                  (unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
              if (pass && cu + (int)pre < cap) {
                // uniform row pointers plus a 32-bit lane offset
                const long at0 = row[u] + cu;
                if constexpr (NL) {
                  int *rp = nl + at0;
                  rp[pre] = j - base;
                }
                if constexpr (NLL) {
                  unsigned short *lp = nll + at0;
                  lp[pre] = (unsigned short)(mbase + idx);  // position in the stencil staging order
                }
                if constexpr (NLB) {
                  int *bp = nlb + at0;
                  bp[pre] = cpos_slot(q.w) - base;  // bin-order slot of the candidate
                }
              }
              c0[u] = cu + (int)__popcll(mask);
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kAPI; ++u) cnt = lane == o0 + u ? c0[u] : cnt;
      }
      __syncwarp();
    }
    if (lane < gsize) {
      const int i = cpos_index(mine.w) - base + obase;
      nn[i] = cnt < cap ? cnt : cap;
      // the longest row is needed only to regrow the list after an overflow:
      // one device-wide atomic per overflowing row, none otherwise
      if (cnt > cap) {
        flags[F_OVERFLOW] = 1;
        atomicMax(flags + F_MAXNB, cnt);
      }
    }
  }
}

// k_list_wave for the outputs that are not null (index rows, staging
// positions, bin-order slots: the combinations the single system and the
// batched replicas use)
static void launch_list_wave(hipStream_t st, int n, int nbins, int cap, double rce2,
                             const double4 *cpos, const int *st_off, const int *stc,
                             const int *start, const int *items, int *nl, int *nn, int *flags,
                             const int *cond, const RepDev *rep, const int *cell_rep,
                             unsigned short *nll, int *nlb)
{
  const dim3 g(list_grid(nbins)), blk(64 * kListWaves);
#define CAV_LISTW(A, B, C)                                                                      \
  k_list_wave<A, B, C><<<g, blk, 0, st>>>(n, nbins, cap, rce2, cpos, st_off, stc, start, items, \
                                          nl, nn, flags, cond, rep, cell_rep, nll, nlb)
  if (nlb && !nl)
    CAV_LISTW(false, false, true);
  else if (nlb)
    CAV_LISTW(true, false, true);
  else if (nll)
    CAV_LISTW(true, true, false);
  else
    CAV_LISTW(true, false, false);
#undef CAV_LISTW
}

// all_atom_test_function (neighborlist.cpp:161-191): every owned j != i and
// every ghost, no distance test.
__global__ void k_list_all(int n, int cap, int *nl, int *nn, int *flags, const int *cond)
{
  if (cond && !*cond) return;
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int tot = n + flags[F_NGHOST];
  int cnt = 0;
  for (int j = 0; j < tot; ++j) {
    if (j == i) continue;
    if (cnt < cap) nl[(long)i * cap + cnt] = j;
    ++cnt;
  }
  nn[i] = cnt < cap ? cnt : cap;
  if (cnt > cap) {
    flags[F_OVERFLOW] = 1;
    atomicMax(flags + F_MAXNB, cnt);
  }
}

// The cells the list was built from, kept for the cell-staged pair kernel:
// later re-binnings (the LJ cell walk, energy and point queries) rewrite
// bin_start/bin_items, while the list's staging positions refer to these.
// binpos is the inverse permutation (slot -> position in bin order).
__global__ void k_list_bins(int n, int nbins, long cap_total, const int *start, const int *items,
                            int *lstart, int *litems, int *binpos, const int *flags, const int *cond)
{
  if (cond && !*cond) return;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t <= nbins) lstart[t] = start[t];
  if (t < (long)n + flags[F_NGHOST] && t < cap_total) {
    const int a = items[t];
    litems[t] = a;
    binpos[a] = (int)t;
  }
}

int launch_list(cav_ctx *c, bool conditional)
{
  const int *cond = conditional ? c->flags + F_REBUILD : nullptr;
  // staging positions (and the cells they refer to) only for the cell-staged
  // pair kernel; a build that writes them makes them valid for its list, one
  // that does not invalidates them (known at once for an unconditional build,
  // after the flags are read back for a conditional one)
  const bool with_nll = c->pair_kind == CAV_PAIRS_CELLS && c->nll && !c->all_atom_test;
  c->nll_pending = with_nll;
  if (!conditional) c->nll_built = with_nll;
  // bin-order slots of the entries, for the row kernel's gathers from the
  // bin-ordered copies: valid on the same terms
  const bool with_nlb = c->pair_kind == CAV_PAIRS_AUTO && c->nlb && !c->all_atom_test;
  // the grid the list build's binning used (launch_bins with for_list)
  const bool xg = list_on_ghost_grid(c);
  c->nlb_pending = with_nlb;
  if (!conditional) c->nlb_built = with_nlb;
  if (c->all_atom_test)
    k_list_all<<<grid_for(c->n), kBlock, 0, c->stream>>>(c->n, c->nl_cap, c->nl, c->nn, c->flags,
                                                        cond);
  else
    launch_list_wave(c->stream, c->n, xg ? c->xnbins : c->nbins, c->nl_cap, c->rce2, c->cpos,
                     xg ? c->d_xst_off : c->d_st_off, xg ? c->d_xst : c->d_st, c->bin_start,
                     c->bin_items, with_nlb ? nullptr : c->nl, c->nn, c->flags, cond, nullptr,
                     nullptr, with_nll ? c->nll : nullptr, with_nlb ? c->nlb : nullptr);
  // (the bin order itself, litems and binpos, was kept by the list build's
  // binning; the cell-staged kernel also needs the cell starts)
  if (with_nll) {
    const long tot = (long)c->n + c->gcap;
    k_list_bins<<<grid_for(tot > c->nbins + 1 ? tot : c->nbins + 1), kBlock, 0, c->stream>>>(
        c->n, c->nbins, tot, c->bin_start, c->bin_items, c->lstart, c->litems, c->binpos, c->flags,
        cond);
  }
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// Index rows from the bin-order slots of the last list build, when that build
// wrote the slots only (F_NL_STALE); the flag is cleared by a kernel of its own
// after every entry is written.
__global__ void k_nl_fill(long entries, int cap, long slots, const int *nlb, const int *litems,
                          const int *nn, int *nl, const int *flags)
{
  if (!flags[F_NL_STALE]) return;
  const long e = (long)blockIdx.x * kBlock + threadIdx.x;
  if (e >= entries || e % cap >= nn[e / cap]) return;
  const int t = nlb[e];
  if (t >= 0 && t < slots) nl[e] = litems[t];
}

__global__ void k_nl_fresh(int *flags) { flags[F_NL_STALE] = 0; }

int launch_nl_fill(cav_ctx *c)
{
  if (!c->nlb || !c->nl || c->all_atom_test) return CAV_OK;
  const long entries = (long)c->n * c->nl_cap;
  k_nl_fill<<<grid_for(entries), kBlock, 0, c->stream>>>(entries, c->nl_cap, (long)c->n + c->gcap,
                                                          c->nlb, c->litems, c->nn, c->nl,
                                                          c->flags);
  k_nl_fresh<<<1, 1, 0, c->stream>>>(c->flags);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// pos_old = pos, ghost_pos_old = ghost pos (neighborlist.cpp:302-303,568-569).
__global__ void k_save_old(int n, long cap_total, const double *x, const double *y, const double *z,
                           double *xo, double *yo, double *zo, int *flags, const int *cond)
{
  if (cond && !*cond) return;
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  const int G = flags[F_NGHOST];
  if (a == 0) flags[F_NGHOST_OLD] = G;
  if (a >= (long)n + G || a >= cap_total) return;
  xo[a] = x[a];
  yo[a] = y[a];
  zo[a] = z[a];
}

int launch_save_old(cav_ctx *c, bool conditional)
{
  if (c->all_atom_test) return CAV_OK;  // the reference returns before saving
  const int *cond = conditional ? c->flags + F_REBUILD : nullptr;
  const long tot = (long)c->n + c->gcap;
  k_save_old<<<grid_for(tot), kBlock, 0, c->stream>>>(c->n, tot, c->x, c->y, c->z, c->xo, c->yo,
                                                     c->zo, c->flags, cond);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}


// ---------------------------------------------------------- batched replicas --
//
// The same steps for every replica of a cav_batch at once: grid row
// blockIdx.y is the replica for per-atom work; per-cell work maps a global
// cell to its replica.  Each replica's decisions (ghost count, rebuild) live
// in its own 16 flag words.

__device__ __forceinline__ GhostRule ghost_rule_of(const RepDev &r, double gc, const double *x,
                                                   const double *y, const double *z)
{
  GhostRule g;
  g.x = x + r.base;
  g.y = y + r.base;
  g.z = z + r.base;
  g.xl = r.lo[0] + gc;
  g.xu = r.hi[0] - gc;
  g.yl = r.lo[1] + gc;
  g.yu = r.hi[1] - gc;
  g.zl = r.lo[2] + gc;
  g.zu = r.hi[2] - gc;
  g.bx = g.by = g.bz = 1;
  return g;
}

__global__ void kb_ghost_sums(const RepDev *rep, double gc, const double *x, const double *y,
                              const double *z, int nblk, int *gblk)
{
  const RepDev &r = rep[blockIdx.y];
  const GhostRule g = ghost_rule_of(r, gc, x, y, z);
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  int tot;
  block_exclusive_scan(i < r.n ? g(i) : 0, &tot);
  if (threadIdx.x == 0) gblk[(long)blockIdx.y * nblk + blockIdx.x] = tot;
}

__global__ void kb_ghost_scan(const RepDev *rep, int nblk, int *gblk, int *flags)
{
  const int rr = blockIdx.x;
  int *blk = gblk + (long)rr * nblk;
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nblk; base += kBlock) {
    const int i = base + threadIdx.x;
    const int v = i < nblk ? blk[i] : 0;
    int tot;
    const int ex = block_exclusive_scan(v, &tot);
    if (i < nblk) blk[i] = ex + carry;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    int *f = flags + 16 * rr;
    f[F_NGHOST] = carry < rep[rr].gcap ? carry : rep[rr].gcap;
    if (carry > rep[rr].gcap) f[F_GOVER] = 1;
  }
}

__global__ void kb_ghost_scatter(const RepDev *rep, double gc, int nblk, const int *gblk, double *x,
                                 double *y, double *z, int *type, int *id, double *q)
{
  const RepDev &r = rep[blockIdx.y];
  const GhostRule rule = ghost_rule_of(r, gc, x, y, z);
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  const int cnt = i < r.n ? rule(i) : 0;
  const int off = block_exclusive_scan(cnt, nullptr) + gblk[(long)blockIdx.y * nblk + blockIdx.x];
  if (i >= r.n || cnt == 0) return;
  int sx, sy, sz;
  rule.sides(i, sx, sy, sz);
  const long own = r.base + i;
  const double px = x[own], py = y[own], pz = z[own];
  const double gx = px - sx * r.L[0], gy = py - sy * r.L[1], gz = pz - sz * r.L[2];
  const int t = type[own];
  const double qq = q[own];
  int g = off;
  auto put = [&](double a, double b, double c) {
    if (g < r.gcap) {
      const long s = (long)r.base + r.n + g;
      x[s] = a;
      y[s] = b;
      z[s] = c;
      type[s] = t;
      id[s] = (int)i;
      q[s] = qq;
    }
    ++g;
  };
  if (sx) put(gx, py, pz);
  if (sy) put(px, gy, pz);
  if (sz) put(px, py, gz);
  if (sx && sy) put(gx, gy, pz);
  if (sx && sz) put(gx, py, gz);
  if (sy && sz) put(px, gy, gz);
  if (sx && sy && sz) put(gx, gy, gz);
}

__global__ void kb_rebuild_check(const RepDev *rep, const double *x, const double *y,
                                 const double *z, const double *xo, const double *yo,
                                 const double *zo, double thr_sq, int *flags)
{
  const RepDev &r = rep[blockIdx.y];
  int *f = flags + 16 * blockIdx.y;
  if (f[F_REBUILD]) return;
  const int G = f[F_NGHOST], Gold = f[F_NGHOST_OLD];
  if (G != Gold) {
    if (blockIdx.x == 0 && threadIdx.x == 0) f[F_REBUILD] = 1;
    return;
  }
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= (long)r.n + G) return;
  const long s = r.base + a;
  const double dx = xsub(x[s], xo[s]), dy = xsub(y[s], yo[s]), dz = xsub(z[s], zo[s]);
  if (dot_rn(dx, dy, dz) > thr_sq) f[F_REBUILD] = 1;
}

__global__ void kb_set_rebuild(int B, int *flags)
{
  const int r = blockIdx.x * kBlock + threadIdx.x;
  if (r < B) flags[16 * r + F_REBUILD] = 1;
}

// A replica's cells: the grid with the ghost layer (r.nb counts it).
__device__ __forceinline__ BinRule bin_rule_of(const RepDev &r, double cinv)
{
  BinRule b;
  b.lox = r.lo[0];
  b.loy = r.lo[1];
  b.loz = r.lo[2];
  b.cinv = cinv;
  b.shift = 1;
  b.nbx = r.nb[0];
  b.nby = r.nb[1];
  b.nbz = r.nb[2];
  return b;
}

// cell counts of the rebuilding replicas (zeroed first by kb_zero_cells)
__global__ void kb_zero_cells(const RepDev *rep, const int *flags, int *cnt, int *cur)
{
  const RepDev &r = rep[blockIdx.y];
  if (!flags[16 * blockIdx.y + F_REBUILD]) return;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < r.nbins) cnt[r.cbase + k] = cur[r.cbase + k] = 0;
}

__global__ void kb_bin_count(const RepDev *rep, double cinv, const double *x, const double *y,
                             const double *z, const int *flags, int *bin_of, int *cnt)
{
  const RepDev &r = rep[blockIdx.y];
  const int *f = flags + 16 * blockIdx.y;
  if (!f[F_REBUILD]) return;
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= (long)r.n + f[F_NGHOST]) return;
  const long s = r.base + a;
  const int c = r.cbase + bin_rule_of(r, cinv)(x[s], y[s], z[s]);
  bin_of[s] = c;
  atomicAdd(cnt + c, 1);
}

// per replica: cell starts (absolute into the item array, replica r's starts
// at index cbase + r .. cbase + r + nbins)
__global__ void kb_bin_start(const RepDev *rep, const int *flags, const int *cnt, int *start)
{
  const int rr = blockIdx.x;
  const RepDev &r = rep[rr];
  if (!flags[16 * rr + F_REBUILD]) return;
  __shared__ int carry;
  if (threadIdx.x == 0) carry = r.base;
  __syncthreads();
  int *st = start + r.cbase + rr;
  for (int k0 = 0; k0 < r.nbins; k0 += kBlock) {
    const int k = k0 + threadIdx.x;
    const int v = k < r.nbins ? cnt[r.cbase + k] : 0;
    int tot;
    const int ex = block_exclusive_scan(v, &tot);
    if (k < r.nbins) st[k] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) st[r.nbins] = carry;
}

__global__ void kb_bin_fill(const RepDev *rep, const int *flags, const int *bin_of,
                            const int *cell_rep, const int *start, int *cur, int *items)
{
  const RepDev &r = rep[blockIdx.y];
  const int *f = flags + 16 * blockIdx.y;
  if (!f[F_REBUILD]) return;
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= (long)r.n + f[F_NGHOST]) return;
  const long s = r.base + a;
  const int c = bin_of[s];
  items[start[c + blockIdx.y] + atomicAdd(cur + c, 1)] = (int)s;
}

// After the sort a cell lists its owned atoms first (they have the lower
// slots); their count goes to ocnt, taken before the sort (it is order-free).
__global__ __launch_bounds__(64 * kSortWaves) void kb_bin_sort(const RepDev *rep, int ncells,
                                                               const int *cell_rep,
                                                               const int *flags, const int *start,
                                                               int *items, int *ocnt)
{
  __shared__ int v_[kSortWaves][kSortWaveMax];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = blockIdx.x * kSortWaves + wv;
  if (c >= ncells) return;
  const int rr = cell_rep[c];
  if (!flags[16 * rr + F_REBUILD]) return;
  const int s0 = start[c + rr], e0 = start[c + rr + 1];
  const int lane = threadIdx.x & 63, own_end = rep[rr].base + rep[rr].n;
  int no = 0;
  for (int t = s0; t < e0; t += 64) no += __popcll(__ballot(t + lane < e0 && items[t + lane] < own_end));
  if (lane == 0) ocnt[c] = no;
  sort_bin_wave(s0, e0, items, v_[wv]);
}

// per replica: each cell's first owned atom in cell order (exclusive scan of
// the owned counts over the replica's cells)
__global__ void kb_owned_start(const RepDev *rep, const int *flags, const int *ocnt, int *ostart)
{
  const int rr = blockIdx.x;
  const RepDev &r = rep[rr];
  if (!flags[16 * rr + F_REBUILD]) return;
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int k0 = 0; k0 < r.nbins; k0 += kBlock) {
    const int k = k0 + threadIdx.x;
    const int v = k < r.nbins ? ocnt[r.cbase + k] : 0;
    int tot;
    const int ex = block_exclusive_scan(v, &tot);
    if (k < r.nbins) ostart[r.cbase + k] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
}

// (x, y, z, index and slot) in bin order for the list build, and the bin
// order the pair kernel reads through: slot -> local atom (litems), local atom
// -> slot (binpos), owned atoms in cell order (order)
__global__ void kb_cell_pack(const RepDev *rep, const int *flags, const int *items,
                             const int *bin_of, const int *start, const int *ostart,
                             const double *x, const double *y, const double *z, double4 *cpos,
                             int *litems, int *binpos, int *order)
{
  const int rr = blockIdx.y;
  const RepDev &r = rep[rr];
  const int *f = flags + 16 * rr;
  if (!f[F_REBUILD]) return;
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= (long)r.n + f[F_NGHOST]) return;
  const long t = r.base + a;
  const int j = items[t];
  cpos[t] = make_double4(x[j], y[j], z[j], cpos_tag(j, t));
  const int lj = j - r.base;
  litems[t] = lj;
  binpos[j] = (int)a;
  if (lj < r.n) {
    // owned atoms come first in their sorted cell
    const int c = bin_of[j];
    order[r.obase + ostart[c] + (int)(t - start[c + rr])] = lj;
  }
}

__global__ void kb_save_old(const RepDev *rep, const double *x, const double *y, const double *z,
                            double *xo, double *yo, double *zo, int *flags)
{
  const RepDev &r = rep[blockIdx.y];
  int *f = flags + 16 * blockIdx.y;
  if (!f[F_REBUILD]) return;
  const int G = f[F_NGHOST];
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a == 0) f[F_NGHOST_OLD] = G;
  if (a >= (long)r.n + G) return;
  const long s = r.base + a;
  xo[s] = x[s];
  yo[s] = y[s];
  zo[s] = z[s];
}

// Ghosts, rebuild decision, cells and Verlet rows of every replica.
int batch_neighbor(cav_batch *b, bool forced)
{
  const dim3 own_grid(b->scan_blocks, b->B);
  kb_ghost_sums<<<own_grid, kBlock, 0, b->stream>>>(b->d_rep, b->ghost_cutoff, b->x, b->y, b->z,
                                                     b->scan_blocks, b->gblk);
  kb_ghost_scan<<<b->B, kBlock, 0, b->stream>>>(b->d_rep, b->scan_blocks, b->gblk, b->flags);
  kb_ghost_scatter<<<own_grid, kBlock, 0, b->stream>>>(b->d_rep, b->ghost_cutoff, b->scan_blocks,
                                                        b->gblk, b->x, b->y, b->z, b->type, b->id,
                                                        b->q);
  if (!b->nl_on) return CAV_OK;
  int maxslots = 0, maxcells = 0;
  for (const RepDev &r : b->reps) {
    maxslots = std::max(maxslots, r.n + r.gcap);
    maxcells = std::max(maxcells, r.nbins);
  }
  const dim3 slot_grid(grid_for(maxslots), b->B), cell_grid(grid_for(maxcells), b->B);
  if (forced)
    kb_set_rebuild<<<grid_for(b->B), kBlock, 0, b->stream>>>(b->B, b->flags);
  else
    kb_rebuild_check<<<slot_grid, kBlock, 0, b->stream>>>(b->d_rep, b->x, b->y, b->z, b->xo, b->yo,
                                                           b->zo, b->thr_sq, b->flags);
  kb_zero_cells<<<cell_grid, kBlock, 0, b->stream>>>(b->d_rep, b->flags, b->bin_cnt, b->bin_cnt + b->ncells);
  kb_bin_count<<<slot_grid, kBlock, 0, b->stream>>>(b->d_rep, b->cinv, b->x, b->y, b->z, b->flags,
                                                     b->bin_of, b->bin_cnt);
  kb_bin_start<<<b->B, kBlock, 0, b->stream>>>(b->d_rep, b->flags, b->bin_cnt, b->bin_start);
  kb_bin_fill<<<slot_grid, kBlock, 0, b->stream>>>(b->d_rep, b->flags, b->bin_of, b->cell_rep,
                                                    b->bin_start, b->bin_cnt + b->ncells, b->bin_items);
  kb_bin_sort<<<(b->ncells + kSortWaves - 1) / kSortWaves, 64 * kSortWaves, 0, b->stream>>>(
      b->d_rep, b->ncells, b->cell_rep, b->flags, b->bin_start, b->bin_items, b->ocnt);
  kb_owned_start<<<b->B, kBlock, 0, b->stream>>>(b->d_rep, b->flags, b->ocnt, b->ostart);
  kb_cell_pack<<<slot_grid, kBlock, 0, b->stream>>>(b->d_rep, b->flags, b->bin_items, b->bin_of,
                                                     b->bin_start, b->ostart, b->x, b->y, b->z,
                                                     b->cpos, b->litems, b->binpos, b->order);
  launch_list_wave(b->stream, 0, b->ncells, b->cap, b->rce2, b->cpos, b->st_off, b->st,
                   b->bin_start, b->bin_items, b->nl, b->nn, b->flags, nullptr, b->d_rep,
                   b->cell_rep, nullptr, b->nlb);
  kb_save_old<<<slot_grid, kBlock, 0, b->stream>>>(b->d_rep, b->x, b->y, b->z, b->xo, b->yo, b->zo,
                                                    b->flags);
  BAT_HIP(hipGetLastError());
  return CAV_OK;
}

}  // namespace cav
