// Electrostatic_ewald_k on gfx950 as two complex GEMMs on MFMA tiles.
//
// The k lattice is separable: exp(i k.r) = X(nx) Y(ny) Z(nz) with
// X(n) = exp(2 pi i n x / lx) etc., and S(-k) = conj S(k) for real charges, so
// only the half lattice nx >= 0 is evaluated and each k carries a multiplicity
// m(k) = [k in K] + [-k in K] over the reference's truncated set K (the first
// (2kx)(2ky)(2kz) lexicographic vectors, electrostatic_ewald_k/initialize.cpp:
// 191,201-221).  With rows r = (nx, ny) and columns c = nz:
//
//   S pass:     S[r][c]  = sum_j (q_j X_j(nx) Y_j(ny)) Z_j(c)     (reduce over atoms)
//   force pass: U_i[r]   = sum_c g[r][c] Z_i(c),  g = m f conj(S)  (reduce over nz)
//               Uz_i[r]  = sum_c nz g[r][c] Z_i(c)
//               F_i      = C q_i sum_r (2pi nx/lx, 2pi ny/ly, 2pi/lz) Im(P_i(r) (U, U, Uz))
//               phi_i    = C sum_r Re(P_i(r) U_i[r]),  P_i(r) = X_i(nx) Y_i(ny)
//
// Complex products are real-embedded (re/im interleaved along the reduction
// axis) and run on v_mfma_f64_16x16x4f64 (plain form) or
// v_mfma_f32_16x16x4f32 (mixed form).  The S pass splits the atom axis into
// chunks; chunk partials are summed in FP64 in a fixed order, so the plain
// form is reproducible run to run.
#include <algorithm>
#include <cmath>

#include "cav_internal.hpp"

namespace cav {

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4 V;
  static __device__ __forceinline__ V op(double a, double b, V c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  typedef f4 V;
  static __device__ __forceinline__ V op(float a, float b, V c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <typename T> struct Cx { T re, im; };

// Output row held in element e of lane group lk = lane/16 of a 16x16x4 MFMA:
// the f32 form keeps four consecutive rows per lane group, the f64 form
// interleaves them (row = lk + 4e).  Columns are lane%16 in both.
template <typename T>
__device__ __forceinline__ int mfma_row(int lk, int e)
{
  return sizeof(T) == 8 ? lk + 4 * e : 4 * lk + e;
}

// ------------------------------------------------------------- phase tables --

// Per atom a record of TS complex values [X(0..kx) | Y(-ky..ky) | Z(-kz..kz)].
// The phase fraction n*x/l is reduced to [-1/2, 1/2] in FP64 before sincospi,
// so the FP32 tables (point fields and the adjoint of the mixed form) carry
// ~1e-7 absolute phase error even where n k.r reaches hundreds of radians; the
// split form's blocked tables (k_tables_h) are products of FP64 phase factors
// rounded once, within 1 FP32 ulp.
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned pack_h2(_Float16 a, _Float16 b)
{
  const h2 v = {a, b};
  return __builtin_bit_cast(unsigned, v);
}

// hi = fp16(v), lo = fp16(v - hi) for a pair of values, packed (first value low)
__device__ __forceinline__ void split_pair(float a, float b, unsigned &hi, unsigned &lo)
{
  const _Float16 ah = (_Float16)a, bh = (_Float16)b;
  hi = pack_h2(ah, bh);
  lo = pack_h2((_Float16)(a - (float)ah), (_Float16)(b - (float)bh));
}

// The same split in three instructions: hi by one packed round-to-nearest
// conversion, each half of lo by a mixed-precision fma, fp16(a * 1 - hi.x),
// reading hi's half directly (a - hi is exact in FP32, so the result equals
// split_pair's).
__device__ __forceinline__ void split_pair_mix(float a, float b, unsigned &hi, unsigned &lo)
{
  asm("v_cvt_pk_f16_f32 %0, %1, %2" : "=v"(hi) : "v"(a), "v"(b));
  unsigned l;
  asm("v_fma_mixlo_f16 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]" : "=v"(l) : "v"(a), "v"(hi));
  asm("v_fma_mixhi_f16 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
               : "+v"(l)
               : "v"(b), "v"(hi));
  lo = l;
}

// Atom-major phase tables of the FP64 form and of the FP32 tiles: per atom
// [X(0..kx) | Y(-ky..ky) | Z(-kz..kz)], TS complex values.
template <typename T>
__global__ void k_tables(long npts, int TS, int kx, int ky, int kz, double ilx, double ily,
                         double ilz, const double *x, const double *y, const double *z,
                         Cx<T> *tab)
{
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= npts * TS) return;
  const long a = t / TS;
  const int e = (int)(t - a * TS);
  const int nX = kx + 1, nY = 2 * ky + 1;
  double u;
  int m;
  if (e < nX) {
    m = e;
    u = x[a] * ilx;
  } else if (e < nX + nY) {
    m = e - nX - ky;
    u = y[a] * ily;
  } else {
    m = e - nX - nY - kz;
    u = z[a] * ilz;
  }
  double f = (double)m * u;
  f -= rint(f);
  Cx<T> v;
  if constexpr (sizeof(T) == 8) {
    double s, c;
    sincospi(2.0 * f, &s, &c);
    v.re = c;
    v.im = s;
  } else {
    float s, c;
    sincospif((float)(2.0 * f), &s, &c);
    v.re = c;
    v.im = s;
  }
  tab[t] = v;
}

// ------------------------------------------------------------------- S pass --

constexpr int kWavesMax = 8;  // waves per S-pass workgroup (upper bound)

struct SArgs {
  int n, TS, nX, nY, nZ, ky, kz, Rtiles, chunk, stage, tiles_per_group;
  const double *q;
  const void *tab;
  double *part;  // [chunk][Rpad][SC]
  int Rpad, SC;
  const int4 *items;  // batched: per workgroup column (first atom, end atom, chunk slot, replica)
  int ncols, groups;  // logical grid: atom chunks (or items) x row-tile groups
  int ct0;            // first column tile of this launch (launches cover 8 tiles each)
};

// Issue the asynchronous copy of `bytes` contiguous bytes (a multiple of 16,
// source padded) from global memory straight into LDS, 16 bytes per lane.
__device__ __forceinline__ void copy_to_lds_async(const unsigned char *src, unsigned char *dst,
                                                  int bytes, int wave, int nwaves, int lane)
{
  for (int off = wave * 1024; off < bytes; off += nwaves * 1024) {
    if (off + lane * 16 < bytes)
      __builtin_amdgcn_global_load_lds((const void *)(src + off + lane * 16),
                                       (__attribute__((address_space(3))) void *)(dst + off), 16,
                                       0, 0);
  }
}

// One workgroup: an atom chunk x a group of row tiles; wave w owns RT row
// tiles and the SCT column tiles from ct0 (wider nz axes take one launch per
// 8 column tiles).  A (rows x (atom,re/im)) is generated from
// the staged X, Y tables and charges, B ((atom,re/im) x (nz,re/im)) from Z.
// Stages of atom records are double-buffered in LDS: the next stage streams
// in from HBM (direct-to-LDS loads) while the current one feeds the MFMAs.
// In the mixed form the MFMAs accumulate in FP32 over one chunk (<= 4096
// atoms) and the chunk partials are summed in FP64.
template <typename T, int RT, int SCT>
__global__ __launch_bounds__(64 * kWavesMax) void k_sk(SArgs a)
{
  typedef typename Mfma<T>::V V;
  extern __shared__ unsigned char smem_raw[];
  const int stage = a.stage;
  const size_t stage_bytes = (size_t)stage * a.TS * sizeof(Cx<T>);
  // stage buffers: records at smem_raw + b * stage_bytes, charges after both;
  // addressed from the __shared__ array itself (LDS, not flat, accesses)
  const unsigned char *tab = reinterpret_cast<const unsigned char *>(a.tab);

  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwaves = blockDim.x >> 6;
  // XCD-aware order: workgroups L and L+8 run on the same XCD (dispatch is
  // round-robin over the 8 XCDs), so the row groups of one chunk share its
  // records through one L2
  int chunk, grp;
  {
    const int L = blockIdx.x, q = L >> 3;
    grp = q % a.groups;
    chunk = (q / a.groups) * 8 + (L & 7);
    if (chunk >= a.ncols) return;
  }
  const int col = chunk;
  const int rt0 = grp * a.tiles_per_group + wave * RT;
  const int rt_end = min(a.Rtiles, (grp + 1) * a.tiles_per_group);
  long a0 = (long)chunk * a.chunk;
  long a1 = a0 + a.chunk;
  if (a1 > a.n) a1 = a.n;
  if (a.items) {
    const int4 it = a.items[col];
    a0 = it.x;
    a1 = it.y;
    chunk = it.z;
  }

  V acc[RT][SCT];
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int s = 0; s < SCT; ++s) acc[r][s] = V{0, 0, 0, 0};

  // lane roles inside a 16x16x4 tile: A[row=l%16][k=l/16], B[k=l/16][col=l%16]
  const int lr = lane & 15, lk = lane >> 4;
  const int jj = lk >> 1, part = lk & 1;  // atom within the K-step, re/im half
  int rnx[RT], rny[RT];
  bool rvalid[RT];
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    const int row = (rt0 + r) * 16 + lr;
    rvalid[r] = (rt0 + r) < rt_end && row < a.nX * a.nY;
    rnx[r] = rvalid[r] ? row / a.nY : 0;
    rny[r] = rvalid[r] ? row % a.nY : 0;  // index into Y (ny + ky)
  }
  int zc[SCT], zt[SCT];
  bool zvalid[SCT];
#pragma unroll
  for (int s = 0; s < SCT; ++s) {
    const int col = (a.ct0 + s) * 16 + lr;
    zc[s] = col >> 1;
    zt[s] = col & 1;
    zvalid[s] = zc[s] < a.nZ;
  }
  const int offY = a.nX, offZ = a.nX + a.nY;
  const bool active = rt0 < rt_end;

  auto issue = [&](long base, int b) {
    const int cnt = (int)((a1 - base) < stage ? (a1 - base) : stage);
    const int bytes = (int)(((size_t)cnt * a.TS * sizeof(Cx<T>) + 15) / 16 * 16);
    copy_to_lds_async(tab + (size_t)base * a.TS * sizeof(Cx<T>), smem_raw + b * stage_bytes,
                      bytes, wave, nwaves, lane);
    if (threadIdx.x < stage)
      reinterpret_cast<T *>(smem_raw + 2 * stage_bytes)[b * stage + threadIdx.x] =
          threadIdx.x < cnt ? (T)a.q[base + threadIdx.x] : (T)0;
  };

  int b = 0;
  if (a0 < a1) issue(a0, 0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (long base = a0; base < a1; base += stage) {
    if (base + stage < a1) issue(base + stage, b ^ 1);  // streams in during the MFMAs below
    const int cnt = (int)((a1 - base) < stage ? (a1 - base) : stage);
    const Cx<T> *st = reinterpret_cast<const Cx<T> *>(smem_raw + b * stage_bytes);
    const T *sq = reinterpret_cast<const T *>(smem_raw + 2 * stage_bytes) + b * stage;
    if (active) {
#pragma unroll 2
      for (int ks = 0; ks < stage / 2; ++ks) {
        const int j = 2 * ks + jj;
        const bool jv = j < cnt;
        const Cx<T> *rec = st + (jv ? j : 0) * a.TS;
        const T qj = jv ? sq[j] : (T)0;
        T av[RT];
#pragma unroll
        for (int r = 0; r < RT; ++r) {
          const Cx<T> X = rec[rnx[r]], Y = rec[offY + rny[r]];
          const T pr = X.re * Y.re - X.im * Y.im;
          const T pi = X.re * Y.im + X.im * Y.re;
          av[r] = rvalid[r] ? qj * (part ? pi : pr) : (T)0;
        }
#pragma unroll
        for (int s = 0; s < SCT; ++s) {
          const Cx<T> Z = rec[offZ + (zvalid[s] ? zc[s] : 0)];
          // re column: (Zr, -Zi); im column: (Zi, Zr) along (re, im) of A
          T bv = zt[s] == 0 ? (part ? -Z.im : Z.re) : (part ? Z.re : Z.im);
          bv = (zvalid[s] && jv) ? bv : (T)0;
#pragma unroll
          for (int r = 0; r < RT; ++r) acc[r][s] = Mfma<T>::op(av[r], bv, acc[r][s]);
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    b ^= 1;
  }
  if (!active) return;
  // D[row = mfma_row(l/16, e)][col = l%16]
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    if (rt0 + r >= rt_end) break;
#pragma unroll
    for (int s = 0; s < SCT; ++s)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = (rt0 + r) * 16 + mfma_row<T>(lk, e);
        const int col = (a.ct0 + s) * 16 + lr;
        const double v = (double)acc[r][s][e];
        a.part[((long)chunk * a.Rpad + row) * a.SC + col] = v;
      }
  }
}

// --------------------------------------------------- reduction and operand --

// S = sum over chunks (fixed order, FP64); g = W conj(S); write the force-pass
// operand B[kappa][col] (kappa = 2c+s, col = rb*64 + t*16 + rr) in fragment-major
// order: for row block rb and component t, lane (kappa%4)*16 + rr holds its
// KS values (kappa/4 = 0..) contiguously, so the force pass loads them as
// 16-byte vectors.  Also the per-thread k-space energy term W |S|^2.
// Split force-pass operand of one 16-row block: 4 components x nkh K-steps x
// (hi, lo) x 64 lanes x 8 halves, then the z = 0 column's values of
// components 0 and 1 (FP32, 16 rows each).  The z = 0 column of Z is exactly
// (1, 0) for every atom, so its product is that per-row constant: it seeds
// the accumulators instead of taking a K slot, and the K axis holds the
// 2 (nZ - 1) other kappa (kappa' = 2 c' + re/im, c' = nz below kz, nz - 1 above).
__host__ __device__ constexpr int split_blk_bytes(int nkh) { return 4 * nkh * 2 * 64 * 16 + 128; }

template <typename T, typename P>
__global__ void k_sk_reduce(int nchunks, int Rpad, int SC, int R, int nZ, int kz, int KSS, int KSP,
                            const P *part, const double *W, double *sk, T *bm,
                            double *epart, double inv_sa, _Float16 *bmh, int nkh, double sb,
                            const int2 *rchunks, long bm_stride, long bmh_stride)
{
  if (rchunks) {  // batched: grid row b reduces replica b's chunks into its own S and operands
    const int b = blockIdx.y;
    part += (long)rchunks[b].x * Rpad * SC;
    nchunks = rchunks[b].y;
    W += (long)b * R * nZ;
    sk += (long)b * Rpad * 2 * nZ;
    bm += b * bm_stride;
    if (bmh) bmh += b * bmh_stride;
    epart += (long)b * gridDim.x;
  }
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double e = 0;
  if (t < Rpad * nZ) {
    const int r = t / nZ, c = t - r * nZ;
    double sr = 0, si = 0;
    for (int k = 0; k < nchunks; ++k) {
      const P *p = part + ((long)k * Rpad + r) * SC + 2 * c;
      sr += (double)p[0];
      si += (double)p[1];
    }
    sr *= inv_sa;
    si *= inv_sa;
    sk[(long)r * 2 * nZ + 2 * c] = sr;
    sk[(long)r * 2 * nZ + 2 * c + 1] = si;
    const double w = r < R ? W[(long)r * nZ + c] : 0.0;
    e = w * (sr * sr + si * si);
    const double gr = w * sr, gi = -w * si, nz = (double)(c - kz);
    const int rb = r >> 4, rr = r & 15;
    // values for kappa = 2c (s=0) and 2c+1 (s=1) of the four components
    const double v[4][2] = {{gr, -gi}, {gi, gr}, {nz * gr, -nz * gi}, {nz * gi, nz * gr}};
#pragma unroll
    for (int comp = 0; comp < 4; ++comp)
#pragma unroll
      for (int sidx = 0; sidx < 2; ++sidx) {
        const int kap = 2 * c + sidx;
        const int lanef = (kap & 3) * 16 + rr;
        const int ks = kap >> 2, sl = ks / KSS;  // K-step and its slab
        bm[((((long)sl * (Rpad >> 4) + rb) * 4 + comp) * 64 + lanef) * KSP + (ks - sl * KSS)] =
            (T)v[comp][sidx];
        if (bmh) {
          // split form: K-step kap'/32, lane (kap'%32/8)*16 + rr, element
          // kap'%8; 64 lanes x 8 hi halves, then 64 lanes x 8 lo halves
          _Float16 *blk = bmh + (long)rb * (split_blk_bytes(nkh) / 2);
          const float vf = (float)(v[comp][sidx] * sb);
          if (c == kz) {
            if (comp < 2 && sidx == 0)
              reinterpret_cast<float *>(blk + 4 * nkh * 2 * 64 * 8)[comp * 16 + rr] = vf;
          } else {
            const int kq = 2 * (c < kz ? c : c - 1) + sidx;
            const int unit = (comp * nkh + (kq >> 5)) * 2;
            const int ln = ((kq & 31) >> 3) * 16 + rr, el = kq & 7;
            const _Float16 h = (_Float16)vf;
            blk[(unit * 64 + ln) * 8 + el] = h;
            blk[((unit + 1) * 64 + ln) * 8 + el] = (_Float16)(vf - (float)h);
          }
        }
      }
  }
  __shared__ double red[kBlock];
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) epart[blockIdx.x] = red[0];
}

// Fixed-order sum of the S-pass chunk partials ([chunk][plane] -> [plane]):
// 64 entries per block, wave w sums chunks w, w+4, ..., then the four wave sums
// are combined in one fixed order (coalesced, and 4x the parallelism of one
// thread per entry walking every chunk).
template <typename P>
__global__ void k_chunk_sum(const P *part, int nchunks, long plane, double *out)
{
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long e = (long)blockIdx.x * 64 + lane;
  double s = 0;
  if (e < plane)
    for (int k = wave; k < nchunks; k += 4) s += (double)part[(long)k * plane + e];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && e < plane) out[e] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// The same sums for FP32 partials, four adjacent entries per lane (16-byte
// loads; plane is a multiple of 4); per entry the summation order is
// k_chunk_sum's.
__global__ void k_chunk_sum4(const float *part, int nchunks, long plane, double *out)
{
  __shared__ double red[4][4][64];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long e = ((long)blockIdx.x * 64 + lane) * 4;
  double s[4] = {0, 0, 0, 0};
  if (e < plane)
    for (int k = wave; k < nchunks; k += 4) {
      const float4 v = *reinterpret_cast<const float4 *>(part + (long)k * plane + e);
      s[0] += (double)v.x;
      s[1] += (double)v.y;
      s[2] += (double)v.z;
      s[3] += (double)v.w;
    }
#pragma unroll
  for (int i = 0; i < 4; ++i) red[wave][i][lane] = s[i];
  __syncthreads();
  if (wave == 0 && e < plane)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      out[e + i] = (red[0][i][lane] + red[1][i][lane]) + (red[2][i][lane] + red[3][i][lane]);
}

// zero the padding rows of the operand (kappa >= 2 nZ)
template <typename T>
__global__ void k_zero_t(long n, T *p)
{
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = (T)0;
}

// --------------------------------------------------------------- force pass --

constexpr int kWavesF = 8;  // waves per force-pass workgroup (share each operand block)

struct FArgs {
  int n, TS, nX, nY, nZ, ky, R, Rpad, KZ, KSP;
  int s0, nks;     // this launch's K slab: first K-step and K-steps (<= KS)
  int out4_add;    // add into out4 (K slabs after the first) instead of writing it
  const void *tab;
  const void *bm;  // this slab's operand blocks
  const double *q, *minv;
  const int *type;
  double c0, kxs, kys, kzs;  // C = 4 pi ke / V; 2 pi / l per axis
  double *ax, *ay, *az;
  double *out4;  // optional per-atom (fx, fy, fz, phi) without charge/mass
  // batched: per workgroup (first atom, end atom, -, replica), the replicas'
  // constants, and the operand stride (bytes) between replicas
  const int4 *items;
  const RepDev *rep;
  long bm_stride;
};

template <typename T, int KS, int NT>
__global__ __launch_bounds__(64 * kWavesF) void k_kforce(FArgs a)
{
  typedef typename Mfma<T>::V V;
  const Cx<T> *tab = reinterpret_cast<const Cx<T> *>(a.tab);
  const unsigned char *bm = reinterpret_cast<const unsigned char *>(a.bm);
  extern __shared__ unsigned char smem_raw[];
  const int blk_elems = 4 * 64 * a.KSP;
  // operand buffers at smem_raw + {0, blk_elems}: addressed from the __shared__
  // array itself so the compiler emits LDS (not flat) accesses
  const int nwaves = blockDim.x >> 6;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  long base = ((long)blockIdx.x * nwaves + wave) * (16 * NT);
  if (a.items) {
    const int4 it = a.items[blockIdx.x];
    const RepDev &r = a.rep[it.w];
    base = it.x + (long)wave * (16 * NT);
    a.n = it.y;
    bm += it.w * a.bm_stride;
    a.c0 = r.c0;
    a.kxs = r.kxs;
    a.kys = r.kys;
    a.kzs = r.kzs;
  }
  const int offY = a.nX, offZ = a.nX + a.nY;

  // A fragments: A[atom = l%16][kappa = 4 (s0 + s) + l/16] = (Zr, Zi)(c = kappa/2)
  T am[NT][KS];
#pragma unroll
  for (int tau = 0; tau < NT; ++tau) {
    const long atom = base + tau * 16 + lr;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int kap = 4 * (a.s0 + s) + lk, c = kap >> 1;
      T v = 0;
      if (atom < a.n && c < a.nZ && s < a.nks) {
        const Cx<T> Z = tab[atom * a.TS + offZ + c];
        v = (kap & 1) ? Z.im : Z.re;
      }
      am[tau][s] = v;
    }
  }
  T fx[NT][4], fy[NT][4], fz[NT][4], ph[NT][4];
#pragma unroll
  for (int tau = 0; tau < NT; ++tau)
#pragma unroll
    for (int e = 0; e < 4; ++e) fx[tau][e] = fy[tau][e] = fz[tau][e] = ph[tau][e] = 0;

  constexpr int VW = 16 / (int)sizeof(T);  // values per 16-byte vector
  typedef T Tvw __attribute__((ext_vector_type(VW)));
  const int nrb = a.Rpad >> 4;
  const int nks = a.nks;  // K-steps in use (<= KS)
  const int blk_bytes = blk_elems * (int)sizeof(T);
  // the operand of row block rb streams into LDS while rb-1 is multiplied
  copy_to_lds_async(bm, smem_raw, blk_bytes, wave, nwaves, lane);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int rb = 0; rb < nrb; ++rb) {
    const int cur = rb & 1;
    if (rb + 1 < nrb)
      copy_to_lds_async(bm + (size_t)(rb + 1) * blk_bytes,
                        smem_raw + (cur ^ 1) * blk_bytes, blk_bytes, wave, nwaves,
                        lane);
    V acc[4][NT];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int tau = 0; tau < NT; ++tau) acc[t][tau] = V{0, 0, 0, 0};
    // operand fragments of this row block: component t, this lane, KSP values
    const T *bfr = reinterpret_cast<const T *>(smem_raw) + cur * blk_elems + lane * a.KSP;
#pragma unroll
    for (int q = 0; q < (KS + VW - 1) / VW; ++q) {
      if (VW * q >= nks) break;
      Tvw bq[4];
#pragma unroll
      for (int t = 0; t < 4; ++t)
        bq[t] = *reinterpret_cast<const Tvw *>(bfr + t * 64 * a.KSP + VW * q);
#pragma unroll
      for (int u = 0; u < VW; ++u) {
        const int s = VW * q + u;
        if (s >= nks || s >= KS) break;
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
          for (int tau = 0; tau < NT; ++tau)
            acc[t][tau] = Mfma<T>::op(am[tau][s], bq[t][u], acc[t][tau]);
      }
    }
    // epilogue: lane holds rows r = rb*16 + l%16 for atoms mfma_row(l/16, e) of each tile
    const int r = rb * 16 + lr;
    if (r < a.R) {
      const int nxi = r / a.nY, nyi = r - (r / a.nY) * a.nY;
      const T wx = (T)nxi, wy = (T)(nyi - a.ky);
#pragma unroll
      for (int tau = 0; tau < NT; ++tau)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const long atom = base + tau * 16 + mfma_row<T>(lk, e);
          if (atom >= a.n) continue;
          const Cx<T> X = tab[atom * a.TS + nxi], Y = tab[atom * a.TS + offY + nyi];
          const T pr = X.re * Y.re - X.im * Y.im;
          const T pi = X.re * Y.im + X.im * Y.re;
          const T ur = acc[0][tau][e], ui = acc[1][tau][e];
          const T uzr = acc[2][tau][e], uzi = acc[3][tau][e];
          const T h = pr * ui + pi * ur;
          fx[tau][e] += wx * h;
          fy[tau][e] += wy * h;
          fz[tau][e] += pr * uzi + pi * uzr;
          ph[tau][e] += pr * ur - pi * ui;
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
  }
  // reduce over the 16 lanes that share l/16 (they hold different rows)
#pragma unroll
  for (int tau = 0; tau < NT; ++tau)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double vx = fx[tau][e], vy = fy[tau][e], vz = fz[tau][e], vp = ph[tau][e];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        vx += __shfl_xor(vx, o, 64);
        vy += __shfl_xor(vy, o, 64);
        vz += __shfl_xor(vz, o, 64);
        vp += __shfl_xor(vp, o, 64);
      }
      const long atom = base + tau * 16 + mfma_row<T>(lk, e);
      if (lr == 0 && atom < a.n) {
        const double gx = a.c0 * a.kxs * vx, gy = a.c0 * a.kys * vy, gz = a.c0 * a.kzs * vz;
        if (a.out4 && a.out4_add) {
          a.out4[4 * atom + 0] += gx;
          a.out4[4 * atom + 1] += gy;
          a.out4[4 * atom + 2] += gz;
          a.out4[4 * atom + 3] += a.c0 * vp;
        } else if (a.out4) {
          a.out4[4 * atom + 0] = gx;
          a.out4[4 * atom + 1] = gy;
          a.out4[4 * atom + 2] = gz;
          a.out4[4 * atom + 3] = a.c0 * vp;
        } else {
          const double qm = a.q[atom] * a.minv[a.type[atom]];
          a.ax[atom] += gx * qm;
          a.ay[atom] += gy * qm;
          a.az[atom] += gz * qm;
        }
      }
    }
}


// ------------------------------------------------- split-FP16 mixed form --
//
// The mixed form runs both GEMMs on v_mfma_f32_16x16x32_f16 (16 atoms or 16
// kappa values per K-step instead of 2 / 4 on the FP32 tiles).  Each FP32
// operand v is split into hi = fp16(v) and lo = fp16(v - hi) and a product is
// taken as hi*hi + hi*lo + lo*hi, which keeps ~22 significant bits (the
// dropped lo*lo term is below 2^-22 relative); accumulation is FP32 within
// an atom chunk and FP64 across chunks, as on the FP32 tiles.  Operands are
// pre-scaled by powers of two (sA on q X Y, sB on the force-pass operand,
// sB chosen from the bound |g| <= max W * sum |q|) so that no half overflows
// and the scales are removed exactly afterwards.

// Tables of the split form, blocked: atoms in blocks of kTabAtoms = 32.  Per
// block, each component of the FP32 record [X(0..kx) | Y(-ky..ky)] is one row
// of kTabRow = 272 bytes (32 complex values, then 16 bytes of padding so that
// the rows a wave reads side by side start in different LDS banks), and each
// nz one row [32 hi words | 32 lo words | 16 bytes] of the split pair
// (Zr, -Zi).  The second real-embedding column, (Zi, Zr), is the same halves
// swapped with one sign flipped, so the S pass derives it in LDS.  A block is
// copied to LDS as it lies and read there with compile-time offsets; atoms
// past the end of a block range are zero.
constexpr int kTabAtoms = 32, kTabRow = 272;

// Bytes between the (Zr, -Zi) rows and the (Zi, Zr) rows of an S-pass LDS
// buffer: a multiple of 16 that puts the second set 32 banks (of 64 x 4 B)
// from the first.
__host__ __device__ constexpr unsigned split_zpad(int nZ)
{
  return 4u * (unsigned)((32 - (nZ * (kTabRow / 4)) % 64 + 64) % 64);
}

__device__ __forceinline__ void put_split(unsigned char *xy, unsigned char *zt, long blk, int e,
                                          int j, int TXY, int nZ, float cs, float sn)
{
  if (e < TXY) {
    *reinterpret_cast<Cx<float> *>(xy + (blk * TXY + e) * kTabRow + j * 8) = Cx<float>{cs, sn};
  } else {
    unsigned h0, l0;
    split_pair(cs, -sn, h0, l0);
    unsigned char *row = zt + (blk * nZ + (e - TXY)) * kTabRow + j * 4;
    *reinterpret_cast<unsigned *>(row) = h0;
    *reinterpret_cast<unsigned *>(row + 128) = l0;
  }
}

// The rows of one axis for one atom: e^{2 pi i m u} for m = 0..k, by complex
// multiplication in FP64 from e^{2 pi i u} (u reduced to [-1/2, 1/2] first,
// one FP64 sincospi), each entry rounded to FP32 once; the entries of -m are
// the conjugates.  The FP64 products carry ~k ulp of FP64 phase error, far
// below the FP32 rounding of the entries.  Axis 0 (X, m >= 0 only) carries
// the atom's scaled charge.  e0 is the entry of m = 0; dead (padding) atoms
// write zeros.
__device__ __forceinline__ void split_axis_rows(unsigned char *xy, unsigned char *zt, long blk,
                                                int j, int TXY, int nZ, int axis, int k, int e0,
                                                double u, bool live, float qs)
{
  u -= rint(u);
  double s1, c1;
  sincospi(2.0 * u, &s1, &c1);
  double cr = 1.0, ci = 0.0;
  for (int m = 0; m <= k; ++m) {
    float cs = live ? (float)cr : 0.f, sn = live ? (float)ci : 0.f;
    if (axis == 0) {
      cs *= qs;
      sn *= qs;
    }
    put_split(xy, zt, blk, e0 + m, j, TXY, nZ, cs, sn);
    if (axis != 0 && m > 0) put_split(xy, zt, blk, e0 - m, j, TXY, nZ, cs, -sn);
    const double nr = cr * c1 - ci * s1, ni = cr * s1 + ci * c1;
    cr = nr;
    ci = ni;
  }
}

// One thread per (atom, axis); the grid's y index is the axis, so its kind
// (X, Y or Z rows) and coordinate are uniform per workgroup.  Consecutive
// threads write consecutive atoms of one row.
__global__ void k_tables_h(long npts, int kx, int ky, int kz, double ilx, double ily, double ilz,
                           const double *x, const double *y, const double *z, const double *q,
                           float sa, unsigned char *xy, unsigned char *zt)
{
  const int nX = kx + 1, nY = 2 * ky + 1, nZ = 2 * kz + 1;
  const int axis = (int)blockIdx.y;
  const long nat = (npts + kTabAtoms - 1) / kTabAtoms * kTabAtoms;
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= nat) return;
  const bool live = a < npts;
  double u = 0.0;
  float qs = 0.f;
  if (live) {
    u = axis == 0 ? x[a] * ilx : (axis == 1 ? y[a] * ily : z[a] * ilz);
    if (axis == 0) qs = (float)q[a] * sa;
  }
  const int k = axis == 0 ? kx : (axis == 1 ? ky : kz);
  const int e0 = axis == 0 ? 0 : (axis == 1 ? nX + ky : nX + nY + kz);
  split_axis_rows(xy, zt, a / kTabAtoms, (int)(a & (kTabAtoms - 1)), nX + nY, nZ, axis, k, e0, u,
                  live, qs);
}

__device__ __forceinline__ f4 mfma_h(u4v a, u4v b, f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a),
                                                __builtin_bit_cast(h8, b), c, 0, 0, 0);
}

struct SArgsH {
  int n, TS, nX, nY, nZ, Rtiles, chunk, stage, tiles_per_group, Rpad, SC;
  const void *tab;   // blocked FP32 records [X | Y] (TS = nX + nY rows per block)
  const void *ztab;  // blocked split (Zr, -Zi) rows (nZ per block)
  float *part;       // [chunk][Rpad][SC]: FP32 chunk sums, summed over chunks in FP64
  const int4 *items; // batched: (first, end) table position, chunk slot, replica
  int ncols, groups; // logical grid, as SArgs
};

// Same work split as k_sk: a workgroup is an atom chunk x a group of row
// tiles, wave w owns RT row tiles and every column tile; a K-step is 16 atoms
// and an LDS stage is one 32-atom table block.  Lane l feeds rows / columns
// l%16 with atoms 4(l/16)..4(l/16)+3 of the step, (re, im) of each atom
// adjacent along the reduction axis.  In the blocked layout a lane's four
// atoms are adjacent in a row, so each operand is read with two (records) or
// one (Z half) 16-byte LDS read at a compile-time offset from a per-lane base.
// The matrix cores take the 2 (nZ - 1) columns of nz != 0 (SCT tiles); the
// nz = 0 column, Z = 1 for every atom, is the plain sum of the A entries:
// each K-step's four atoms of a lane are summed in FP32 and accumulated in
// FP64 beside the products.
template <int RT, int SCT>
__global__ __launch_bounds__(64 * kWavesMax) void k_sk_h(SArgsH a)
{
  extern __shared__ unsigned char smem_raw[];
  constexpr int S = kTabAtoms;
  const unsigned rec_bytes = (unsigned)a.TS * kTabRow;  // [X | Y] rows of a block
  const unsigned zrow_bytes = (unsigned)a.nZ * kTabRow;  // (Zr, -Zi) rows of a block
  // per buffer: records (X already carries each atom's scaled charge); nZ
  // (Zr, -Zi) rows then nZ (Zi, Zr) rows, padded to nZ + 8 ceil(2 nZ / 16)
  // rows as split_sk_lds allocates them
  const unsigned zrows = (unsigned)(a.nZ + 8 * ((2 * a.nZ + 15) / 16));
  // the (Zi, Zr) rows start zpad bytes later, so that they sit 32 banks away
  // from the (Zr, -Zi) rows: a column tile's even and odd lanes then read
  // disjoint banks
  const unsigned zpad = split_zpad(a.nZ);
  const unsigned buf_bytes = rec_bytes + zrows * kTabRow + zpad;
  const unsigned char *tab = reinterpret_cast<const unsigned char *>(a.tab);
  const unsigned char *ztab = reinterpret_cast<const unsigned char *>(a.ztab);

  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nwaves = blockDim.x >> 6;
  // XCD-aware order: workgroups L and L+8 run on the same XCD (dispatch is
  // round-robin over the 8 XCDs), so the row groups of one chunk share its
  // records through one L2
  int chunk, grp;
  {
    const int L = blockIdx.x, q = L >> 3;
    grp = q % a.groups;
    chunk = (q / a.groups) * 8 + (L & 7);
    if (chunk >= a.ncols) return;
  }
  const int col = chunk;
  const int rt0 = grp * a.tiles_per_group + wave * RT;
  const int rt_end = min(a.Rtiles, (grp + 1) * a.tiles_per_group);
  long a0 = (long)chunk * a.chunk;
  long a1 = a0 + a.chunk;
  if (a1 > a.n) a1 = a.n;
  if (a.items) {
    const int4 it = a.items[col];
    a0 = it.x;
    a1 = it.y;
    chunk = it.z;
  }

  f4 acc[RT][SCT];
#pragma unroll
  for (int r = 0; r < RT; ++r)
#pragma unroll
    for (int s = 0; s < SCT; ++s) acc[r][s] = f4{0, 0, 0, 0};

  const int lr = lane & 15, lk = lane >> 4;
  // per-lane LDS offsets inside a buffer; rows past the lattice read row 0
  // (finite) and are masked on output
  unsigned xo[RT], yo[RT];
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    const int row = (rt0 + r) * 16 + lr;
    const bool ok = (rt0 + r) < rt_end && row < a.nX * a.nY;
    xo[r] = (unsigned)(ok ? row / a.nY : 0) * kTabRow + 32 * lk;
    yo[r] = (unsigned)(a.nX + (ok ? row % a.nY : 0)) * kTabRow + 32 * lk;
  }
  // product column s*16 + l%16: even columns are (Zr, -Zi) rows, odd ones
  // (Zi, Zr); column pair c' = s*8 + l%16/2 is nz != 0 number c' (table row
  // c' below kz, c' + 1 above); pairs past nZ - 1 read row 0 and are masked
  // on output
  const int kz = (a.nZ - 1) >> 1, ncp = a.nZ - 1;
  unsigned zo[SCT];
#pragma unroll
  for (int s = 0; s < SCT; ++s) {
    const int cp = s * 8 + (lr >> 1);
    const int zc = cp >= ncp ? 0 : (cp < kz ? cp : cp + 1);
    zo[s] = rec_bytes + (lr & 1) * zpad + (unsigned)((lr & 1) * a.nZ + zc) * kTabRow + 16 * lk;
  }
  double z0r[RT], z0i[RT];  // the lane's share of the nz = 0 column per row tile
#pragma unroll
  for (int r = 0; r < RT; ++r) z0r[r] = z0i[r] = 0.0;
  const bool active = rt0 < rt_end;

  // The last wave of the workgroup is the producer: it streams each block's
  // record and (Zr, -Zi) rows into the other LDS buffer while the compute
  // waves multiply the current one, then derives the (Zi, Zr) rows.  A chunk
  // starts on a block and ends on one or at the end of the atoms, whose block
  // is padded with zero entries, so no atom outside the chunk is staged.  The
  // compute waves issue no asynchronous LDS writes themselves, so their LDS
  // reads never wait on them.
  const bool producer = wave == nwaves - 1;
  auto issue = [&](long base, int b) {
    unsigned char *dst = smem_raw + b * buf_bytes;
    const long blk = base / S;
    copy_to_lds_async(tab + blk * rec_bytes, dst, rec_bytes, 0, 1, lane);
    copy_to_lds_async(ztab + blk * zrow_bytes, dst + rec_bytes, zrow_bytes, 0, 1, lane);
  };
  // (Zi, Zr) halves from (Zr, -Zi): swap the two fp16 halves of each word and
  // negate the new low one (rounding to fp16 is odd-symmetric, so this is
  // exactly the split of (Zi, Zr)).  Four 16-byte pieces per lane are read
  // before any is written, so the reads overlap.
  auto finish = [&](int b) {
    unsigned char *zd = smem_raw + b * buf_bytes + rec_bytes;
    const int total = a.nZ * 16;
    for (int e0 = lane; e0 < total; e0 += 256) {
      u4v w[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int e = e0 + 64 * k;
        if (e < total) w[k] = *reinterpret_cast<const u4v *>(zd + (e >> 4) * kTabRow + (e & 15) * 16);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int e = e0 + 64 * k;
        if (e < total) {
          u4v o;
#pragma unroll
          for (int m = 0; m < 4; ++m) o[m] = ((w[k][m] >> 16) | (w[k][m] << 16)) ^ 0x8000u;
          *reinterpret_cast<u4v *>(zd + zpad + (a.nZ + (e >> 4)) * kTabRow + (e & 15) * 16) = o;
        }
      }
    }
  };

  int b = 0;
  if (producer && a0 < a1) {
    issue(a0, 0);
    __builtin_amdgcn_s_waitcnt(0);
    finish(0);
  }
  __syncthreads();
  // three loops with the same barriers: producer, compute waves, and compute
  // waves with no row tile in this group (no accumulators live across them)
  if (producer) {
    for (long base = a0; base < a1; base += S) {
      if (base + S < a1) {
        issue(base + S, b ^ 1);
        __builtin_amdgcn_s_waitcnt(0);
        finish(b ^ 1);
      }
      __syncthreads();
      b ^= 1;
    }
    return;
  }
  if (!active) {
    for (long base = a0; base < a1; base += S) __syncthreads();
    return;
  }
  for (long base = a0; base < a1; base += S) {
    const unsigned cur = (unsigned)b * buf_bytes;
#pragma unroll
    for (int ks = 0; ks < S / 16; ++ks) {
      u4v ahi[RT], alo[RT];
#pragma unroll
      for (int r = 0; r < RT; ++r) {
        const float4 *px = reinterpret_cast<const float4 *>(smem_raw + cur + xo[r] + 128 * ks);
        const float4 *py = reinterpret_cast<const float4 *>(smem_raw + cur + yo[r] + 128 * ks);
        const float4 x01 = px[0], x23 = px[1], y01 = py[0], y23 = py[1];
        const float xr[4] = {x01.x, x01.z, x23.x, x23.z}, xi[4] = {x01.y, x01.w, x23.y, x23.w};
        const float yr[4] = {y01.x, y01.z, y23.x, y23.z}, yi[4] = {y01.y, y01.w, y23.y, y23.w};
        float tr = 0.f, ti = 0.f;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float pr = xr[u] * yr[u] - xi[u] * yi[u];
          const float pi = xr[u] * yi[u] + xi[u] * yr[u];
          unsigned h, l;
          split_pair_mix(pr, pi, h, l);
          ahi[r][u] = h;
          alo[r][u] = l;
          tr += pr;
          ti += pi;
        }
        z0r[r] += (double)tr;
        z0i[r] += (double)ti;
      }
#pragma unroll
      for (int s = 0; s < SCT; ++s) {
        const unsigned char *pz = smem_raw + cur + zo[s] + 64 * ks;
        const u4v bhi = *reinterpret_cast<const u4v *>(pz);
        const u4v blo = *reinterpret_cast<const u4v *>(pz + 128);
#pragma unroll
        for (int r = 0; r < RT; ++r) {
          acc[r][s] = mfma_h(ahi[r], bhi, acc[r][s]);
          acc[r][s] = mfma_h(ahi[r], blo, acc[r][s]);
          acc[r][s] = mfma_h(alo[r], bhi, acc[r][s]);
        }
      }
    }
    __syncthreads();
    b ^= 1;
  }
  // nz = 0 column: the four lanes of a row hold atoms 4 (l/16) .. +3 of
  // every K-step; summed in a fixed order
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    z0r[r] += __shfl_xor(z0r[r], 16, 64);
    z0i[r] += __shfl_xor(z0i[r], 16, 64);
    z0r[r] += __shfl_xor(z0r[r], 32, 64);
    z0i[r] += __shfl_xor(z0i[r], 32, 64);
  }
  const int nrow = a.nX * a.nY;
#pragma unroll
  for (int r = 0; r < RT; ++r) {
    if (rt0 + r >= rt_end) break;
    float *out = a.part + ((long)chunk * a.Rpad + (rt0 + r) * 16) * a.SC;
#pragma unroll
    for (int s = 0; s < SCT; ++s)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = 4 * lk + e;
        const int cp = s * 8 + (lr >> 1);
        const bool ok = (rt0 + r) * 16 + row < nrow;
        if (cp < ncp) out[(long)row * a.SC + 2 * (cp < kz ? cp : cp + 1) + (lr & 1)] = ok ? acc[r][s][e] : 0.f;
      }
    // lanes 0..15 write the nz = 0 pair of row l%16, lanes 16..63 clear the
    // padding columns 2 nZ .. SC - 1
    const bool ok = (rt0 + r) * 16 + lr < nrow;
    if (lk == 0) {
      out[(long)lr * a.SC + 2 * kz] = ok ? (float)z0r[r] : 0.f;
      out[(long)lr * a.SC + 2 * kz + 1] = ok ? (float)z0i[r] : 0.f;
    } else {
      for (int col = 2 * a.nZ + lk - 1; col < a.SC; col += 3) out[(long)lr * a.SC + col] = 0.f;
    }
  }
}

struct FArgsH {
  int n, TS, nX, nY, nZ, ky, kz, R, Rpad, nkh;  // TS = nX + nY: FP32 [X | Y] record length
  const void *tab, *ztab, *bmh;
  const double *minv;
  const int *type;
  double c0, kxs, kys, kzs;  // c0 includes 1/(sA sB): X carries q sA, the operand sB
  double *ax, *ay, *az;
  const int4 *items;  // batched, as FArgs
  const RepDev *rep;
  long bm_stride;
  double inv_sb;      // batched: c0 of a replica is multiplied by it (1/(sA sB))
};

// Force pass of the split form: A = the atoms' Z (kappa = 2 nz + re/im, from
// the split records), B = the operand blocks staged in LDS per row block and
// shared by the workgroup's waves.  The workgroup's blocked X rows and Y rows
// of ny >= 0 are staged in LDS once (Y(-ny) = conj Y(ny)), so the epilogue's
// per-(atom, row) phase factors are LDS reads.  Staging a third fewer rows
// leaves LDS on the CU for list-build waves when both run concurrently.
template <int NK, int NT, int W>
__global__ __launch_bounds__(64 * W) void k_kforce_h(FArgsH a)
{
  const unsigned char *tab = reinterpret_cast<const unsigned char *>(a.tab);
  const unsigned char *ztab = reinterpret_cast<const unsigned char *>(a.ztab);
  const unsigned char *bm = reinterpret_cast<const unsigned char *>(a.bmh);
  extern __shared__ unsigned char smem_raw[];
  const int nwaves = blockDim.x >> 6;
  const int blk_bytes = split_blk_bytes(NK);
  const int wg_atoms = 16 * NT * nwaves;  // a multiple of kTabAtoms
  // staged rows per block: X(0..kx) and Y(0..ky); Y(-ny) = conj Y(ny)
  const int TSs = a.nX + a.ky + 1;
  const int xy_bytes = wg_atoms / kTabAtoms * TSs * kTabRow;
  // operand buffers at smem_raw + xy_bytes + {0, blk_bytes}, addressed from the
  // __shared__ array itself so the compiler emits LDS (not flat) accesses
  const Cx<float> *lxy = reinterpret_cast<const Cx<float> *>(smem_raw);
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lr = lane & 15, lk = lane >> 4;
  long wg_base = (long)blockIdx.x * wg_atoms;
  long toff = 0;  // table position - atom index
  if (a.items) {
    const int4 it = a.items[blockIdx.x];
    const RepDev &r = a.rep[it.w];
    wg_base = it.x;
    toff = (long)r.tbase - r.obase;
    a.n = it.y;
    bm += it.w * a.bm_stride;
    a.c0 = r.c0 * a.inv_sb;
    a.kxs = r.kxs;
    a.kys = r.kys;
    a.kzs = r.kzs;
  }
  const long base = wg_base + (long)wave * (16 * NT);
  const int offY = a.nX;
  {
    // the workgroup's table blocks (its first atom starts a block)
    const long cnt = min((long)wg_atoms, a.n - wg_base);
    const int nb = (int)((cnt + kTabAtoms - 1) / kTabAtoms);
    // (the X rows and the Y rows of ny >= 0 of each block)
    const unsigned char *src = tab + (wg_base + toff) / kTabAtoms * a.TS * kTabRow;
    for (int bk = 0; bk < nb; ++bk) {
      copy_to_lds_async(src + (size_t)bk * a.TS * kTabRow, smem_raw + bk * TSs * kTabRow,
                        a.nX * kTabRow, wave, nwaves, lane);
      copy_to_lds_async(src + ((size_t)bk * a.TS + a.nX + a.ky) * kTabRow,
                        smem_raw + (bk * TSs + a.nX) * kTabRow, (a.ky + 1) * kTabRow, wave,
                        nwaves, lane);
    }
  }

  // A fragments: atom l%16 of tile tau, kappa = 32 s + 8 (l/16) + i, i.e.
  // nz = 16 s + 4 (l/16) + i/2, plain (Zr, Zi) = the (Zr, -Zi) record with the
  // sign of its upper half flipped
  u4v ahi[NT][NK], alo[NT][NK];
#pragma unroll
  for (int tau = 0; tau < NT; ++tau) {
    const long atom = base + tau * 16 + lr;
#pragma unroll
    for (int s = 0; s < NK; ++s)
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int c = 16 * s + 4 * lk + u;  // c' (the z = 0 column is left out)
        unsigned h = 0, l = 0;
        if (atom < a.n && c < a.nZ - 1) {
          const long t = atom + toff;
          const int zc = c < a.kz ? c : c + 1;
          const unsigned char *row =
              ztab + ((t / kTabAtoms) * a.nZ + zc) * kTabRow + (t % kTabAtoms) * 4;
          h = *reinterpret_cast<const unsigned *>(row) ^ 0x80000000u;
          l = *reinterpret_cast<const unsigned *>(row + 128) ^ 0x80000000u;
        }
        ahi[tau][s][u] = h;
        alo[tau][s][u] = l;
      }
  }
  // the split pass computes accelerations only (point fields use the FP32 /
  // FP64 tiles), so the potential column is not accumulated
  float fx[NT][4], fy[NT][4], fz[NT][4];
#pragma unroll
  for (int tau = 0; tau < NT; ++tau)
#pragma unroll
    for (int e = 0; e < 4; ++e) fx[tau][e] = fy[tau][e] = fz[tau][e] = 0;

  // LDS offset (in complex values) of the lane's atoms 4 (l/16) .. +3 of each
  // tile in the blocked [X | Y] rows: adjacent, so one 32-byte read per row
  int xoff[NT];
#pragma unroll
  for (int tau = 0; tau < NT; ++tau) {
    const int rel = wave * 16 * NT + tau * 16 + 4 * lk;
    xoff[tau] = (rel / kTabAtoms) * TSs * (kTabRow / 8) + rel % kTabAtoms;
  }
  // lattice row of the lane, r = rb*16 + l%16 = nxi nY + nyi, kept incrementally;
  // rows past R have zero operands (U = 0) and read the phase factors of row 0
  // (finite), so they add exactly nothing
  int nxi = lr / a.nY, nyi = lr % a.nY;

  const int nrb = a.Rpad >> 4;
  copy_to_lds_async(bm, smem_raw + xy_bytes, blk_bytes, wave, nwaves, lane);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int rb = 0; rb < nrb; ++rb) {
    const int cur = rb & 1;
    // The next row block's operand (CH = 512 NK pieces of 16 B) is fetched
    // into registers here and stored to the other LDS buffer after this block
    // is consumed: with no LDS write in flight, reads of the current block
    // need no wait on the prefetch.
    constexpr int CH = split_blk_bytes(NK) / 16, PER = (CH + 64 * W - 1) / (64 * W);
    const bool more = rb + 1 < nrb;
    u4v pf[PER];
    if (more) {
      const u4v *src = reinterpret_cast<const u4v *>(bm + (size_t)(rb + 1) * blk_bytes);
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int idx = threadIdx.x + k * (64 * W);
        if (idx < CH) pf[k] = src[idx];
      }
    }
    // phase factors X Y of the lane's (atom, row) pairs, formed while the
    // products below run (atoms past n give sums that are never written)
    const bool rin = rb * 16 + lr < a.R;
    const int rx = rin ? nxi : 0, yk = rin ? nyi - a.ky : 0, ya = yk < 0 ? -yk : yk;
    const float wx = (float)rx, wy = (float)yk, ys = yk < 0 ? -1.f : 1.f;
    float pr[NT][4], pi[NT][4];
#pragma unroll
    for (int tau = 0; tau < NT; ++tau) {
      const float4 *px = reinterpret_cast<const float4 *>(lxy + xoff[tau] + rx * (kTabRow / 8));
      const float4 *py = reinterpret_cast<const float4 *>(lxy + xoff[tau] + (offY + ya) * (kTabRow / 8));
      const float4 x01 = px[0], x23 = px[1], y01 = py[0], y23 = py[1];
      const float xr[4] = {x01.x, x01.z, x23.x, x23.z}, xi[4] = {x01.y, x01.w, x23.y, x23.w};
      const float yr[4] = {y01.x, y01.z, y23.x, y23.z};
      const float yi[4] = {ys * y01.y, ys * y01.w, ys * y23.y, ys * y23.w};
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        pr[tau][e] = xr[e] * yr[e] - xi[e] * yi[e];
        pi[tau][e] = xr[e] * yi[e] + xi[e] * yr[e];
      }
    }
    // accumulators seeded with the z = 0 column's products (components 0, 1)
    const float *z0 = reinterpret_cast<const float *>(smem_raw + xy_bytes + cur * blk_bytes +
                                                      4 * NK * 2 * 1024);
    const float b0 = z0[lr], b1 = z0[16 + lr];
    f4 acc[4][NT];
#pragma unroll
    for (int tau = 0; tau < NT; ++tau) {
      acc[0][tau] = f4{b0, b0, b0, b0};
      acc[1][tau] = f4{b1, b1, b1, b1};
      acc[2][tau] = acc[3][tau] = f4{0, 0, 0, 0};
    }
    // Per K-step the four components' fragments are loaded first; each split
    // product then sweeps all 4 x NT accumulators, so consecutive MFMAs are
    // independent (per accumulator the order stays hi.hi, hi.lo, lo.hi).
#pragma unroll
    for (int s = 0; s < NK; ++s) {
      u4v bh[4], bl[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const unsigned char *u = smem_raw + xy_bytes + cur * blk_bytes + ((t * NK + s) * 2) * 1024 + lane * 16;
        bh[t] = *reinterpret_cast<const u4v *>(u);
        bl[t] = *reinterpret_cast<const u4v *>(u + 1024);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int tau = 0; tau < NT; ++tau) acc[t][tau] = mfma_h(ahi[tau][s], bh[t], acc[t][tau]);
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int tau = 0; tau < NT; ++tau) acc[t][tau] = mfma_h(ahi[tau][s], bl[t], acc[t][tau]);
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int tau = 0; tau < NT; ++tau) acc[t][tau] = mfma_h(alo[tau][s], bh[t], acc[t][tau]);
    }
#pragma unroll
    for (int tau = 0; tau < NT; ++tau)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float ur = acc[0][tau][e], ui = acc[1][tau][e];
        const float uzr = acc[2][tau][e], uzi = acc[3][tau][e];
        const float h = pr[tau][e] * ui + pi[tau][e] * ur;
        fx[tau][e] += wx * h;
        fy[tau][e] += wy * h;
        fz[tau][e] += pr[tau][e] * uzi + pi[tau][e] * uzr;
      }
    nyi += 16;
    while (nyi >= a.nY) {
      nyi -= a.nY;
      ++nxi;
    }
    if (more) {
      u4v *dst = reinterpret_cast<u4v *>(smem_raw + xy_bytes + (cur ^ 1) * blk_bytes);
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int idx = threadIdx.x + k * (64 * W);
        if (idx < CH) dst[idx] = pf[k];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int tau = 0; tau < NT; ++tau)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double vx = fx[tau][e], vy = fy[tau][e], vz = fz[tau][e];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        vx += __shfl_xor(vx, o, 64);
        vy += __shfl_xor(vy, o, 64);
        vz += __shfl_xor(vz, o, 64);
      }
      const long atom = base + tau * 16 + 4 * lk + e;
      if (lr == 0 && atom < a.n) {
        const double gx = a.c0 * a.kxs * vx, gy = a.c0 * a.kys * vy, gz = a.c0 * a.kzs * vz;
        // the phase factors carry the atom's charge (times sA, removed in c0)
        const double m = a.minv[a.type[atom]];
        a.ax[atom] += gx * m;
        a.ay[atom] += gy * m;
        a.az[atom] += gz * m;
      }
    }
}

// ------------------------------------------------------------ dipole term --

// calculate_dipole_sum (initialize.cpp:94-115): per-block FP64 partials of q r.
__global__ void k_dipole_part(int n, const double *q, const double *x, const double *y,
                              const double *z, double *part)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  __shared__ double red[3][kBlock];
  double v[3] = {0, 0, 0};
  if (i < n) {
    v[0] = q[i] * x[i];
    v[1] = q[i] * y[i];
    v[2] = q[i] * z[i];
  }
  for (int d = 0; d < 3; ++d) red[d][threadIdx.x] = v[d];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int d = 0; d < 3; ++d) red[d][threadIdx.x] += red[d][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int d = 0; d < 3; ++d) part[3 * blockIdx.x + d] = red[d][0];
}

// The reference adds q_i * dipole_field inside its k loop, i.e. once per used
// k-vector (electrostatic_ewald_k.cpp:197,216-219); the same total is added here.
__global__ void k_dipole_apply(int n, int nblk, const double *part, double dip_coef, double nk,
                               const double *q, const double *minv, const int *type, double *ax,
                               double *ay, double *az, double *dsum_out)
{
  __shared__ double f[3];
  if (threadIdx.x < 3) {
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += part[3 * b + threadIdx.x];
    f[threadIdx.x] = -s * dip_coef;
    if (blockIdx.x == 0) dsum_out[threadIdx.x] = s;
  }
  __syncthreads();
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double qm = q[i] * minv[type[i]] * nk;
  ax[i] += f[0] * qm;
  ay[i] += f[1] * qm;
  az[i] += f[2] * qm;
}

// ------------------------------------------------------------------- host --

static int roundup(int v, int m) { return (v + m - 1) / m * m; }

// Elements of one K slab of the FP32/FP64 force-pass operand (all row blocks).
static size_t slab_elems(const EwaldKPlan &p) { return (size_t)(p.Rpad / 16) * 4 * 64 * p.KSP; }

// Atoms per LDS stage of the FP32/FP64 S pass: the largest of 32, 16, ..., 2
// whose two buffers of records (TS complex values each) and charges fit in
// 150 KB, or 0 if none does.
static int s_stage_for(int TS, size_t tsz)
{
  for (int st = 32; st >= 2; st /= 2)
    if (2 * (size_t)st * TS * 2 * tsz + 2 * (size_t)st * tsz <= 150 * 1024) return st;
  return 0;
}

// The k lattice of one box: sizes, W = m(k) f(k) over the half lattice and
// the map from the reference's k order.  Returns an error message or nullptr.
const char *plan_k_lattice(EwaldKPlan &p, int kx, int ky, int kz, int precision, double alpha,
                           double lx, double ly, double lz)
{
  p.kx = kx;
  p.ky = ky;
  p.kz = kz;
  if (p.kx < 0 || p.ky < 0 || p.kz < 0) return "negative k_max";
  p.nX = p.kx + 1;
  p.nY = 2 * p.ky + 1;
  p.nZ = 2 * p.kz + 1;
  p.R = p.nX * p.nY;
  p.Rpad = roundup(p.R, 16);
  p.KZ = roundup(2 * p.nZ, 4);
  // the force pass holds at most 32 K-steps (128 kappa) of A fragments in
  // registers and of operand in LDS: longer nz axes are split into slabs of
  // equal depth, one launch each, adding into the same outputs
  p.nslab = (p.KZ / 4 + 31) / 32;
  p.KSS = (p.KZ / 4 + p.nslab - 1) / p.nslab;
  // fragment stride: whole 16-byte vectors, an odd number of them per lane so
  // that the lanes' LDS reads of the force-pass operand hit distinct banks
  {
    const int vw = precision == CAV_FP64 ? 2 : 4;
    int ksp = roundup(p.KSS, vw);
    if ((ksp / vw) % 2 == 0) ksp += vw;
    p.KSP = ksp;
  }
  // the S pass stages at least two atom records of nX + nY + nZ complex values
  // per LDS buffer (double-buffered)
  if (s_stage_for(p.nX + p.nY + p.nZ, precision == CAV_FP64 ? 8 : 4) == 0)
    return "k lattice too large: one S-pass stage of two atom records exceeds LDS";
  p.nk_ref = (2 * p.kx) * (2 * p.ky) * (2 * p.kz);
  p.lx = lx;
  p.ly = ly;
  p.lz = lz;
  const double lxi = 1.0 / p.lx, lyi = 1.0 / p.ly, lzi = 1.0 / p.lz;
  p.lxyz_inv = 1.0 / (p.lx * p.ly * p.lz);
  // make_k_vectors (initialize.cpp:184-222): f(k) as the reference computes it;
  // the index of (kx,ky,kz) in the lexicographic list without 0 decides membership.
  const double a2 = alpha * alpha, f4a2 = 1.0 / (4.0 * a2);
  const long zero_idx = ((long)p.kx * p.nY + p.ky) * p.nZ + p.kz;
  auto ref_index = [&](int nx, int ny, int nz) -> long {
    const long g = ((long)(nx + p.kx) * p.nY + (ny + p.ky)) * p.nZ + (nz + p.kz);
    return g > zero_idx ? g - 1 : g;
  };
  auto in_K = [&](int nx, int ny, int nz) -> bool {
    if (nx == 0 && ny == 0 && nz == 0) return false;
    return ref_index(nx, ny, nz) < p.nk_ref;
  };
  auto fk = [&](int nx, int ny, int nz) -> double {
    const double k0 = CAV_2PI * nx * lxi, k1 = CAV_2PI * ny * lyi, k2 = CAV_2PI * nz * lzi;
    const double ksq = k0 * k0 + k1 * k1 + k2 * k2;
    return (1.0 / ksq) * std::exp(-ksq * f4a2);
  };
  p.W.assign((size_t)p.R * p.nZ, 0.0);
  for (int nx = 0; nx <= p.kx; ++nx)
    for (int ny = -p.ky; ny <= p.ky; ++ny)
      for (int nz = -p.kz; nz <= p.kz; ++nz) {
        const bool half = nx > 0 || (nx == 0 && ny > 0) || (nx == 0 && ny == 0 && nz > 0);
        if (!half) continue;
        const int m = (int)in_K(nx, ny, nz) + (int)in_K(-nx, -ny, -nz);
        if (!m) continue;
        const int r = nx * p.nY + (ny + p.ky);
        p.W[(size_t)r * p.nZ + (nz + p.kz)] = m * fk(nx, ny, nz);
      }
  // reference order of K for cav_get_sk: each k maps to a computed (row, col),
  // directly (nx >= 0) or through S(k) = conj S(-k)
  p.ref_r.clear();
  p.ref_c.clear();
  p.ref_conj.clear();
  for (int nx = -p.kx; nx <= p.kx; ++nx)
    for (int ny = -p.ky; ny <= p.ky; ++ny)
      for (int nz = -p.kz; nz <= p.kz; ++nz) {
        if (!in_K(nx, ny, nz)) continue;
        const bool direct = nx >= 0;
        const int ax_ = direct ? nx : -nx, ay_ = direct ? ny : -ny, az_ = direct ? nz : -nz;
        p.ref_r.push_back(ax_ * p.nY + (ay_ + p.ky));
        p.ref_c.push_back(az_ + p.kz);
        p.ref_conj.push_back(direct ? 0 : 1);
      }
  return nullptr;
}

// Power-of-two operand scales of the split form from the bounds
// |q X Y| <= max|q| and |g| <= max W * sum_j |q_j|.
void split_scales(double qmax, double qsum, double wmax, double *sA, double *sB)
{
  *sA = qmax > 0 ? std::ldexp(1.0, -(int)std::ceil(std::log2(qmax))) : 1.0;
  const double gb = wmax * qsum;
  *sB = gb > 0 ? std::ldexp(1.0, 14 - (int)std::ceil(std::log2(gb))) : 1.0;
}

int ewald_k_plan(cav_ctx *c)
{
  EwaldKPlan &p = c->ek;
  // initialize.cpp:42-77: lengths from size_global, slab override of one axis
  double l[3] = {c->L[0], c->L[1], c->L[2]};
  if (c->slab_axis >= 0 && c->slab_axis < 3) l[c->slab_axis] = c->slab_lz;
  if (const char *e = plan_k_lattice(p, c->kx, c->ky, c->kz, c->precision, c->ek_alpha, l[0], l[1], l[2]))
    return fail(c, CAV_ERR_ARG, e);
  if (c->dipole) p.dip_coef = c->ek_ke * CAV_4PI * p.lxyz_inv / (1 + 2.0 * c->eps_dipole);
  // split form: K-steps of 32 along the 2 (nZ - 1) kappa other than z = 0
  c->nkh = std::max(1, (2 * (p.nZ - 1) + 31) / 32);
  double qmax = 0, qsum = 0, wmax = 0;
  for (int i = 0; i < c->n; ++i) {
    const double qa = std::fabs(c->charge[c->h_type[i]]);
    qmax = std::max(qmax, qa);
    qsum += qa;
  }
  for (double w : p.W) wmax = std::max(wmax, std::fabs(w));
  split_scales(qmax, qsum, wmax, &c->sA, &c->sB);
  return CAV_OK;
}

constexpr int kRT = 3;  // row tiles per S-pass wave

// Row-tile groups of the S pass: as few groups as the 8-wave limit allows,
// split evenly so that no wave of a group idles.
static void s_layout(const cav_ctx *c, int *groups, int *tiles_per_group, int *waves)
{
  const int tiles = c->ek.Rpad / 16;
  const int per_max = kWavesMax * kRT;
  const int g = (tiles + per_max - 1) / per_max;
  int tpg = (tiles + g - 1) / g;
  tpg = (tpg + kRT - 1) / kRT * kRT;
  *groups = g;
  *tiles_per_group = tpg;
  *waves = tpg / kRT;
}

// Split S pass: at most kWavesMax - 1 compute waves (the producer wave is
// added at launch).
static void s_layout_h(int Rpad, int RT, int *groups, int *tiles_per_group, int *waves)
{
  const int tiles = Rpad / 16;
  const int per_max = (kWavesMax - 1) * RT;
  const int g = (tiles + per_max - 1) / per_max;
  int tpg = (tiles + g - 1) / g;
  tpg = (tpg + RT - 1) / RT * RT;
  *groups = g;
  *tiles_per_group = tpg;
  *waves = tpg / RT;
}

static void s_layout_rt(int Rpad, int RT, int *groups, int *tiles_per_group, int *waves)
{
  const int tiles = Rpad / 16;
  const int per_max = kWavesMax * RT;
  const int g = (tiles + per_max - 1) / per_max;
  int tpg = (tiles + g - 1) / g;
  tpg = (tpg + RT - 1) / RT * RT;
  *groups = g;
  *tiles_per_group = tpg;
  *waves = tpg / RT;
}

static int s_stage(const cav_ctx *c, size_t tsz)
{
  return s_stage_for(c->ek.nX + c->ek.nY + c->ek.nZ, tsz);
}

static int s_chunk(const cav_ctx *c, int ntarget)
{
  // aim for >= ntarget workgroups over the whole grid; chunk multiple of 32
  int groups, tpg, waves;
  s_layout(c, &groups, &tpg, &waves);
  long per = ((long)c->n * groups + ntarget - 1) / ntarget;
  per = (per + 31) / 32 * 32;
  if (per < 32) per = 32;
  return (int)per;
}

int ewald_k_alloc(cav_ctx *c)
{
  const EwaldKPlan &p = c->ek;
  const size_t tsz = c->precision == CAV_FP64 ? sizeof(double) : sizeof(float);
  const int TS = p.nX + p.nY + p.nZ;
  // padded: the S pass streams whole 16-byte chunks past the last record; the
  // split form's blocked rows (whole 32-atom blocks) fit in the same buffer
  const size_t nblk = ((size_t)c->n + kTabAtoms - 1) / kTabAtoms;
  const size_t tab_bytes =
      std::max((size_t)c->n * TS * 2 * tsz, nblk * (p.nX + p.nY) * kTabRow) + 4096;
  if (c->tab) (void)hipFree(c->tab);
  CAV_HIP(hipMalloc(&c->tab, tab_bytes));
  c->tab_bytes = tab_bytes;
  c->tab_cap = c->n;
  const int chunk = c->tune_chunk > 0 ? c->tune_chunk : s_chunk(c, 512);
  c->sk_chunks = (c->n + chunk - 1) / chunk;
  if (c->sk_chunks < 1) c->sk_chunks = 1;
  const int SC = roundup(2 * p.nZ, 16);
  if (c->sk_part) (void)hipFree(c->sk_part);
  CAV_HIP(hipMalloc(&c->sk_part, sizeof(double) * (size_t)c->sk_chunks * p.Rpad * SC));
  CAV_HIP(hipMemsetAsync(c->sk_part, 0, sizeof(double) * (size_t)c->sk_chunks * p.Rpad * SC,
                         c->stream));
  if (c->sk) (void)hipFree(c->sk);
  CAV_HIP(hipMalloc(&c->sk, sizeof(double) * (size_t)p.Rpad * 2 * p.nZ));
  if (c->sk_sum) (void)hipFree(c->sk_sum);
  CAV_HIP(hipMalloc(&c->sk_sum, sizeof(double) * (size_t)p.Rpad * SC));
  if (c->bmat) (void)hipFree(c->bmat);
  const size_t bm_elems = p.nslab * slab_elems(p);
  CAV_HIP(hipMalloc(&c->bmat, bm_elems * tsz));
  if (c->precision == CAV_FP64)
    k_zero_t<double><<<grid_for(bm_elems), kBlock, 0, c->stream>>>(bm_elems, (double *)c->bmat);
  else
    k_zero_t<float><<<grid_for(bm_elems), kBlock, 0, c->stream>>>(bm_elems, (float *)c->bmat);
  if (c->ztab) (void)hipFree(c->ztab);
  if (c->bmat_h) (void)hipFree(c->bmat_h);
  c->ztab = c->bmat_h = nullptr;
  if (c->precision == CAV_MIXED) {
    CAV_HIP(hipMalloc(&c->ztab, nblk * p.nZ * kTabRow + 4096));
    const size_t bh_elems = (size_t)(p.Rpad / 16) * (split_blk_bytes(c->nkh) / 2);
    CAV_HIP(hipMalloc(&c->bmat_h, bh_elems * sizeof(_Float16)));
    k_zero_t<_Float16><<<grid_for(bh_elems), kBlock, 0, c->stream>>>(bh_elems, (_Float16 *)c->bmat_h);
  }
  if (c->d_W) (void)hipFree(c->d_W);
  CAV_HIP(hipMalloc(&c->d_W, sizeof(double) * (size_t)p.R * p.nZ));
  CAV_HIP(hipMemcpyAsync(c->d_W, p.W.data(), sizeof(double) * p.W.size(), hipMemcpyHostToDevice,
                         c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  return CAV_OK;
}

template <typename K>
static void allow_smem(K kern, size_t smem)
{
  if (smem > 48 * 1024)
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)smem);
}

template <typename T, int RT>
static void launch_sk_rt(hipStream_t st, SArgs a, dim3 lgrid, size_t smem, int SCT, int waves)
{
  a.ncols = (int)lgrid.x;
  a.groups = (int)lgrid.y;
  const dim3 grid((lgrid.x + 7) / 8 * 8 * lgrid.y);
  const int block = 64 * waves;
  allow_smem(k_sk<T, RT, 1>, smem);
  allow_smem(k_sk<T, RT, 2>, smem);
  allow_smem(k_sk<T, RT, 3>, smem);
  allow_smem(k_sk<T, RT, 4>, smem);
  allow_smem(k_sk<T, RT, 5>, smem);
  allow_smem(k_sk<T, RT, 6>, smem);
  allow_smem(k_sk<T, RT, 7>, smem);
  allow_smem(k_sk<T, RT, 8>, smem);
  // one launch per 8 column tiles; each writes its own columns of the partials
  for (a.ct0 = 0; a.ct0 < SCT; a.ct0 += 8) {
    switch (std::min(8, SCT - a.ct0)) {
    case 1: k_sk<T, RT, 1><<<grid, block, smem, st>>>(a); break;
    case 2: k_sk<T, RT, 2><<<grid, block, smem, st>>>(a); break;
    case 3: k_sk<T, RT, 3><<<grid, block, smem, st>>>(a); break;
    case 4: k_sk<T, RT, 4><<<grid, block, smem, st>>>(a); break;
    case 5: k_sk<T, RT, 5><<<grid, block, smem, st>>>(a); break;
    case 6: k_sk<T, RT, 6><<<grid, block, smem, st>>>(a); break;
    case 7: k_sk<T, RT, 7><<<grid, block, smem, st>>>(a); break;
    default: k_sk<T, RT, 8><<<grid, block, smem, st>>>(a); break;
    }
  }
}

// Atom tiles per force-pass wave: two in FP32 (operand reuse), one in FP64
// (keeps the kernel at 2+ waves per SIMD).
template <typename T> constexpr int kforce_nt() { return sizeof(T) == 8 ? 1 : 2; }

// The force pass, one launch per K slab (one for nz axes of up to 128 kappa);
// `a.bm` is the first slab's operand, the others follow it.
template <typename T>
static void launch_kforce_t(hipStream_t st, FArgs a, const EwaldKPlan &p, int g)
{
  constexpr int NT = kforce_nt<T>();
  const size_t smem = 2 * (size_t)4 * 64 * a.KSP * sizeof(T);
  const int KS = p.KSS, nks = p.KZ / 4;
  const unsigned char *bm0 = reinterpret_cast<const unsigned char *>(a.bm);
  for (int sl = 0; sl < p.nslab; ++sl) {
    a.s0 = sl * p.KSS;
    a.nks = std::min(p.KSS, nks - a.s0);
    a.out4_add = sl > 0;
    a.bm = bm0 + sl * slab_elems(p) * sizeof(T);
    if (KS <= 8) {
      allow_smem(k_kforce<T, 8, NT>, smem);
      k_kforce<T, 8, NT><<<g, 64 * kWavesF, smem, st>>>(a);
    } else if (KS <= 12) {
      allow_smem(k_kforce<T, 12, NT>, smem);
      k_kforce<T, 12, NT><<<g, 64 * kWavesF, smem, st>>>(a);
    } else if (KS <= 17) {
      allow_smem(k_kforce<T, 17, NT>, smem);
      k_kforce<T, 17, NT><<<g, 64 * kWavesF, smem, st>>>(a);
    } else if (KS <= 24) {
      allow_smem(k_kforce<T, 24, NT>, smem);
      k_kforce<T, 24, NT><<<g, 64 * kWavesF, smem, st>>>(a);
    } else {
      allow_smem(k_kforce<T, 32, NT>, smem);
      k_kforce<T, 32, NT><<<g, 64 * kWavesF, smem, st>>>(a);
    }
  }
}

// Phase tables + S pass + reduction for npts points (the owned atoms, or query
// points when tab/positions are the caller's).
template <typename T>
static int sk_pipeline(cav_ctx *c)
{
  const EwaldKPlan &p = c->ek;
  const int TS = p.nX + p.nY + p.nZ;
  timer_begin(c, c->t_tab);
  k_tables<T><<<grid_for((long)c->n * TS), kBlock, 0, c->stream>>>(
      c->n, TS, p.kx, p.ky, p.kz, 1.0 / p.lx, 1.0 / p.ly, 1.0 / p.lz, c->x, c->y, c->z,
      (Cx<T> *)c->tab);
  timer_end(c, c->t_tab);
  SArgs a;
  a.n = c->n;
  a.TS = TS;
  a.nX = p.nX;
  a.nY = p.nY;
  a.nZ = p.nZ;
  a.ky = p.ky;
  a.kz = p.kz;
  a.Rtiles = p.Rpad / 16;
  a.chunk = (c->n + c->sk_chunks - 1) / c->sk_chunks;
  a.chunk = (a.chunk + 31) / 32 * 32;
  a.stage = s_stage(c, sizeof(T));
  a.q = c->q;
  a.tab = c->tab;
  a.part = c->sk_part;
  a.items = nullptr;
  a.Rpad = p.Rpad;
  a.SC = roundup(2 * p.nZ, 16);
  int groups, waves;
  s_layout(c, &groups, &a.tiles_per_group, &waves);
  const int nch = (c->n + a.chunk - 1) / a.chunk;
  dim3 grid(nch, groups);
  const size_t smem = 2 * (size_t)a.stage * TS * sizeof(Cx<T>) + 2 * (size_t)a.stage * sizeof(T);
  timer_begin(c, c->t_sk);
  launch_sk_rt<T, kRT>(c->stream, a, grid, smem, a.SC / 16, waves);
  timer_end(c, c->t_sk);
  const int nthr = p.Rpad * p.nZ;
  const long plane = (long)p.Rpad * a.SC;
  k_chunk_sum<double><<<(plane + 63) / 64, kBlock, 0, c->stream>>>(c->sk_part, nch, plane, c->sk_sum);
  k_sk_reduce<T, double><<<grid_for(nthr), kBlock, 0, c->stream>>>(
      1, p.Rpad, a.SC, p.R, p.nZ, p.kz, p.KSS, p.KSP, c->sk_sum, c->d_W, c->sk, (T *)c->bmat, c->part,
      1.0, nullptr, 0, 1.0, nullptr, 0, 0);
  CAV_HIP(hipGetLastError());
  return reduce_partials(c, grid_for(nthr), 3);
}

template <typename T>
static int kforce(cav_ctx *c, int npts, const void *tab, double *out4)
{
  const EwaldKPlan &p = c->ek;
  FArgs f;
  f.n = npts;
  f.TS = p.nX + p.nY + p.nZ;
  f.nX = p.nX;
  f.nY = p.nY;
  f.nZ = p.nZ;
  f.ky = p.ky;
  f.R = p.R;
  f.Rpad = p.Rpad;
  f.KZ = p.KZ;
  f.KSP = p.KSP;
  f.tab = tab;
  f.bm = c->bmat;
  f.q = c->q;
  f.minv = c->d_minv;
  f.type = c->type;
  f.c0 = CAV_4PI * c->ek_ke * p.lxyz_inv;
  f.kxs = CAV_2PI / p.lx;
  f.kys = CAV_2PI / p.ly;
  f.kzs = CAV_2PI / p.lz;
  f.ax = c->ax;
  f.ay = c->ay;
  f.az = c->az;
  f.out4 = out4;
  f.items = nullptr;
  f.rep = nullptr;
  f.bm_stride = 0;
  timer_begin(c, c->t_force);
  constexpr int NT = kforce_nt<T>();
  launch_kforce_t<T>(c->stream, f, p, (npts + 16 * NT * kWavesF - 1) / (16 * NT * kWavesF));
  timer_end(c, c->t_force);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// LDS of a split S-pass workgroup: two buffers of one table block each
// ([X | Y] rows, nZ + 8 SCT Z rows and the bank padding between the two Z
// row sets).  (A third buffer, to request blocks two ahead, measured slower:
// it halves the workgroups per CU.)
// Product column tiles of the split S pass: the 2 (nZ - 1) columns of nz != 0
// (the nz = 0 column is summed on the vector ALUs).
static int split_sct(const EwaldKPlan &p) { return std::max(1, (2 * (p.nZ - 1) + 15) / 16); }

static size_t split_sk_lds(const EwaldKPlan &p)
{
  const int zrows = p.nZ + 8 * (roundup(2 * p.nZ, 16) / 16);
  return 2 * ((size_t)(p.nX + p.nY + zrows) * kTabRow + split_zpad(p.nZ));
}

template <int RT>
static void launch_sk_h(hipStream_t st, SArgsH a, dim3 lgrid, size_t smem, int SCT, int waves)
{
  a.ncols = (int)lgrid.x;
  a.groups = (int)lgrid.y;
  const dim3 grid((lgrid.x + 7) / 8 * 8 * lgrid.y);
  const int block = 64 * (waves + 1);  // compute waves + the producer wave
#define CAV_SKH(S)                                                      \
  case S:                                                               \
    allow_smem(k_sk_h<RT, S>, smem);                                    \
    k_sk_h<RT, S><<<grid, block, smem, st>>>(a);                 \
    break;
  switch (SCT) {
    CAV_SKH(1) CAV_SKH(2) CAV_SKH(3) CAV_SKH(4) CAV_SKH(5) CAV_SKH(6) CAV_SKH(7)
    default:
      allow_smem(k_sk_h<RT, 8>, smem);
      k_sk_h<RT, 8><<<grid, block, smem, st>>>(a);
  }
#undef CAV_SKH
}

// Mixed form: FP32 tables + split Z records, split-FP16 S pass, FP64 chunk
// reduction writing S, the FP32 operand (point queries, adjoint) and the split one.
static int sk_pipeline_h(cav_ctx *c)
{
  const EwaldKPlan &p = c->ek;
  const long nblk = (c->n + kTabAtoms - 1) / kTabAtoms;
  timer_begin(c, c->t_tab);
  k_tables_h<<<dim3(grid_for(nblk * kTabAtoms), 3), kBlock, 0, c->stream>>>(
      c->n, p.kx, p.ky, p.kz, 1.0 / p.lx, 1.0 / p.ly, 1.0 / p.lz, c->x, c->y, c->z, c->q,
      (float)c->sA, (unsigned char *)c->tab, (unsigned char *)c->ztab);
  timer_end(c, c->t_tab);
  SArgsH a;
  a.n = c->n;
  a.TS = p.nX + p.nY;
  a.nX = p.nX;
  a.nY = p.nY;
  a.nZ = p.nZ;
  a.Rtiles = p.Rpad / 16;
  a.chunk = (c->n + c->sk_chunks - 1) / c->sk_chunks;
  a.chunk = (a.chunk + kTabAtoms - 1) / kTabAtoms * kTabAtoms;
  a.stage = kTabAtoms;
  a.tab = c->tab;
  a.ztab = c->ztab;
  a.part = reinterpret_cast<float *>(c->sk_part);
  a.items = nullptr;
  a.Rpad = p.Rpad;
  a.SC = roundup(2 * p.nZ, 16);
  const int RT = c->tune_rt;
  int groups, waves;
  s_layout_h(p.Rpad, RT, &groups, &a.tiles_per_group, &waves);
  const int nch = (c->n + a.chunk - 1) / a.chunk;
  const size_t smem = split_sk_lds(p);
  const dim3 lg(nch, groups);
  timer_begin(c, c->t_sk);
  if (RT == 1)
    launch_sk_h<1>(c->stream, a, lg, smem, split_sct(p), waves);
  else if (RT == 2)
    launch_sk_h<2>(c->stream, a, lg, smem, split_sct(p), waves);
  else if (RT == 4)
    launch_sk_h<4>(c->stream, a, lg, smem, split_sct(p), waves);
  else
    launch_sk_h<3>(c->stream, a, lg, smem, split_sct(p), waves);
  timer_end(c, c->t_sk);
  const int nthr = p.Rpad * p.nZ;
  const long plane = (long)p.Rpad * a.SC;
  k_chunk_sum4<<<(plane / 4 + 63) / 64, kBlock, 0, c->stream>>>(
      reinterpret_cast<const float *>(c->sk_part), nch, plane, c->sk_sum);
  k_sk_reduce<float, double><<<grid_for(nthr), kBlock, 0, c->stream>>>(
      1, p.Rpad, a.SC, p.R, p.nZ, p.kz, p.KSS, p.KSP, c->sk_sum, c->d_W, c->sk, (float *)c->bmat,
      c->part, 1.0 / c->sA, (_Float16 *)c->bmat_h, c->nkh, c->sB, nullptr, 0, 0);
  CAV_HIP(hipGetLastError());
  return reduce_partials(c, grid_for(nthr), 3);
}

// LDS of a split force-pass workgroup: two operand blocks plus the blocked X
// rows and Y rows of ny >= 0 of its 16 nt w atoms (w waves).
static size_t split_force_lds_p(const EwaldKPlan &p, int nkh, int nt = 2, int w = kWavesF)
{
  return 2 * (size_t)split_blk_bytes(nkh) +
         (size_t)16 * nt * w / kTabAtoms * (p.nX + p.ky + 1) * kTabRow;
}

static size_t split_force_lds(const cav_ctx *c)
{
  return split_force_lds_p(c->ek, c->nkh, c->tune_fnt, c->tune_fw);
}

// Whether the split form's fixed tiling covers this k lattice: at most 8 S-pass
// column tiles and 4 force-pass K-steps (|nz| <= 31), and both passes' LDS
// within a CU's 160 KB.  Wider lattices take the FP32 tiles.
static bool split_fits(const EwaldKPlan &p, int nkh, int nt, int w)
{
  return roundup(2 * p.nZ, 16) / 16 <= 8 && nkh <= 4 &&
         split_force_lds_p(p, nkh, nt, w) <= 160 * 1024 && split_sk_lds(p) <= 160 * 1024;
}

template <int NT, int W>
static void launch_kforce_h(hipStream_t st, const FArgsH &f, int nkh, int g, size_t smem)
{
  switch (nkh) {
  case 1: allow_smem(k_kforce_h<1, NT, W>, smem); k_kforce_h<1, NT, W><<<g, 64 * W, smem, st>>>(f); break;
  case 2: allow_smem(k_kforce_h<2, NT, W>, smem); k_kforce_h<2, NT, W><<<g, 64 * W, smem, st>>>(f); break;
  case 3: allow_smem(k_kforce_h<3, NT, W>, smem); k_kforce_h<3, NT, W><<<g, 64 * W, smem, st>>>(f); break;
  default: allow_smem(k_kforce_h<4, NT, W>, smem); k_kforce_h<4, NT, W><<<g, 64 * W, smem, st>>>(f); break;
  }
}

static int kforce_h(cav_ctx *c)
{
  const EwaldKPlan &p = c->ek;
  FArgsH f;
  f.n = c->n;
  f.TS = p.nX + p.nY;
  f.nX = p.nX;
  f.nY = p.nY;
  f.nZ = p.nZ;
  f.ky = p.ky;
  f.kz = p.kz;
  f.R = p.R;
  f.Rpad = p.Rpad;
  f.nkh = c->nkh;
  f.tab = c->tab;
  f.ztab = c->ztab;
  f.bmh = c->bmat_h;
  f.minv = c->d_minv;
  f.type = c->type;
  f.c0 = CAV_4PI * c->ek_ke * p.lxyz_inv / c->sB / c->sA;  // both powers of two
  f.kxs = CAV_2PI / p.lx;
  f.kys = CAV_2PI / p.ly;
  f.kzs = CAV_2PI / p.lz;
  f.ax = c->ax;
  f.ay = c->ay;
  f.az = c->az;
  f.items = nullptr;
  f.rep = nullptr;
  f.bm_stride = 0;
  f.inv_sb = 1.0;
  const int NT = c->tune_fnt, W = c->tune_fw;
  const int g = (c->n + 16 * NT * W - 1) / (16 * NT * W);
  const size_t smem = split_force_lds(c);
  timer_begin(c, c->t_force);
  if (W == 16) {
    if (NT == 1)
      launch_kforce_h<1, 16>(c->stream, f, c->nkh, g, smem);
    else
      launch_kforce_h<2, 16>(c->stream, f, c->nkh, g, smem);
  } else if (NT == 1) {
    launch_kforce_h<1, kWavesF>(c->stream, f, c->nkh, g, smem);
  } else {
    launch_kforce_h<2, kWavesF>(c->stream, f, c->nkh, g, smem);
  }
  timer_end(c, c->t_force);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

int launch_ewald_k(cav_ctx *c)
{
  if (c->n == 0) return CAV_OK;
  if (c->precision == CAV_FP64) {
    CAV_TRY(sk_pipeline<double>(c));
    CAV_TRY(kforce<double>(c, c->n, c->tab, nullptr));
  } else if (split_fits(c->ek, c->nkh, c->tune_fnt, c->tune_fw)) {
    CAV_TRY(sk_pipeline_h(c));
    CAV_TRY(kforce_h(c));
  } else {  // k lattices too wide for the split form's LDS tiles: FP32 tiles
    CAV_TRY(sk_pipeline<float>(c));
    CAV_TRY(kforce<float>(c, c->n, c->tab, nullptr));
  }
  if (c->dipole && !c->defer_dipole) CAV_TRY(launch_dipole(c));
  return CAV_OK;
}

// The epsilon_dipole term, added into the accelerations after the k-space sum
// (the reference adds it inside the same per-atom loop, after the k terms).
int launch_dipole(cav_ctx *c)
{
  const int nb = grid_for(c->n);
  k_dipole_part<<<nb, kBlock, 0, c->stream>>>(c->n, c->q, c->x, c->y, c->z, c->part);
  k_dipole_apply<<<nb, kBlock, 0, c->stream>>>(c->n, nb, c->part, c->ek.dip_coef,
                                               (double)c->ek.nk_ref, c->q, c->d_minv, c->type,
                                               c->ax, c->ay, c->az, c->scal + 8);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// S over the half lattice with arbitrary real per-atom weights u_j in place
// of the charges (the adjoint's species and residual-weighted structure
// factors); written as [Rpad][nZ] complex FP64.
__global__ void k_sk_sum(int nchunks, int Rpad, int SC, int nZ, const double *part, double *out)
{
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= Rpad * nZ) return;
  const int r = t / nZ, c = t - r * nZ;
  double sr = 0, si = 0;
  for (int k = 0; k < nchunks; ++k) {
    const double *p = part + ((long)k * Rpad + r) * SC + 2 * c;
    sr += p[0];
    si += p[1];
  }
  out[(long)r * 2 * nZ + 2 * c] = sr;
  out[(long)r * 2 * nZ + 2 * c + 1] = si;
}

template <typename T>
static int weighted_sk_t(cav_ctx *c, const double *weights, double *out)
{
  const EwaldKPlan &p = c->ek;
  const int TS = p.nX + p.nY + p.nZ;
  k_tables<T><<<grid_for((long)c->n * TS), kBlock, 0, c->stream>>>(
      c->n, TS, p.kx, p.ky, p.kz, 1.0 / p.lx, 1.0 / p.ly, 1.0 / p.lz, c->x, c->y, c->z,
      (Cx<T> *)c->tab);
  SArgs a;
  a.n = c->n;
  a.TS = TS;
  a.nX = p.nX;
  a.nY = p.nY;
  a.nZ = p.nZ;
  a.ky = p.ky;
  a.kz = p.kz;
  a.Rtiles = p.Rpad / 16;
  a.chunk = (c->n + c->sk_chunks - 1) / c->sk_chunks;
  a.chunk = (a.chunk + 31) / 32 * 32;
  a.stage = s_stage(c, sizeof(T));
  a.q = weights;
  a.tab = c->tab;
  a.part = c->sk_part;
  a.items = nullptr;
  a.Rpad = p.Rpad;
  a.SC = roundup(2 * p.nZ, 16);
  int groups, waves;
  s_layout(c, &groups, &a.tiles_per_group, &waves);
  const int nch = (c->n + a.chunk - 1) / a.chunk;
  const size_t smem = 2 * (size_t)a.stage * TS * sizeof(Cx<T>) + 2 * (size_t)a.stage * sizeof(T);
  launch_sk_rt<T, kRT>(c->stream, a, dim3(nch, groups), smem, a.SC / 16, waves);
  const long plane = (long)p.Rpad * a.SC;
  k_chunk_sum<double><<<(plane + 63) / 64, kBlock, 0, c->stream>>>(c->sk_part, nch, plane, c->sk_sum);
  k_sk_sum<<<grid_for((long)p.Rpad * p.nZ), kBlock, 0, c->stream>>>(1, p.Rpad, a.SC, p.nZ,
                                                                    c->sk_sum, out);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

int ewald_k_weighted_sk(cav_ctx *c, const double *weights, double *out)
{
  if (c->precision == CAV_FP64) return weighted_sk_t<double>(c, weights, out);
  return weighted_sk_t<float>(c, weights, out);
}

// Per-atom k-space field (C sum k Im(g e^{ik.r}), and the potential) with the
// operand of the last force evaluation.
int ewald_k_field(cav_ctx *c, double *out4)
{
  if (c->precision == CAV_FP64) return kforce<double>(c, c->n, c->tab, out4);
  // the split form leaves [X | Y] records in tab: rebuild the full FP32 ones
  const EwaldKPlan &p = c->ek;
  const int TS = p.nX + p.nY + p.nZ;
  k_tables<float><<<grid_for((long)c->n * TS), kBlock, 0, c->stream>>>(
      c->n, TS, p.kx, p.ky, p.kz, 1.0 / p.lx, 1.0 / p.ly, 1.0 / p.lz, c->x, c->y, c->z,
      (Cx<float> *)c->tab);
  CAV_HIP(hipGetLastError());
  return kforce<float>(c, c->n, c->tab, out4);
}

// E_k = 1/2 C sum W |S|^2 + self (+ dipole), valid after a force evaluation at
// the same positions (the reference's S(k) coupling, energy.cpp:61-116).
int launch_energy_k(cav_ctx *c, double *e)
{
  double v[12];
  CAV_HIP(hipMemcpyAsync(v, c->scal, sizeof(double) * 12, hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  double q2 = 0;
  for (int i = 0; i < c->n; ++i) {
    const double qq = c->charge[c->h_type[i]];
    q2 += qq * qq;
  }
  const double self = -CAV_PIS_INV * c->ek_alpha * q2 * c->ek_ke;
  double ek = 0.5 * CAV_4PI * c->ek_ke * c->ek.lxyz_inv * v[3] + self;
  if (c->dipole) {
    const double d2 = v[8] * v[8] + v[9] * v[9] + v[10] * v[10];
    ek += 0.5 * c->ek.dip_coef * d2;
  }
  *e = ek;
  return CAV_OK;
}

// Point queries: tables for the points, then the force pass with q = 1 into a
// per-point (field, potential) buffer; the dipole field is added on the host.
int point_field_k(cav_ctx *c, int np, const double *pts, double *field, double *pot)
{
  const EwaldKPlan &p = c->ek;
  const int TS = p.nX + p.nY + p.nZ;
  const size_t tsz = c->precision == CAV_FP64 ? sizeof(double) : sizeof(float);
  double *d;
  void *tab;
  CAV_HIP(hipMalloc(&d, sizeof(double) * 7 * (size_t)(np > 0 ? np : 1)));
  CAV_HIP(hipMalloc(&tab, (size_t)(np > 0 ? np : 1) * TS * 2 * tsz));
  std::vector<double> soa(3 * (size_t)np);
  for (int i = 0; i < np; ++i)
    for (int k = 0; k < 3; ++k) soa[(size_t)k * np + i] = pts[3 * i + k];
  CAV_HIP(hipMemcpyAsync(d, soa.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice, c->stream));
  if (c->precision == CAV_FP64) {
    k_tables<double><<<grid_for((long)np * TS), kBlock, 0, c->stream>>>(
        np, TS, p.kx, p.ky, p.kz, 1.0 / p.lx, 1.0 / p.ly, 1.0 / p.lz, d, d + np, d + 2 * np,
        (Cx<double> *)tab);
    CAV_TRY(kforce<double>(c, np, tab, d + 3 * np));
  } else {
    k_tables<float><<<grid_for((long)np * TS), kBlock, 0, c->stream>>>(
        np, TS, p.kx, p.ky, p.kz, 1.0 / p.lx, 1.0 / p.ly, 1.0 / p.lz, d, d + np, d + 2 * np,
        (Cx<float> *)tab);
    CAV_TRY(kforce<float>(c, np, tab, d + 3 * np));
  }
  std::vector<double> o4(4 * (size_t)np);
  CAV_HIP(hipMemcpyAsync(o4.data(), d + 3 * np, sizeof(double) * 4 * np, hipMemcpyDeviceToHost,
                         c->stream));
  double dsum[3] = {0, 0, 0};
  if (c->dipole)
    CAV_HIP(hipMemcpyAsync(dsum, c->scal + 8, sizeof(double) * 3, hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < np; ++i) {
    for (int k = 0; k < 3; ++k) {
      double f = o4[4 * (size_t)i + k];
      if (c->dipole) f += -dsum[k] * p.dip_coef;
      if (field) field[3 * i + k] = f;
    }
    if (pot) pot[i] = o4[4 * (size_t)i + 3];
  }
  CAV_HIP(hipFree(d));
  CAV_HIP(hipFree(tab));
  return CAV_OK;
}


// ---------------------------------------------------------- batched replicas --
//
// All replicas share the k lattice (kx, ky, kz) but not the box: W, the phase
// table lengths and the force constants are per replica.  The S pass runs over
// a table of (replica, atom chunk) work items, the reduction over a grid row
// per replica, and the force pass over (replica, atom group) items, so one
// launch of each kernel covers every replica.

// Phase tables of every replica: grid row = replica.  Full FP64/FP32 records
// (plain form) or FP32 [X | Y] + split Z records (mixed form).
template <typename T>
__global__ void kb_tables(const RepDev *rep, int kx, int ky, int kz, const double *x,
                          const double *y, const double *z, Cx<T> *tab)
{
  const RepDev &r = rep[blockIdx.y];
  const int nX = kx + 1, nY = 2 * ky + 1, nZ = 2 * kz + 1, TS = nX + nY + nZ;
  const long t = (long)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (long)r.n * TS) return;
  const long a = t / TS;
  const int e = (int)(t - a * TS);
  const long s = r.base + a;
  double u;
  int m;
  if (e < nX) {
    m = e;
    u = x[s] * r.il[0];
  } else if (e < nX + nY) {
    m = e - nX - ky;
    u = y[s] * r.il[1];
  } else {
    m = e - nX - nY - kz;
    u = z[s] * r.il[2];
  }
  double f = (double)m * u;
  f -= rint(f);
  Cx<T> v;
  if constexpr (sizeof(T) == 8) {
    double sn, cs;
    sincospi(2.0 * f, &sn, &cs);
    v.re = cs;
    v.im = sn;
  } else {
    float sn, cs;
    sincospif((float)(2.0 * f), &sn, &cs);
    v.re = cs;
    v.im = sn;
  }
  tab[(r.obase + a) * TS + e] = v;
}

// Batched split tables: replica r's atoms fill the blocks from r.tbase.  The
// grid's y index is the replica and its z index the axis, as in k_tables_h.
__global__ void kb_tables_h(const RepDev *rep, int kx, int ky, int kz, const double *x,
                            const double *y, const double *z, const double *q, float sa,
                            unsigned char *xy, unsigned char *zt)
{
  const RepDev &r = rep[blockIdx.y];
  const int nX = kx + 1, nY = 2 * ky + 1, nZ = 2 * kz + 1;
  const int axis = (int)blockIdx.z;
  const long nat = (r.n + kTabAtoms - 1) / kTabAtoms * kTabAtoms;
  const long a = (long)blockIdx.x * kBlock + threadIdx.x;
  if (a >= nat) return;
  const long s = r.base + a;
  const bool live = a < r.n;
  double u = 0.0;
  float qs = 0.f;
  if (live) {
    u = axis == 0 ? x[s] * r.il[0] : (axis == 1 ? y[s] * r.il[1] : z[s] * r.il[2]);
    if (axis == 0) qs = (float)q[r.obase + a] * sa;
  }
  const int k = axis == 0 ? kx : (axis == 1 ? ky : kz);
  const int e0 = axis == 0 ? 0 : (axis == 1 ? nX + ky : nX + nY + kz);
  split_axis_rows(xy, zt, r.tbase / kTabAtoms + a / kTabAtoms, (int)(a & (kTabAtoms - 1)),
                  nX + nY, nZ, axis, k, e0, u, live, qs);
}


static bool batch_split(const cav_batch *b)
{
  return b->precision == CAV_MIXED && split_fits(b->ek, b->nkh, b->f_nt, kWavesF);
}

int batch_ewald_k_plan(cav_batch *b)
{
  if (!b->ek_on) return CAV_OK;
  b->W.clear();
  double qmax = 0, gb = 0;
  for (int r = 0; r < b->B; ++r) {
    RepDev &d = b->reps[r];
    EwaldKPlan p;
    if (const char *e = plan_k_lattice(p, b->kx, b->ky, b->kz, b->precision, b->ek_alpha, d.L[0],
                                       d.L[1], d.L[2]))
      return bfail(b, CAV_ERR_ARG, e);
    if (r == 0) b->ek = p;
    b->W.insert(b->W.end(), p.W.begin(), p.W.end());
    d.c0 = CAV_4PI * b->ek_ke * p.lxyz_inv;
    d.kxs = CAV_2PI / p.lx;
    d.kys = CAV_2PI / p.ly;
    d.kzs = CAV_2PI / p.lz;
    double qsum = 0, wmax = 0;
    for (int t : b->hreps[r].type) {
      const double qa = std::fabs(b->charge[t]);
      qmax = std::max(qmax, qa);
      qsum += qa;
    }
    for (double w : p.W) wmax = std::max(wmax, std::fabs(w));
    gb = std::max(gb, wmax * qsum);
  }
  b->nkh = std::max(1, (2 * (b->ek.nZ - 1) + 31) / 32);
  split_scales(qmax, gb > 0 ? 1.0 : 0.0, gb, &b->sA, &b->sB);
  return CAV_OK;
}

static int batch_free_k(cav_batch *b)
{
  void *ptrs[] = {b->d_W, b->tab, b->ztab, b->bmat, b->bmat_h, b->sk_part, b->sk, b->epart,
                  b->s_items, b->s_items_o, b->f_items, b->rchunks};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  b->d_W = b->sk_part = b->sk = b->epart = nullptr;
  b->tab = b->ztab = b->bmat = b->bmat_h = nullptr;
  b->s_items = b->s_items_o = b->f_items = nullptr;
  b->rchunks = nullptr;
  return CAV_OK;
}

// Work tables and buffers of the batched Ewald_k passes for the current tiles.
int batch_ewald_k_alloc(cav_batch *b)
{
  batch_free_k(b);
  if (!b->ek_on) return CAV_OK;
  const EwaldKPlan &p = b->ek;
  const bool split = batch_split(b);
  const size_t tsz = b->precision == CAV_FP64 ? sizeof(double) : sizeof(float);
  const int TS = p.nX + p.nY + p.nZ;
  const size_t nblk = (size_t)b->ntab / kTabAtoms;
  BAT_HIP(hipMalloc(&b->tab, std::max((size_t)b->nown * TS * 2 * tsz,
                                      nblk * (p.nX + p.nY) * kTabRow) + 4096));
  if (split) BAT_HIP(hipMalloc(&b->ztab, nblk * p.nZ * kTabRow + 4096));
  // S-pass items: chunks of s_chunk atoms per replica, largest replicas first
  std::vector<int4> sitems;
  std::vector<int2> rch(b->B);
  std::vector<int> order(b->B);
  for (int r = 0; r < b->B; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(),
                   [&](int u, int v) { return b->reps[u].n > b->reps[v].n; });
  int slot = 0;
  for (int r = 0; r < b->B; ++r) {
    const RepDev &d = b->reps[r];
    rch[r].x = slot;
    int cnt = 0;
    for (int a0 = 0; a0 < d.n; a0 += b->s_chunk) {
      const int a1 = std::min(d.n, a0 + b->s_chunk);
      sitems.push_back(make_int4(d.obase + a0, d.obase + a1, slot + cnt, r));
      ++cnt;
    }
    rch[r].y = cnt;
    slot += cnt;
  }
  b->n_chunks = slot;
  // issue the items of the largest replicas first
  std::vector<int4> sorted;
  for (int r : order)
    for (const int4 &it : sitems)
      if (it.w == r) sorted.push_back(it);
  b->n_sitems = (int)sorted.size();
  const int SC = roundup(2 * p.nZ, 16);
  BAT_HIP(hipMalloc(&b->s_items_o, sizeof(int4) * std::max(1, b->n_sitems)));
  BAT_HIP(hipMemcpy(b->s_items_o, sorted.data(), sizeof(int4) * sorted.size(), hipMemcpyHostToDevice));
  if (split)  // the split form's blocked tables hold replica r from tbase
    for (int4 &it : sorted) {
      const int sh = b->reps[it.w].tbase - b->reps[it.w].obase;
      it.x += sh;
      it.y += sh;
    }
  BAT_HIP(hipMalloc(&b->s_items, sizeof(int4) * std::max(1, b->n_sitems)));
  BAT_HIP(hipMemcpy(b->s_items, sorted.data(), sizeof(int4) * sorted.size(), hipMemcpyHostToDevice));
  BAT_HIP(hipMalloc(&b->rchunks, sizeof(int2) * b->B));
  BAT_HIP(hipMemcpy(b->rchunks, rch.data(), sizeof(int2) * b->B, hipMemcpyHostToDevice));
  BAT_HIP(hipMalloc(&b->sk_part, sizeof(double) * (size_t)std::max(1, b->n_chunks) * p.Rpad * SC));
  BAT_HIP(hipMalloc(&b->sk, sizeof(double) * (size_t)b->B * p.Rpad * 2 * p.nZ));
  BAT_HIP(hipMalloc(&b->epart, sizeof(double) * (size_t)b->B * grid_for((long)p.Rpad * p.nZ)));
  // force-pass operands per replica (padding entries stay zero)
  b->bm_stride = (long)(p.nslab * slab_elems(p));
  BAT_HIP(hipMalloc(&b->bmat, b->bm_stride * tsz * b->B));
  BAT_HIP(hipMemset(b->bmat, 0, b->bm_stride * tsz * b->B));
  if (split) {
    b->bmh_stride = (long)(p.Rpad / 16) * (split_blk_bytes(b->nkh) / 2);
    BAT_HIP(hipMalloc(&b->bmat_h, b->bmh_stride * sizeof(_Float16) * b->B));
    BAT_HIP(hipMemset(b->bmat_h, 0, b->bmh_stride * sizeof(_Float16) * b->B));
  }
  BAT_HIP(hipMalloc(&b->d_W, sizeof(double) * b->W.size()));
  BAT_HIP(hipMemcpy(b->d_W, b->W.data(), sizeof(double) * b->W.size(), hipMemcpyHostToDevice));
  // force-pass items: groups of one workgroup's atoms per replica
  // (item ends are clipped, so a workgroup spanning more atoms than wg, as the
  // field pass's FP32 tiles may, computes each atom once)
  const int wg = split ? 16 * b->f_nt * kWavesF
                       : 16 * (b->precision == CAV_FP64 ? kforce_nt<double>() : kforce_nt<float>()) *
                             kWavesF;
  std::vector<int4> fitems;
  for (int r : order) {
    const RepDev &d = b->reps[r];
    for (int a0 = 0; a0 < d.n; a0 += wg)
      fitems.push_back(make_int4(d.obase + a0, d.obase + std::min(d.n, a0 + wg), 0, r));
  }
  b->n_fitems = (int)fitems.size();
  BAT_HIP(hipMalloc(&b->f_items, sizeof(int4) * std::max(1, b->n_fitems)));
  BAT_HIP(hipMemcpy(b->f_items, fitems.data(), sizeof(int4) * fitems.size(), hipMemcpyHostToDevice));
  return CAV_OK;
}

template <int RT>
static void batch_sk_h(cav_batch *b, const SArgsH &a, int groups, size_t smem, int SCT, int waves)
{
  launch_sk_h<RT>(b->stream, a, dim3(b->n_sitems, groups), smem, SCT, waves);
}

template <typename T>
static int batch_ewald_k_t(cav_batch *b)
{
  const EwaldKPlan &p = b->ek;
  const int TS = p.nX + p.nY + p.nZ;
  const int SC = roundup(2 * p.nZ, 16);
  const dim3 tgrid(grid_for((long)b->maxn * TS), b->B);
  int groups, tpg, waves;
  s_layout_rt(p.Rpad, kRT, &groups, &tpg, &waves);
  const bool split = batch_split(b);
  if (!split) {
    kb_tables<T><<<tgrid, kBlock, 0, b->stream>>>(b->d_rep, p.kx, p.ky, p.kz, b->x, b->y, b->z,
                                                  (Cx<T> *)b->tab);
    SArgs a;
    a.n = 0;
    a.TS = TS;
    a.nX = p.nX;
    a.nY = p.nY;
    a.nZ = p.nZ;
    a.ky = p.ky;
    a.kz = p.kz;
    a.Rtiles = p.Rpad / 16;
    a.chunk = b->s_chunk;
    a.stage = s_stage_for(TS, sizeof(T));
    a.tiles_per_group = tpg;
    a.q = b->q_o;
    a.tab = b->tab;
    a.part = b->sk_part;
    a.Rpad = p.Rpad;
    a.SC = SC;
    a.items = b->s_items;
    const size_t smem = 2 * (size_t)a.stage * TS * sizeof(Cx<T>) + 2 * (size_t)a.stage * sizeof(T);
    launch_sk_rt<T, kRT>(b->stream, a, dim3(b->n_sitems, groups), smem, SC / 16, waves);
  } else {
    const dim3 hgrid(grid_for((long)(b->maxn + kTabAtoms - 1) / kTabAtoms * kTabAtoms), b->B, 3);
    kb_tables_h<<<hgrid, kBlock, 0, b->stream>>>(b->d_rep, p.kx, p.ky, p.kz, b->x, b->y, b->z,
                                                 b->q_o, (float)b->sA, (unsigned char *)b->tab,
                                                 (unsigned char *)b->ztab);
    SArgsH a;
    a.n = 0;
    a.TS = p.nX + p.nY;
    a.nX = p.nX;
    a.nY = p.nY;
    a.nZ = p.nZ;
    a.Rtiles = p.Rpad / 16;
    a.chunk = b->s_chunk;
    a.stage = kTabAtoms;
    a.tiles_per_group = 0;
    a.tab = b->tab;
    a.ztab = b->ztab;
    a.part = reinterpret_cast<float *>(b->sk_part);
    a.Rpad = p.Rpad;
    a.SC = SC;
    a.items = b->s_items;
    const int RT = b->s_rt;
    s_layout_h(p.Rpad, RT, &groups, &tpg, &waves);
    a.tiles_per_group = tpg;
    const size_t smem = split_sk_lds(p);
    if (RT == 1)
      batch_sk_h<1>(b, a, groups, smem, split_sct(p), waves);
    else if (RT == 2)
      batch_sk_h<2>(b, a, groups, smem, split_sct(p), waves);
    else
      batch_sk_h<3>(b, a, groups, smem, split_sct(p), waves);
  }
  const int nthr = p.Rpad * p.nZ;
  // partials: FP32 from the split S pass, FP64 from the FP32 / FP64 tiles
  if (split)
    k_sk_reduce<T, float><<<dim3(grid_for(nthr), b->B), kBlock, 0, b->stream>>>(
        0, p.Rpad, SC, p.R, p.nZ, p.kz, p.KSS, p.KSP, reinterpret_cast<const float *>(b->sk_part),
        b->d_W, b->sk, (T *)b->bmat, b->epart, 1.0 / b->sA, (_Float16 *)b->bmat_h, b->nkh, b->sB,
        b->rchunks, b->bm_stride, b->bmh_stride);
  else
    k_sk_reduce<T, double><<<dim3(grid_for(nthr), b->B), kBlock, 0, b->stream>>>(
        0, p.Rpad, SC, p.R, p.nZ, p.kz, p.KSS, p.KSP, b->sk_part, b->d_W, b->sk, (T *)b->bmat,
        b->epart, 1.0, nullptr, b->nkh, 1.0, b->rchunks, b->bm_stride, b->bmh_stride);
  if (!split) {
    FArgs f;
    f.n = 0;
    f.TS = TS;
    f.nX = p.nX;
    f.nY = p.nY;
    f.nZ = p.nZ;
    f.ky = p.ky;
    f.R = p.R;
    f.Rpad = p.Rpad;
    f.KZ = p.KZ;
    f.KSP = p.KSP;
    f.tab = b->tab;
    f.bm = b->bmat;
    f.q = b->q_o;
    f.minv = b->d_minv;
    f.type = b->type_o;
    f.c0 = f.kxs = f.kys = f.kzs = 0;
    f.ax = b->ax;
    f.ay = b->ay;
    f.az = b->az;
    f.out4 = nullptr;
    f.items = b->f_items;
    f.rep = b->d_rep;
    f.bm_stride = b->bm_stride * (long)sizeof(T);
    launch_kforce_t<T>(b->stream, f, p, b->n_fitems);
  } else {
    FArgsH f;
    f.n = 0;
    f.TS = p.nX + p.nY;
    f.nX = p.nX;
    f.nY = p.nY;
    f.nZ = p.nZ;
    f.ky = p.ky;
    f.kz = p.kz;
    f.R = p.R;
    f.Rpad = p.Rpad;
    f.nkh = b->nkh;
    f.tab = b->tab;
    f.ztab = b->ztab;
    f.bmh = b->bmat_h;
    f.minv = b->d_minv;
    f.type = b->type_o;
    f.c0 = f.kxs = f.kys = f.kzs = 0;
    f.ax = b->ax;
    f.ay = b->ay;
    f.az = b->az;
    f.items = b->f_items;
    f.rep = b->d_rep;
    f.bm_stride = b->bmh_stride * (long)sizeof(_Float16);
    f.inv_sb = 1.0 / b->sB / b->sA;
    if (b->f_nt == 1)
      launch_kforce_h<1, kWavesF>(b->stream, f, b->nkh, b->n_fitems, split_force_lds_p(p, b->nkh, 1));
    else
      launch_kforce_h<2, kWavesF>(b->stream, f, b->nkh, b->n_fitems, split_force_lds_p(p, b->nkh, 2));
  }
  BAT_HIP(hipGetLastError());
  return CAV_OK;
}


// Per replica, fixed-order sum of its S-pass chunk partials into [B][plane].
__global__ void kb_chunk_sum(const int2 *rchunks, const double *part, long plane, double *out)
{
  __shared__ double red[4][64];
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long e = (long)blockIdx.x * 64 + lane;
  const int2 rc = rchunks[blockIdx.y];
  double s = 0;
  if (e < plane)
    for (int k = wave; k < rc.y; k += 4) s += part[(long)(rc.x + k) * plane + e];
  red[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && e < plane)
    out[(long)blockIdx.y * plane + e] = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
}

// Per-atom k-space field of every replica (C sum k Im(g e^{ik.r}) and the
// potential, without charge or mass) into out4[owned][4], from the operand of
// the last force evaluation; rebuilds full FP32/FP64 phase tables first.
template <typename T>
static int batch_field_t(cav_batch *b, double *out4)
{
  const EwaldKPlan &p = b->ek;
  const int TS = p.nX + p.nY + p.nZ;
  kb_tables<T><<<dim3(grid_for((long)b->maxn * TS), b->B), kBlock, 0, b->stream>>>(
      b->d_rep, p.kx, p.ky, p.kz, b->x, b->y, b->z, (Cx<T> *)b->tab);
  FArgs f;
  f.n = 0;
  f.TS = TS;
  f.nX = p.nX;
  f.nY = p.nY;
  f.nZ = p.nZ;
  f.ky = p.ky;
  f.R = p.R;
  f.Rpad = p.Rpad;
  f.KZ = p.KZ;
  f.KSP = p.KSP;
  f.tab = b->tab;
  f.bm = b->bmat;
  f.q = b->q_o;
  f.minv = b->d_minv;
  f.type = b->type_o;
  f.c0 = f.kxs = f.kys = f.kzs = 0;
  f.ax = f.ay = f.az = nullptr;
  f.out4 = out4;
  f.items = b->f_items;
  f.rep = b->d_rep;
  f.bm_stride = b->bm_stride * (long)sizeof(T);
  launch_kforce_t<T>(b->stream, f, p, b->n_fitems);
  BAT_HIP(hipGetLastError());
  return CAV_OK;
}

int batch_ewald_k_field(cav_batch *b, double *out4)
{
  if (b->precision == CAV_FP64) return batch_field_t<double>(b, out4);
  return batch_field_t<float>(b, out4);
}

// S over the half lattice with per-atom weights u (owned-indexed) for every
// replica, into out[B][Rpad][SC] (re, im at columns 2c, 2c+1).  Uses the full
// phase tables of batch_ewald_k_field.
template <typename T>
static int batch_wsk_t(cav_batch *b, const double *u, double *out)
{
  const EwaldKPlan &p = b->ek;
  const int TS = p.nX + p.nY + p.nZ;
  const int SC = roundup(2 * p.nZ, 16);
  int groups, tpg, waves;
  s_layout_rt(p.Rpad, kRT, &groups, &tpg, &waves);
  SArgs a;
  a.n = 0;
  a.TS = TS;
  a.nX = p.nX;
  a.nY = p.nY;
  a.nZ = p.nZ;
  a.ky = p.ky;
  a.kz = p.kz;
  a.Rtiles = p.Rpad / 16;
  a.chunk = b->s_chunk;
  a.stage = s_stage_for(TS, sizeof(T));
  a.tiles_per_group = tpg;
  a.q = u;
  a.tab = b->tab;
  a.part = b->sk_part;
  a.Rpad = p.Rpad;
  a.SC = SC;
  a.items = b->s_items_o;
  const size_t smem = 2 * (size_t)a.stage * TS * sizeof(Cx<T>) + 2 * (size_t)a.stage * sizeof(T);
  launch_sk_rt<T, kRT>(b->stream, a, dim3(b->n_sitems, groups), smem, SC / 16, waves);
  const long plane = (long)p.Rpad * SC;
  kb_chunk_sum<<<dim3((plane + 63) / 64, b->B), kBlock, 0, b->stream>>>(b->rchunks, b->sk_part, plane, out);
  BAT_HIP(hipGetLastError());
  return CAV_OK;
}

int batch_weighted_sk(cav_batch *b, const double *u, double *out)
{
  if (b->precision == CAV_FP64) return batch_wsk_t<double>(b, u, out);
  return batch_wsk_t<float>(b, u, out);
}

// ------------------------------------------------------------------ autotuner --

// Times the Ewald_k passes (tables, S pass, reduction, force pass) for each
// candidate tile shape on this context's own system, as cav_step runs them
// (concurrently with a list build and the pair kernel), and keeps the fastest:
// S-pass row tiles per wave, atoms per LDS stage and atoms per chunk, and in
// the split form the force pass's atom tiles per wave.
int ewald_k_autotune(cav_ctx *c)
{
  struct Cand { int rt, stage, chunk, fnt, fw; };
  std::vector<Cand> cands;
  // atoms per S-pass chunk: FP32 accumulation within a chunk (its rounding
  // error grows as sqrt(chunk)), FP64 across chunks; fewer chunks, fewer partials
  const int chunks[5] = {256, 512, 1024, 2048, 4096};
  if (c->precision == CAV_MIXED && split_fits(c->ek, c->nkh, 1, kWavesF)) {
    // the split S pass stages one 32-atom table block; 4096-atom chunks only
    // pay beside the concurrent list build, where they tie with 2048 (within
    // the timing spread) while running 2x slower in order
    for (int rt : {1, 2, 3, 4})
      for (int ch : {256, 512, 1024, 2048})
        for (int nt : {1, 2})
          for (int fw : {kWavesF, 16}) cands.push_back({rt, kTabAtoms, ch, nt, fw});
  } else {
    for (int ch : chunks) cands.push_back({3, 0, ch, 2, kWavesF});
  }
  hipEvent_t e0, e1;
  CAV_HIP(hipEventCreate(&e0));
  CAV_HIP(hipEventCreate(&e1));
  const bool timing = c->timing;
  c->timing = false;
  // Two rounds: every candidate timed over 3 launches, then the 4 fastest
  // again over 10 each (single timings of ~1 ms phases vary by a few percent)
  auto apply = [&](const Cand &k) -> int {
    c->tune_rt = k.rt;
    c->tune_stage = k.stage;
    c->tune_fnt = k.fnt;
    c->tune_fw = k.fw;
    if (c->tune_chunk != k.chunk) {
      c->tune_chunk = k.chunk;
      CAV_TRY(ewald_k_alloc(c));
    }
    return CAV_OK;
  };
  auto time_it = [&](int reps, float *ms) -> int {
    for (int r = 0; r < 2; ++r) CAV_TRY(tuning_phase(c));
    CAV_HIP(hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps; ++r) CAV_TRY(tuning_phase(c));
    CAV_HIP(hipEventRecord(e1, c->stream));
    CAV_HIP(hipEventSynchronize(e1));
    CAV_HIP(hipEventElapsedTime(ms, e0, e1));
    *ms /= (float)reps;
    return CAV_OK;
  };
  std::vector<std::pair<float, int>> timed;
  for (int q = 0; q < (int)cands.size(); ++q) {
    CAV_TRY(apply(cands[q]));
    if (c->precision == CAV_MIXED && cands[q].stage == kTabAtoms &&
        !split_fits(c->ek, c->nkh, c->tune_fnt, c->tune_fw))
      continue;
    float ms = 0;
    CAV_TRY(time_it(3, &ms));
    timed.push_back({ms, q});
  }
  std::sort(timed.begin(), timed.end());
  Cand pick = {c->tune_rt, c->tune_stage, c->tune_chunk, c->tune_fnt, c->tune_fw};
  std::vector<std::pair<float, int>> finals;
  float best = 1e30f;
  for (int q = 0; q < (int)timed.size() && q < 4; ++q) {
    CAV_TRY(apply(cands[timed[q].second]));
    float ms = 0;
    CAV_TRY(time_it(10, &ms));
    finals.push_back({ms, timed[q].second});
    best = std::min(best, ms);
  }
  // within 0.5% of the fastest the smallest chunk wins (fewer atoms per FP32
  // accumulation), then the fewest S-pass row tiles (the S pass runs faster
  // in order with fewer); a tie is below the run-to-run spread of these timings
  auto preferred = [&](int a, int b) {
    return cands[a].chunk < cands[b].chunk ||
           (cands[a].chunk == cands[b].chunk && cands[a].rt < cands[b].rt);
  };
  int pq = -1;
  for (const auto &f : finals)
    if (f.first <= best * 1.005f && (pq < 0 || preferred(f.second, pq))) pq = f.second;
  if (pq >= 0) pick = cands[pq];
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  c->timing = timing;
  c->tune_rt = pick.rt;
  c->tune_stage = pick.stage;
  c->tune_fnt = pick.fnt;
  c->tune_fw = pick.fw;
  c->tune_chunk = pick.chunk;
  c->tuned = true;
  return ewald_k_alloc(c);
}

// Ewald_k accelerations of every replica, added into the owned accelerations.
int batch_ewald_k(cav_batch *b)
{
  if (!b->ek_on) return CAV_OK;
  if (b->precision == CAV_FP64) return batch_ewald_k_t<double>(b);
  return batch_ewald_k_t<float>(b);
}

}  // namespace cav
