// Force-matching adjoint: the vector-Jacobian product of the forces with
// respect to the force-field parameters,
//
//   g_theta = sum_i w_i . dF_i/dtheta,   w_i = dL/dF_i (given),
//
// i.e. the mixed second derivatives d2U/(dr dtheta) contracted with the
// residual weights.  Parameters: lj_cell_list epsilon[a][b] and sigma[a][b]
// (including the sigma dependence of the force shift at rc), and the type
// charges through Electrostatic_ewald_r and Electrostatic_ewald_k.  The
// reference has no differentiation; this is the gradient of its force formulas
// (lj_cell_list.cpp:143-155, electrostatic_ewald_r.cpp:183-211,
// electrostatic_ewald_k.cpp:197-226).
#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "cav_internal.hpp"

namespace cav {

// Ewald_k pieces exposed by cav_ewald_k.hip
int ewald_k_weighted_sk(cav_ctx *c, const double *weights, double *sk_out);
int ewald_k_field(cav_ctx *c, double *out4);

constexpr int kLPAa = 16;

static int roundup16(int v) { return (v + 15) / 16 * 16; }

struct AdjArgs {
  int n, nt, cap;
  const double4 *xq;
  const int2 *tid;
  const int *mol, *nl, *nn;
  const double *w;        // 3n residual weights (AoS)
  const double *eps, *sig;
  double lj_rc, lj_rc2;
  int lj_on, er_on, lambda_set;
  double alpha, a2, ke, two_a_rpi;
  const double *lambda, *charge;
  double *part;  // [block][nvals]: g_eps[nt2] | g_sig[nt2] | g_q[nt]
  const RepDev *rep;  // batched replicas (blockIdx.y) or nullptr
};

// One list entry: force on i is coef * (r_j - r_i); its derivative along the
// residual weight w_i is (w_i . r_ij) dcoef/dtheta.
//
// NT (the type count, <= 4) is a template parameter so that every per-lane
// accumulator has a compile-time index and lives in registers: for a fixed
// atom i the type-pair entry depends only on the partner's type u and on
// whether the partner is the lower owned index (the reference's row rule), so
// the LJ terms accumulate per (lower, u) and the charge terms per u; they are
// mapped to the [eps | sigma | charge] entries once per atom.
template <int NT>
__global__ __launch_bounds__(kBlock) void k_adjoint_pairs(AdjArgs a)
{
  constexpr int NT2 = NT * NT, NV = 2 * NT2 + NT;
  extern __shared__ double sh[];  // [kBlock/kLPAa][NV] per-atom-group partials
  if (a.rep) {  // one replica per grid row, arrays shifted to it
    const RepDev &r = a.rep[blockIdx.y];
    a.n = r.n;
    a.xq += r.base;
    a.tid += r.base;
    a.nl += (long)r.obase * a.cap;
    a.nn += r.obase;
    a.w += 3L * r.obase;
    a.part += (long)blockIdx.y * gridDim.x * NV;
  }
  const long gid = (long)blockIdx.x * kBlock + threadIdx.x;
  const int t = (int)(gid / kLPAa), sub = (int)(gid % kLPAa);
  // [lower][u]: d/d eps and d/d sigma of entry (lower ? u*NT+ti : ti*NT+u)
  double ge[2][NT], gs[2][NT];
  double gqi = 0, gqj[NT];  // d/dq of type ti (this atom) and of partner type u
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    ge[0][u] = ge[1][u] = gs[0][u] = gs[1][u] = 0;
    gqj[u] = 0;
  }
  int ti = 0;
  if (t < a.n) {
    const int i = t;
    const double4 pi4 = a.xq[i];
    ti = a.tid[i].x;
    const int moli = a.lambda_set ? a.mol[i] : 0;
    const double wx = a.w[3 * i], wy = a.w[3 * i + 1], wz = a.w[3 * i + 2];
    const double qti = a.charge[ti];
    const int cnt = a.nn[i];
    const int *row = a.nl + (long)i * a.cap;
    for (int k = sub; k < cnt; k += kLPAa) {
      const int j = row[k];
      const double4 pj = a.xq[j];
      const double dx = pj.x - pi4.x, dy = pj.y - pi4.y, dz = pj.z - pi4.z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double wr = wx * dx + wy * dy + wz * dz;  // w_i . r_ij
      const int tj = a.tid[j].x;
      // type-pair entry as the reference indexes it: [lower owned index][partner]
      const bool lower = j < a.n && j < i;
      const int p = lower ? tj * NT + ti : ti * NT + tj;
      if (a.lj_on && !(r2 > a.lj_rc2)) {
        const double e = a.eps[p], s = a.sig[p];
        const double ri2 = 1.0 / r2, rci2 = 1.0 / a.lj_rc2;
        const double s2 = s * s, s5 = s2 * s2 * s, s6 = s5 * s, s11 = s6 * s5, s12 = s6 * s6;
        const double r6 = ri2 * ri2 * ri2, r8 = r6 * ri2, r14 = r6 * r6 * ri2;
        const double c6 = rci2 * rci2 * rci2, c8 = c6 * rci2, c14 = c6 * c6 * rci2;
        // coef = 4e(-12 s^12 r^-14 + 6 s^6 r^-8 + 12 s^12 rc^-14 - 6 s^6 rc^-8)
        const double dce = wr * 4.0 * (-12 * s12 * r14 + 6 * s6 * r8 + 12 * s12 * c14 - 6 * s6 * c8);
        const double dcs = wr * 4.0 * e * (-144 * s11 * r14 + 36 * s5 * r8 + 144 * s11 * c14 - 36 * s5 * c8);
#pragma unroll
        for (int u = 0; u < NT; ++u) {
          const bool hit = tj == u;
          ge[0][u] += (hit && !lower) ? dce : 0.0;
          ge[1][u] += (hit && lower) ? dce : 0.0;
          gs[0][u] += (hit && !lower) ? dcs : 0.0;
          gs[1][u] += (hit && lower) ? dcs : 0.0;
        }
      }
      if (a.er_on && !(dx == 0 && dy == 0 && dz == 0)) {
        double lam = 1.0;
        bool skip = false;
        if (a.lambda_set && moli == a.mol[j]) {
          lam = a.lambda[p];
          if (lam == 0.0) skip = true;
        }
        if (!skip) {
          const double rn = sqrt(r2);
          const double sr = (a.two_a_rpi * exp(-a.a2 * r2) + erfc(a.alpha * rn) / rn) / r2;
          // coef = -lam ke q_a q_b sr: d/dq_a = -lam ke q_b sr, d/dq_b = -lam ke q_a sr
          const double base = -lam * a.ke * sr * wr;
          double qtj = 0;
#pragma unroll
          for (int u = 0; u < NT; ++u) qtj = tj == u ? a.charge[u] : qtj;
          gqi += base * qtj;
#pragma unroll
          for (int u = 0; u < NT; ++u) gqj[u] += tj == u ? base * qti : 0.0;
        }
      }
    }
  }
  // fixed-order reduction: the 16 lanes of an atom (shuffle tree), then the
  // block's atoms in group order
#pragma unroll
  for (int o = 1; o < kLPAa; o <<= 1) {
#pragma unroll
    for (int u = 0; u < NT; ++u) {
      ge[0][u] += __shfl_xor(ge[0][u], o, 64);
      ge[1][u] += __shfl_xor(ge[1][u], o, 64);
      gs[0][u] += __shfl_xor(gs[0][u], o, 64);
      gs[1][u] += __shfl_xor(gs[1][u], o, 64);
      gqj[u] += __shfl_xor(gqj[u], o, 64);
    }
    gqi += __shfl_xor(gqi, o, 64);
  }
  const int grp = threadIdx.x / kLPAa;
  if (sub == 0) {
    double *g = sh + grp * NV;
    for (int v = 0; v < NV; ++v) g[v] = 0;
#pragma unroll
    for (int u = 0; u < NT; ++u) {
      g[ti * NT + u] += ge[0][u];
      g[u * NT + ti] += ge[1][u];
      g[NT2 + ti * NT + u] += gs[0][u];
      g[NT2 + u * NT + ti] += gs[1][u];
      g[2 * NT2 + u] += gqj[u];
    }
    g[2 * NT2 + ti] += gqi;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    double x = 0;
    for (int gi = 0; gi < kBlock / kLPAa; ++gi) x += sh[gi * NV + threadIdx.x];
    a.part[(long)blockIdx.x * NV + threadIdx.x] = x;
  }
}

static void launch_adjoint_pairs(int nt, dim3 grid, size_t sh, hipStream_t st, const AdjArgs &a)
{
  switch (nt) {
  case 1: k_adjoint_pairs<1><<<grid, kBlock, sh, st>>>(a); break;
  case 2: k_adjoint_pairs<2><<<grid, kBlock, sh, st>>>(a); break;
  case 3: k_adjoint_pairs<3><<<grid, kBlock, sh, st>>>(a); break;
  default: k_adjoint_pairs<4><<<grid, kBlock, sh, st>>>(a); break;
  }
}

__global__ void k_reduce_cols(int rows, int cols, const double *part, double *out)
{
  const int v = threadIdx.x;
  if (v >= cols) return;
  double x = 0;
  for (int r = 0; r < rows; ++r) x += part[(long)r * cols + v];
  out[v] = x;
}

// k-space term of dL/dq_s through S: C sum_H W Im(conj(S^(s)) T), with
// T = sum_d (2 pi n_d / l_d) S^(q w_d).  Per-thread partial over (row, col).
__global__ void k_adjoint_kspace(int R, int Rpad, int nY, int nZ, int ky, int kz, int nt,
                                 double kxs, double kys, double kzs, const double *W,
                                 const double *ss, const double *sw, double *part)
{
  const int t = blockIdx.x * kBlock + threadIdx.x;
  __shared__ double red[kBlock];
  const long plane = (long)Rpad * 2 * nZ;  // doubles per weighted S
  for (int s = 0; s < nt; ++s) {
    double v = 0;
    if (t < R * nZ) {
      const int r = t / nZ, c = t - r * nZ;
      const double w = W[(long)r * nZ + c];
      if (w != 0) {
        const int nx = r / nY, ny = r % nY - ky, nz = c - kz;
        const long o = (long)r * 2 * nZ + 2 * c;
        const double tr = kxs * nx * sw[o] + kys * ny * sw[plane + o] + kzs * nz * sw[2 * plane + o];
        const double ti = kxs * nx * sw[o + 1] + kys * ny * sw[plane + o + 1] +
                          kzs * nz * sw[2 * plane + o + 1];
        const double sr = ss[s * plane + o], si = ss[s * plane + o + 1];
        v = w * (sr * ti - si * tr);  // Im(conj(S) T)
      }
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int q = kBlock / 2; q > 0; q >>= 1) {
      if (threadIdx.x < q) red[threadIdx.x] += red[threadIdx.x + q];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.x * nt + s] = red[0];
    __syncthreads();
  }
}

// Per-atom term: sum over atoms of species s of w_i . G_i (G = k-space field).
__global__ void k_adjoint_field(int n, int nt, const double *w, const double *f4, const int *type,
                                double *part)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  __shared__ double red[kBlock];
  for (int s = 0; s < nt; ++s) {
    double v = 0;
    if (i < n && type[i] == s)
      v = w[3 * i] * f4[4 * i] + w[3 * i + 1] * f4[4 * i + 1] + w[3 * i + 2] * f4[4 * i + 2];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int q = kBlock / 2; q > 0; q >>= 1) {
      if (threadIdx.x < q) red[threadIdx.x] += red[threadIdx.x + q];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(long)blockIdx.x * nt + s] = red[0];
    __syncthreads();
  }
}

__global__ void k_fill_weights(int n, int mode, int s, const int *type, const double *q,
                               const double *w, double *u)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  u[i] = mode < 0 ? (type[i] == s ? 1.0 : 0.0) : q[i] * w[3 * i + mode];
}


// ---------------------------------------------------------- batched replicas --

__global__ void kb_reduce_cols(int rows, int cols, const double *part, double *out)
{
  const int v = threadIdx.x;
  if (v >= cols) return;
  const double *p = part + (long)blockIdx.x * rows * cols;
  double x = 0;
  for (int r = 0; r < rows; ++r) x += p[(long)r * cols + v];
  out[(long)blockIdx.x * cols + v] = x;
}

// k_adjoint_kspace for every replica (grid row): planes [B][Rpad][SC]
__global__ void kb_adjoint_kspace(const RepDev *rep, int R, int Rpad, int SC, int nY, int nZ, int ky,
                                  int kz, int nt, const double *W, const double *ss,
                                  const double *sw, double *part)
{
  const int b = blockIdx.y, B = gridDim.y;
  const RepDev &rr = rep[b];
  const double kxs = rr.kxs, kys = rr.kys, kzs = rr.kzs;
  W += (long)b * R * nZ;
  const long plane = (long)Rpad * SC;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  __shared__ double red[kBlock];
  for (int s = 0; s < nt; ++s) {
    double v = 0;
    if (t < R * nZ) {
      const int r = t / nZ, c = t - r * nZ;
      const double w = W[(long)r * nZ + c];
      if (w != 0) {
        const int nx = r / nY, ny = r % nY - ky, nz = c - kz;
        const long o = (long)r * SC + 2 * c;
        const double *w0 = sw + ((long)0 * B + b) * plane, *w1 = sw + ((long)1 * B + b) * plane,
                     *w2 = sw + ((long)2 * B + b) * plane;
        const double tr = kxs * nx * w0[o] + kys * ny * w1[o] + kzs * nz * w2[o];
        const double ti = kxs * nx * w0[o + 1] + kys * ny * w1[o + 1] + kzs * nz * w2[o + 1];
        const double *sp = ss + ((long)s * B + b) * plane;
        v = w * (sp[o] * ti - sp[o + 1] * tr);
      }
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int q = kBlock / 2; q > 0; q >>= 1) {
      if (threadIdx.x < q) red[threadIdx.x] += red[threadIdx.x + q];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[((long)b * gridDim.x + blockIdx.x) * nt + s] = red[0];
    __syncthreads();
  }
}

// k_adjoint_field for every replica (grid row)
__global__ void kb_adjoint_field(const RepDev *rep, int nt, const double *w, const double *f4,
                                 const int *type_o, double *part)
{
  const RepDev &r = rep[blockIdx.y];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  __shared__ double red[kBlock];
  for (int s = 0; s < nt; ++s) {
    double v = 0;
    if (i < r.n) {
      const long o = r.obase + i;
      if (type_o[o] == s) v = w[3 * o] * f4[4 * o] + w[3 * o + 1] * f4[4 * o + 1] + w[3 * o + 2] * f4[4 * o + 2];
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int q = kBlock / 2; q > 0; q >>= 1) {
      if (threadIdx.x < q) red[threadIdx.x] += red[threadIdx.x + q];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[((long)blockIdx.y * gridDim.x + blockIdx.x) * nt + s] = red[0];
    __syncthreads();
  }
}

__global__ void kb_fill_weights(long nown, int mode, int s, const int *type_o, const double *q_o,
                                const double *w, double *u)
{
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nown) return;
  u[i] = mode < 0 ? (type_o[i] == s ? 1.0 : 0.0) : q_o[i] * w[3 * i + mode];
}
}  // namespace cav

using namespace cav;

// Device buffers of one adjoint call, in doubles: residual weights, per-block
// partials, reduced outputs, and for Ewald_k the per-atom weights, the species
// and residual-weighted structure factors and the per-atom k-space field.
struct AdjWork {
  double *w = nullptr, *part = nullptr, *out = nullptr, *u = nullptr, *ss = nullptr,
         *sw = nullptr, *f4 = nullptr;
};

struct AdjSizes {
  size_t w = 0, part = 0, out = 0, u = 0, ss = 0, sw = 0, f4 = 0;
};

static AdjSizes adjoint_sizes(const cav_ctx *c)
{
  const int n = c->n, nt = c->nt, nv = 2 * nt * nt + nt;
  const int nblk = grid_for((long)n * kLPAa);
  AdjSizes z;
  z.w = 3 * (size_t)n + 1;
  z.part = (size_t)nblk * nv + 4 * (size_t)grid_for(n) * nt +
           (size_t)grid_for((long)c->ek.R * c->ek.nZ + 1) * nt + 64;
  z.out = nv + 2 * (size_t)nt;
  if (c->ek_on) {
    const size_t plane = (size_t)c->ek.Rpad * 2 * c->ek.nZ;
    z.u = (size_t)n + 1;
    z.ss = plane * nt;
    z.sw = plane * 3;
    z.f4 = 4 * (size_t)(n + 1);
  }
  return z;
}

static int adjoint_checks(cav_ctx *c)
{
  if (!c->setup_done) return fail(c, CAV_ERR_STATE, "cav_setup must run before cav_adjoint");
  if (c->lj_on && c->lj_cellwalk)
    return fail(c, CAV_ERR_STATE, "the adjoint needs the LJ cutoff within the neighbour-list cutoff");
  if (c->dipole) return fail(c, CAV_ERR_STATE, "the adjoint does not cover the epsilon_dipole term");
  if (c->ljl_on || c->sr_on)
    return fail(c, CAV_ERR_STATE, "the adjoint covers lj_cell_list, ewald_r and ewald_k only");
  if (c->nt > 4) return fail(c, CAV_ERR_ARG, "the adjoint supports up to 4 atom types");
  return CAV_OK;
}

// The adjoint on a given workspace (every buffer written before it is read,
// so its previous contents do not matter).  hout has nv + 2 nt doubles.
static int adjoint_run(cav_ctx *c, const double *w_host, const AdjWork &k, double *hout)
{
  const int n = c->n, nt = c->nt, nt2 = nt * nt, nv = 2 * nt2 + nt;
  const int nblk = grid_for((long)n * kLPAa);
  CAV_HIP(hipMemcpyAsync(k.w, w_host, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  for (int v = 0; v < nv + 2 * nt; ++v) hout[v] = 0.0;
  if (c->lj_on || c->er_on) {
    CAV_TRY(launch_nl_fill(c));  // the index rows, if the last build wrote slots only
    AdjArgs a;
    a.n = n;
    a.nt = nt;
    a.cap = c->nl_cap;
    a.xq = c->xq;
    a.tid = c->tid;
    a.mol = c->mol;
    a.nl = c->nl;
    a.nn = c->nn;
    a.w = k.w;
    a.eps = c->d_raw_eps;
    a.sig = c->d_raw_sig;
    a.lj_rc = c->lj_rc;
    a.lj_rc2 = c->lj_rc * c->lj_rc;
    a.lj_on = c->lj_on;
    a.er_on = c->er_on;
    a.lambda_set = c->lambda_set;
    a.alpha = c->er_alpha;
    a.a2 = c->er_alpha * c->er_alpha;
    a.ke = c->er_ke;
    a.two_a_rpi = 2 * c->er_alpha * CAV_PIS_INV;
    a.lambda = c->d_lambda;
    a.charge = c->d_charge;
    a.part = k.part;
    a.rep = nullptr;
    const size_t sh = sizeof(double) * (kBlock / kLPAa) * nv;
    launch_adjoint_pairs(c->nt, dim3(nblk), sh, c->stream, a);
    k_reduce_cols<<<1, 64, 0, c->stream>>>(nblk, nv, k.part, k.out);
    CAV_HIP(hipGetLastError());
    CAV_HIP(hipMemcpyAsync(hout, k.out, sizeof(double) * nv, hipMemcpyDeviceToHost, c->stream));
    CAV_HIP(hipStreamSynchronize(c->stream));
  }
  if (c->ek_on) {
    const EwaldKPlan &p = c->ek;
    const size_t plane = (size_t)p.Rpad * 2 * p.nZ;
    // per-atom k-space field G_i (uses the operand of the last force evaluation)
    CAV_TRY(ewald_k_field(c, k.f4));
    for (int s = 0; s < nt; ++s) {
      k_fill_weights<<<grid_for(n), kBlock, 0, c->stream>>>(n, -1, s, c->type, c->q, k.w, k.u);
      CAV_TRY(ewald_k_weighted_sk(c, k.u, k.ss + s * plane));
    }
    for (int d = 0; d < 3; ++d) {
      k_fill_weights<<<grid_for(n), kBlock, 0, c->stream>>>(n, d, 0, c->type, c->q, k.w, k.u);
      CAV_TRY(ewald_k_weighted_sk(c, k.u, k.sw + d * plane));
    }
    const int nb = grid_for((long)p.R * p.nZ);
    k_adjoint_kspace<<<nb, kBlock, 0, c->stream>>>(p.R, p.Rpad, p.nY, p.nZ, p.ky, p.kz, nt,
                                                   CAV_2PI / p.lx, CAV_2PI / p.ly, CAV_2PI / p.lz,
                                                   c->d_W, k.ss, k.sw, k.part);
    k_reduce_cols<<<1, 64, 0, c->stream>>>(nb, nt, k.part, k.out + nv);
    const int nb2 = grid_for(n);
    double *part2 = k.part + (size_t)nb * nt + 16;
    k_adjoint_field<<<nb2, kBlock, 0, c->stream>>>(n, nt, k.w, k.f4, c->type, part2);
    k_reduce_cols<<<1, 64, 0, c->stream>>>(nb2, nt, part2, k.out + nv + nt);
    CAV_HIP(hipGetLastError());
    CAV_HIP(hipMemcpyAsync(hout + nv, k.out + nv, sizeof(double) * 2 * nt,
                           hipMemcpyDeviceToHost, c->stream));
    CAV_HIP(hipStreamSynchronize(c->stream));
    const double C = CAV_4PI * c->ek_ke * p.lxyz_inv;
    for (int s = 0; s < nt; ++s) hout[2 * nt2 + s] += C * hout[nv + s] + hout[nv + nt + s];
  }
  return CAV_OK;
}

static void adjoint_outputs(const cav_ctx *c, const double *hout, double *g_eps, double *g_sigma,
                            double *g_charge)
{
  const int nt = c->nt, nt2 = nt * nt;
  if (g_eps)
    for (int p = 0; p < nt2; ++p) g_eps[p] = hout[p];
  if (g_sigma)
    for (int p = 0; p < nt2; ++p) g_sigma[p] = hout[nt2 + p];
  if (g_charge)
    for (int s = 0; s < nt; ++s) g_charge[s] = hout[2 * nt2 + s];
}

extern "C" int cav_adjoint(cav_ctx *c, const double *w_host, double *g_eps, double *g_sigma,
                           double *g_charge)
{
  if (!c || !w_host) return CAV_ERR_ARG;
  CAV_TRY(adjoint_checks(c));
  (void)hipSetDevice(c->device);
  // parameter values changed since the last force call: charges (owned and
  // ghosts), their packed copies and the k-space weights follow them
  const bool dirty = c->params_dirty;
  CAV_TRY(refresh_params(c));
  if (dirty && (c->lj_on || c->er_on)) CAV_TRY(launch_pack(c));
  // the workspace of this call, allocated here and released on every exit
  struct Owned {
    AdjWork k;
    ~Owned()
    {
      double *ps[] = {k.w, k.part, k.out, k.u, k.ss, k.sw, k.f4};
      for (double *q : ps)
        if (q) (void)hipFree(q);
    }
  } own;
  const AdjSizes z = adjoint_sizes(c);
  CAV_HIP(hipMalloc(&own.k.w, sizeof(double) * z.w));
  CAV_HIP(hipMalloc(&own.k.part, sizeof(double) * z.part));
  CAV_HIP(hipMalloc(&own.k.out, sizeof(double) * z.out));
  if (c->ek_on) {
    CAV_HIP(hipMalloc(&own.k.u, sizeof(double) * z.u));
    CAV_HIP(hipMalloc(&own.k.ss, sizeof(double) * z.ss));
    CAV_HIP(hipMalloc(&own.k.sw, sizeof(double) * z.sw));
    CAV_HIP(hipMalloc(&own.k.f4, sizeof(double) * z.f4));
  }
  std::vector<double> hout(z.out, 0.0);
  CAV_TRY(adjoint_run(c, w_host, own.k, hout.data()));
  adjoint_outputs(c, hout.data(), g_eps, g_sigma, g_charge);
  return CAV_OK;
}

// ---- the adjoint without allocation ---------------------------------------

// Workspace of cav_adjoint_noalloc: one array per GPU, held for the life of
// the process, shared by every call on that GPU.  Its host-side name is not a
// device address: host code uses only the address hipGetSymbolAddress gives
// for the current device (kept in adj_scratch_dev).
constexpr size_t kAdjScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) double scratch[kAdjScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
static std::mutex adj_scratch_mu;
static double *adj_scratch_dev[kMaxDevices];

// The device address of this GPU's scratch, looked up once per device with
// that device current.
static int adj_scratch(cav_ctx *c, double **out)
{
  if (c->device < 0 || c->device >= kMaxDevices) return fail(c, CAV_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lock(adj_scratch_mu);
  if (!adj_scratch_dev[c->device]) {
    double *ptr = nullptr;
    CAV_HIP(hipSetDevice(c->device));
    CAV_HIP(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(c, CAV_ERR_HIP, "no device address for the adjoint scratch");
    adj_scratch_dev[c->device] = ptr;
  }
  *out = adj_scratch_dev[c->device];
  return CAV_OK;
}

extern "C" size_t cav_adjoint_noalloc_scratch_bytes(void) { return kAdjScratchBytes; }

extern "C" int cav_adjoint_noalloc(cav_ctx *c, const double *w_host, double *g_eps,
                                   double *g_sigma, double *g_charge)
{
  if (!c || !w_host) return CAV_ERR_ARG;
  CAV_TRY(adjoint_checks(c));
  if (c->params_dirty)
    return fail(c, CAV_ERR_STATE,
                "parameter values changed since the last force evaluation: cav_adjoint_noalloc "
                "allocates nothing, so call cav_recalculate, cav_forces or cav_adjoint first");
  const AdjSizes z = adjoint_sizes(c);
  // each buffer starts on a 256-byte boundary (32 doubles)
  auto up = [](size_t v) { return (v + 31) / 32 * 32; };
  const size_t need = up(z.w) + up(z.part) + up(z.out) + up(z.u) + up(z.ss) + up(z.sw) + up(z.f4);
  if (need * sizeof(double) > kAdjScratchBytes)
    return fail(c, CAV_ERR_CAPACITY,
                "cav_adjoint_noalloc needs " + std::to_string(need * sizeof(double)) +
                    " bytes of workspace, above its fixed " + std::to_string(kAdjScratchBytes) +
                    "-byte scratch: use cav_adjoint");
  (void)hipSetDevice(c->device);
  double *base = nullptr;
  CAV_TRY(adj_scratch(c, &base));
  AdjWork k;
  double *q = base;
  k.w = q;
  q += up(z.w);
  k.part = q;
  q += up(z.part);
  k.out = q;
  q += up(z.out);
  if (c->ek_on) {
    k.u = q;
    q += up(z.u);
    k.ss = q;
    q += up(z.ss);
    k.sw = q;
    q += up(z.sw);
    k.f4 = q;
  }
  double hout[2 * 16 + 4 + 8];  // nv + 2 nt for nt <= 4
  CAV_TRY(adjoint_run(c, w_host, k, hout));
  adjoint_outputs(c, hout, g_eps, g_sigma, g_charge);
  return CAV_OK;
}

// Force-matching adjoint of every replica at its current positions: w holds
// the replicas' residual weights one after another (3 n_r doubles each, in
// replica order); outputs are [B][nt*nt], [B][nt*nt] and [B][nt].
extern "C" int cav_batch_adjoint(cav_batch *b, const double *w_host, double *g_eps, double *g_sigma,
                                 double *g_charge)
{
  if (!b || !w_host) return CAV_ERR_ARG;
  if (!b->setup_done) return bfail(b, CAV_ERR_STATE, "cav_batch_setup must run first");
  if (b->nt > 4) return bfail(b, CAV_ERR_ARG, "the adjoint supports up to 4 atom types");
  (void)hipSetDevice(b->device);
  const int nt = b->nt, nt2 = nt * nt, nv = 2 * nt2 + nt, B = b->B;
  const long nown = b->nown;
  const int pb = grid_for((long)b->maxn * kLPAa);
  const EwaldKPlan &p = b->ek;
  const long plane = b->ek_on ? (long)p.Rpad * roundup16(2 * p.nZ) : 1;
  const int kb = b->ek_on ? grid_for((long)p.R * p.nZ) : 1, fb = grid_for(b->maxn);
  double *dw = nullptr, *dpart = nullptr, *dout = nullptr, *u = nullptr, *ss = nullptr, *f4 = nullptr;
  const size_t npart = (size_t)B * std::max<long>((long)pb * nv, (long)(kb + fb) * nt) + 64;
  BAT_HIP(hipMalloc(&dw, sizeof(double) * 3 * nown));
  BAT_HIP(hipMalloc(&dpart, sizeof(double) * npart));
  BAT_HIP(hipMalloc(&dout, sizeof(double) * (size_t)B * (nv + 2 * nt)));
  BAT_HIP(hipMemcpyAsync(dw, w_host, sizeof(double) * 3 * nown, hipMemcpyHostToDevice, b->stream));
  std::vector<double> hpair((size_t)B * nv, 0.0), hk((size_t)B * nt, 0.0), hf((size_t)B * nt, 0.0);
  if (b->lj_on || b->er_on) {
    AdjArgs a;
    a.n = 0;
    a.nt = nt;
    a.cap = b->cap;
    a.xq = b->xq;
    a.tid = b->tid;
    a.mol = nullptr;
    a.nl = b->nl;
    a.nn = b->nn;
    a.w = dw;
    a.eps = b->d_raw_eps;
    a.sig = b->d_raw_sig;
    a.lj_rc = b->lj_rc;
    a.lj_rc2 = b->lj_rc * b->lj_rc;
    a.lj_on = b->lj_on;
    a.er_on = b->er_on;
    a.lambda_set = 0;
    a.alpha = b->er_alpha;
    a.a2 = b->er_alpha * b->er_alpha;
    a.ke = b->er_ke;
    a.two_a_rpi = 2 * b->er_alpha * CAV_PIS_INV;
    a.lambda = nullptr;
    a.charge = b->d_charge;
    a.part = dpart;
    a.rep = b->d_rep;
    const size_t sh = sizeof(double) * (kBlock / kLPAa) * nv;
    launch_adjoint_pairs(b->nt, dim3(pb, B), sh, b->stream, a);
    kb_reduce_cols<<<B, 64, 0, b->stream>>>(pb, nv, dpart, dout);
    BAT_HIP(hipGetLastError());
    BAT_HIP(hipMemcpyAsync(hpair.data(), dout, sizeof(double) * B * nv, hipMemcpyDeviceToHost, b->stream));
    BAT_HIP(hipStreamSynchronize(b->stream));
  }
  if (b->ek_on) {
    BAT_HIP(hipMalloc(&u, sizeof(double) * nown));
    BAT_HIP(hipMalloc(&ss, sizeof(double) * plane * B * (nt + 3)));
    BAT_HIP(hipMalloc(&f4, sizeof(double) * 4 * nown));
    CAV_TRY(batch_ewald_k_field(b, f4));
    for (int s = 0; s < nt; ++s) {
      kb_fill_weights<<<grid_for(nown), kBlock, 0, b->stream>>>(nown, -1, s, b->type_o, b->q_o, dw, u);
      CAV_TRY(batch_weighted_sk(b, u, ss + (size_t)s * B * plane));
    }
    double *sw = ss + (size_t)nt * B * plane;
    for (int d = 0; d < 3; ++d) {
      kb_fill_weights<<<grid_for(nown), kBlock, 0, b->stream>>>(nown, d, 0, b->type_o, b->q_o, dw, u);
      CAV_TRY(batch_weighted_sk(b, u, sw + (size_t)d * B * plane));
    }
    kb_adjoint_kspace<<<dim3(kb, B), kBlock, 0, b->stream>>>(b->d_rep, p.R, p.Rpad, roundup16(2 * p.nZ),
                                                             p.nY, p.nZ, p.ky, p.kz, nt, b->d_W, ss,
                                                             sw, dpart);
    kb_reduce_cols<<<B, 64, 0, b->stream>>>(kb, nt, dpart, dout + (size_t)B * nv);
    double *dpart2 = dpart + (size_t)B * kb * nt;
    kb_adjoint_field<<<dim3(fb, B), kBlock, 0, b->stream>>>(b->d_rep, nt, dw, f4, b->type_o, dpart2);
    kb_reduce_cols<<<B, 64, 0, b->stream>>>(fb, nt, dpart2, dout + (size_t)B * (nv + nt));
    BAT_HIP(hipGetLastError());
    BAT_HIP(hipMemcpyAsync(hk.data(), dout + (size_t)B * nv, sizeof(double) * B * nt, hipMemcpyDeviceToHost, b->stream));
    BAT_HIP(hipMemcpyAsync(hf.data(), dout + (size_t)B * (nv + nt), sizeof(double) * B * nt,
                           hipMemcpyDeviceToHost, b->stream));
    BAT_HIP(hipStreamSynchronize(b->stream));
  }
  for (int r = 0; r < B; ++r) {
    const double C = b->ek_on ? b->reps[r].c0 : 0.0;  // 4 pi ke / V of this replica
    for (int q = 0; q < nt2; ++q) {
      if (g_eps) g_eps[(size_t)r * nt2 + q] = hpair[(size_t)r * nv + q];
      if (g_sigma) g_sigma[(size_t)r * nt2 + q] = hpair[(size_t)r * nv + nt2 + q];
    }
    for (int s = 0; s < nt; ++s)
      if (g_charge)
        g_charge[(size_t)r * nt + s] = hpair[(size_t)r * nv + 2 * nt2 + s] + C * hk[(size_t)r * nt + s] +
                                       hf[(size_t)r * nt + s];
  }
  void *fr[] = {dw, dpart, dout, u, ss, f4};
  for (void *q : fr)
    if (q) (void)hipFree(q);
  return CAV_OK;
}
