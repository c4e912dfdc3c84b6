// Electrostatic layer correction (ELC) of Ewald sums in slab geometry: the
// reference's Electrostatic_ewald_slab_correction
// (src/caviar/objects/force_field/electrostatic_ewald_slab_correction.cpp and
// electrostatic_ewald_slab_correction/*.cpp).
//
// Per force evaluation:
//   1. chi sums over the owned atoms, one workgroup per in-plane wave vector
//      (kx, ky and the (kx, ky) pairs) with a fixed-order block reduction, and
//      one workgroup for the dipole sum (initialize.cpp:75-268);
//   2. per owned atom the field of the order-N scheme (field.cpp:29-136) times
//      the charge, added to the acceleration.
// Everything is FP64 in both precision forms: per atom the work is
// (kx + ky + kx ky) hyperbolic and circular functions, small beside Ewald_k.
//
// The dipole field of an evaluation comes from the previous evaluation's
// dipole sum (initialize.cpp:86 reads dipole_sum before updating it): the
// context keeps that sum on the device, zero before the first evaluation.
#include <algorithm>
#include <cmath>

#include "cav_internal.hpp"

namespace cav {

struct ElcArgs {
  int n, axis, kx, ky;
  const double *x, *y, *z, *q, *minv;
  const int *type;
  const double *kxv, *kxc, *kyv, *kyc, *kpv, *kpc;  // k grids and coefficients
  double *chi;   // [kx][4] | [ky][4] | [kx ky][8]
  double *dip;   // [0..2] dipole sum of the last evaluation (slab normal only),
                 // [3..5] its dipole field, [6..8] the sum before it
  double dip_coef, sc;  // dipole coefficient; 2 slab_sum_e_coef k_e
  double *ax, *ay, *az;
};

// give_slab_local_coordinates (electrostatic_ewald_slab_correction.h:55-77)
__device__ __forceinline__ void elc_local(int axis, double gx, double gy, double gz, double &lx,
                                          double &ly, double &lz)
{
  if (axis == 0) {
    lx = gy;
    ly = gz;
    lz = gx;
  } else if (axis == 1) {
    lx = gz;
    ly = gx;
    lz = gy;
  } else {
    lx = gx;
    ly = gy;
    lz = gz;
  }
}

// give_slab_global_coordinates (electrostatic_ewald_slab_correction.h:79-101)
// as the reference maps it: for a normal along x it applies the inverse of the
// y map and for y that of the x map.
__device__ __forceinline__ void elc_global(int axis, double lx, double ly, double lz, double &gx,
                                           double &gy, double &gz)
{
  if (axis == 0) {
    gx = ly;
    gy = lz;
    gz = lx;
  } else if (axis == 1) {
    gx = lz;
    gy = lx;
    gz = ly;
  } else {
    gx = lx;
    gy = ly;
    gz = lz;
  }
}

// Fixed-order sum of V per-thread values over the block (kBlock threads).
template <int V>
__device__ __forceinline__ void block_sum(double (&v)[V], double *out)
{
  __shared__ double red[kBlock / 64][V];
#pragma unroll
  for (int k = 0; k < V; ++k)
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < V; ++k) red[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < V; ++k) {
      double s = 0;
      for (int w = 0; w < kBlock / 64; ++w) s += red[w][k];
      out[k] = s;
    }
}

// chi sums (make_slab_chi_vectors, initialize.cpp:162-268) and the dipole sum
// (calculate_dipole_sum, initialize.cpp:75-117): block b < kx is chi_x[b],
// then ky blocks of chi_y, then kx ky blocks of chi_p ([a][b][c] at 4a+2b+c),
// and the last block the dipole.
__global__ __launch_bounds__(kBlock) void k_elc_chi(ElcArgs a)
{
  const int b = blockIdx.x, nx = a.kx, ny = a.ky, np = nx * ny;
  if (b < nx + ny) {
    const bool isx = b < nx;
    const double k = isx ? a.kxv[b] : a.kyv[b - nx];
    double v[4] = {0, 0, 0, 0};
    for (int i = threadIdx.x; i < a.n; i += kBlock) {
      double px, py, pz;
      elc_local(a.axis, a.x[i], a.y[i], a.z[i], px, py, pz);
      const double q = a.q[i];
      const double sh = sinh(k * pz), ch = cosh(k * pz);
      const double arg = isx ? px : py;
      const double sn = sin(k * arg), cs = cos(k * arg);
      v[0] += q * sh * sn;
      v[1] += q * sh * cs;
      v[2] += q * ch * sn;
      v[3] += q * ch * cs;
    }
    block_sum<4>(v, a.chi + 4 * b);
  } else if (b < nx + ny + np) {
    const int ip = b - nx - ny, ix = ip / ny, iy = ip - ix * ny;
    const double kx = a.kxv[ix], ky = a.kyv[iy], kp = a.kpv[ip];
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < a.n; i += kBlock) {
      double px, py, pz;
      elc_local(a.axis, a.x[i], a.y[i], a.z[i], px, py, pz);
      const double q = a.q[i];
      const double sh = sinh(kp * pz), ch = cosh(kp * pz);
      const double sx = sin(kx * px), cx = cos(kx * px);
      const double sy = sin(ky * py), cy = cos(ky * py);
      v[0] += q * sh * sx * sy;
      v[1] += q * sh * sx * cy;
      v[2] += q * sh * cx * sy;
      v[3] += q * sh * cx * cy;
      v[4] += q * ch * sx * sy;
      v[5] += q * ch * sx * cy;
      v[6] += q * ch * cx * sy;
      v[7] += q * ch * cx * cy;
    }
    block_sum<8>(v, a.chi + 4 * (nx + ny) + 8 * ip);
  } else {
    double v[3] = {0, 0, 0};
    for (int i = threadIdx.x; i < a.n; i += kBlock) {
      const double q = a.q[i];
      v[0] += q * a.x[i];
      v[1] += q * a.y[i];
      v[2] += q * a.z[i];
    }
    __shared__ double ds[3];
    block_sum<3>(v, ds);
    __syncthreads();
    if (threadIdx.x == 0) {
      // the field of this evaluation from the previous sum, then the new sum,
      // both kept along the slab normal only; the previous sum is kept too,
      // so that a re-evaluation of the same step can start from it
      for (int d = 0; d < 3; ++d) {
        const bool on = d == a.axis;
        const double prev = a.dip[d];
        a.dip[6 + d] = prev;
        a.dip[3 + d] = on ? -prev * a.dip_coef : 0.0;
        a.dip[d] = on ? ds[d] : 0.0;
      }
    }
  }
}

// field(const Vector&) (field.cpp:29-136) at the global point (gx, gy, gz),
// without the dipole field.
__device__ void elc_field_at(const ElcArgs &a, double gx, double gy, double gz, double &fx,
                             double &fy, double &fz)
{
  const int nx = a.kx, ny = a.ky;
  const double *cxs = a.chi, *cys = a.chi + 4 * nx, *cps = a.chi + 4 * (nx + ny);
  double px, py, pz;
  elc_local(a.axis, gx, gy, gz, px, py, pz);
  double f0 = 0, f1 = 0, f2 = 0;
  int ip = 0;
  for (int ix = 0; ix < nx; ++ix) {
    const double kx = a.kxv[ix];
    const double shx = sinh(kx * pz), chx = cosh(kx * pz);
    const double sx = sin(kx * px), cx = cos(kx * px);
    const double *c = cxs + 4 * ix;
    const double sx_chi = -kx * (-(c[0] * cx * shx) - (-c[1] * sx * shx) + (c[2] * cx * chx) + (-c[3] * sx * chx));
    const double sz_chi = -kx * (-(c[0] * chx * sx) - (c[1] * chx * cx) + (c[2] * shx * sx) + (c[3] * shx * cx));
    f0 += a.kxc[ix] * sx_chi;
    f2 += a.kxc[ix] * sz_chi;
    for (int iy = 0; iy < ny; ++iy, ++ip) {
      const double ky = a.kyv[iy], kp = a.kpv[ip];
      const double shp = sinh(kp * pz), chp = cosh(kp * pz);
      const double sy = sin(ky * py), cy = cos(ky * py);
      const double *q = cps + 8 * ip;
      const double tx = -kx * (-(q[0] * cx * shp * sy) - (q[1] * cx * shp * cy) - (-q[2] * sx * shp * sy) +
                               (q[4] * cx * chp * sy) - (-q[3] * sx * shp * cy) + (q[5] * cx * chp * cy) +
                               (-q[6] * sx * chp * sy) + (-q[7] * sx * chp * cy));
      const double ty = -ky * (-(q[0] * cy * shp * sx) - (-q[1] * sy * shp * sx) - (q[2] * cy * shp * cx) +
                               (q[4] * cy * chp * sx) - (-q[3] * sy * shp * cx) + (-q[5] * sy * chp * sx) +
                               (q[6] * cy * chp * cx) + (-q[7] * sy * chp * cx));
      const double tz = -kp * (-(q[0] * chp * sx * sy) - (q[1] * chp * sx * cy) - (q[2] * chp * cx * sy) +
                               (q[4] * shp * sx * sy) - (q[3] * chp * cx * cy) + (q[5] * shp * sx * cy) +
                               (q[6] * shp * cx * sy) + (q[7] * shp * cx * cy));
      f0 += 2.0 * a.kpc[ip] * tx;
      f1 += 2.0 * a.kpc[ip] * ty;
      f2 += 2.0 * a.kpc[ip] * tz;
      if (ix == 0) {
        const double shy = sinh(ky * pz), chy = cosh(ky * pz);
        const double *cy4 = cys + 4 * iy;
        const double syy = -ky * (-(cy4[0] * cy * shy) - (-cy4[1] * sy * shy) + (cy4[2] * cy * chy) +
                                  (-cy4[3] * sy * chy));
        const double syz = -ky * (-(cy4[0] * sy * chy) - (cy4[1] * cy * chy) + (cy4[2] * sy * shy) +
                                  (cy4[3] * cy * shy));
        f1 += a.kyc[iy] * syy;
        f2 += a.kyc[iy] * syz;
      }
    }
  }
  elc_global(a.axis, f0 * a.sc, f1 * a.sc, f2 * a.sc, fx, fy, fz);
}

// calculate_acceleration (electrostatic_ewald_slab_correction.cpp:134-181):
// force = q_i (field(r_i) + dipole field), added as force * mass_inv.
__global__ __launch_bounds__(kBlock) void k_elc_force(ElcArgs a)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.n) return;
  double fx, fy, fz;
  elc_field_at(a, a.x[i], a.y[i], a.z[i], fx, fy, fz);
  const double qi = a.q[i], mi = a.minv[a.type[i]];
  a.ax[i] += (qi * (fx + a.dip[3])) * mi;
  a.ay[i] += (qi * (fy + a.dip[4])) * mi;
  a.az[i] += (qi * (fz + a.dip[5])) * mi;
}

// Point queries with the state of the last evaluation: field (plus the
// dipole field) per point, and the potential (potential.cpp:30-79) as a
// fixed-order sum over the owned atoms, one workgroup per point.
__global__ __launch_bounds__(kBlock) void k_elc_point(ElcArgs a, int np, const double *pts,
                                                      double *out4, double pot_coef)
{
  const int k = blockIdx.x;
  if (k >= np) return;
  const double rx = pts[3 * k], ry = pts[3 * k + 1], rz = pts[3 * k + 2];
  const int nx = a.kx, ny = a.ky;
  double v[1] = {0};
  for (int j = threadIdx.x; j < a.n; j += kBlock) {
    double px, py, pz;
    elc_local(a.axis, rx - a.x[j], ry - a.y[j], rz - a.z[j], px, py, pz);
    double sum_kx = 0, sum_ky = 0, sum_kp = 0;
    int ip = 0;
    for (int ix = 0; ix < nx; ++ix) {
      const double cxz = cosh(a.kxv[ix] * pz), cxx = cos(a.kxv[ix] * px);
      sum_kx += a.kxc[ix] * cxz * cxx;
      for (int iy = 0; iy < ny; ++iy, ++ip) {
        const double cpz = cosh(a.kpv[ip] * pz), cyy = cos(a.kyv[iy] * py);
        sum_kp += a.kpc[ip] * cpz * cxx * cyy;
        if (ix == 0) sum_ky += a.kyc[iy] * cosh(a.kyv[iy] * pz) * cyy;
      }
    }
    v[0] += a.q[j] * ((2.0 * sum_kp) + sum_kx + sum_ky);
  }
  __shared__ double s[1];
  block_sum<1>(v, s);
  __syncthreads();
  if (threadIdx.x == 0) {
    double fx, fy, fz;
    elc_field_at(a, rx, ry, rz, fx, fy, fz);
    out4[4 * k] = fx + a.dip[3];
    out4[4 * k + 1] = fy + a.dip[4];
    out4[4 * k + 2] = fz + a.dip[5];
    out4[4 * k + 3] = s[0] * pot_coef;
  }
}

static ElcArgs elc_args(cav_ctx *c)
{
  ElcArgs a;
  a.n = c->n;
  a.axis = c->elc_axis;
  a.kx = c->elc_kx;
  a.ky = c->elc_ky;
  a.x = c->x;
  a.y = c->y;
  a.z = c->z;
  a.q = c->q;
  a.minv = c->d_minv;
  a.type = c->type;
  const int nx = c->elc_kx, ny = c->elc_ky;
  a.kxv = c->d_elc_k;
  a.kxc = a.kxv + nx;
  a.kyv = a.kxc + nx;
  a.kyc = a.kyv + ny;
  a.kpv = a.kyc + ny;
  a.kpc = a.kpv + nx * ny;
  a.chi = c->d_elc_chi;
  a.dip = c->d_elc_dip;
  a.dip_coef = c->elc_dip_coef;
  a.sc = 2.0 * c->elc_sum_e_coef * c->elc_ke;
  a.ax = c->ax;
  a.ay = c->ay;
  a.az = c->az;
  return a;
}

// initialize (initialize.cpp:31-73) and make_slab_k_vectors
// (initialize.cpp:119-160): box lengths in slab coordinates, k grids and
// coefficients, on the host in the reference's arithmetic; device buffers.
int elc_plan(cav_ctx *c)
{
  const double glx = c->L[0], gly = c->L[1], glz = c->L[2];
  const int ax = c->elc_axis;
  const double lx = ax == 0 ? gly : glx, ly = ax == 2 ? gly : glz, hz = ax == 0 ? glx : (ax == 1 ? gly : glz);
  const double lz = c->elc_lz;
  if (hz > lz) return fail(c, CAV_ERR_ARG, "expected 'hz > lz' for slab geometry");
  const double lx_ly_inv = 1.0 / (lx * ly);
  c->elc_sum_e_coef = -CAV_4PI * lx_ly_inv;
  c->elc_dip_coef = c->elc_ke * CAV_4PI * (lx_ly_inv / lz);
  const int nx = c->elc_kx, ny = c->elc_ky, np = nx * ny;
  std::vector<double> k(2 * (size_t)(nx + ny + np) + 1, 0.0);
  double *kxv = k.data(), *kxc = kxv + nx, *kyv = kxc + nx, *kyc = kyv + ny, *kpv = kyc + ny,
         *kpc = kpv + np;
  const double lx_inv = 1.0 / lx, ly_inv = 1.0 / ly;
  for (int ix = 0; ix < nx; ++ix) {
    kxv[ix] = CAV_2PI * (ix + 1) * lx_inv;
    kxc[ix] = 1.0 / (kxv[ix] * (std::exp(kxv[ix] * lz) - 1.0));
  }
  for (int iy = 0; iy < ny; ++iy) {
    kyv[iy] = CAV_2PI * (iy + 1) * ly_inv;
    kyc[iy] = 1.0 / (kyv[iy] * (std::exp(kyv[iy] * lz) - 1.0));
  }
  for (int ix = 0, ip = 0; ix < nx; ++ix)
    for (int iy = 0; iy < ny; ++iy, ++ip) {
      kpv[ip] = std::sqrt((kxv[ix] * kxv[ix]) + (kyv[iy] * kyv[iy]));
      kpc[ip] = 1.0 / (kpv[ip] * (std::exp(kpv[ip] * lz) - 1.0));
    }
  for (double **p : {&c->d_elc_k, &c->d_elc_chi, &c->d_elc_dip})
    if (*p) {
      (void)hipFree(*p);
      *p = nullptr;
    }
  CAV_HIP(hipMalloc(&c->d_elc_k, sizeof(double) * k.size()));
  CAV_HIP(hipMalloc(&c->d_elc_chi, sizeof(double) * (4 * (size_t)(nx + ny) + 8 * (size_t)np + 1)));
  CAV_HIP(hipMalloc(&c->d_elc_dip, sizeof(double) * 12));
  CAV_HIP(hipMemcpyAsync(c->d_elc_k, k.data(), sizeof(double) * k.size(), hipMemcpyHostToDevice, c->stream));
  // Vector<double>() starts at zero: no dipole sum before the first evaluation
  std::vector<double> zero(12, 0.0);
  CAV_HIP(hipMemcpyAsync(c->d_elc_dip, zero.data(), sizeof(double) * 12, hipMemcpyHostToDevice, c->stream));
  return CAV_OK;
}

int launch_elc(cav_ctx *c)
{
  if (c->n == 0) return CAV_OK;
  const ElcArgs a = elc_args(c);
  const int blocks = c->elc_kx + c->elc_ky + c->elc_kx * c->elc_ky + 1;
  k_elc_chi<<<blocks, kBlock, 0, c->stream>>>(a);
  k_elc_force<<<grid_for(c->n), kBlock, 0, c->stream>>>(a);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

__global__ void k_elc_rewind(double *dip)
{
  if (threadIdx.x < 3) dip[threadIdx.x] = dip[6 + threadIdx.x];
}

// Before a re-evaluation of the current step that the reference would not
// make (after autotuning or a tile change): the dipole sum goes back to the
// one the last evaluation started from, so the re-evaluation gives the same
// field and leaves the same state.
int elc_rewind(cav_ctx *c)
{
  if (!c->elc_on || !c->d_elc_dip) return CAV_OK;
  k_elc_rewind<<<1, 64, 0, c->stream>>>(c->d_elc_dip);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// energy (energy.cpp:32-138) from the chi sums of the last evaluation, on the
// host in the reference's order, plus 0.5 dipole_coef |dipole_sum|^2.
int energy_elc(cav_ctx *c, double *e)
{
  const int nx = c->elc_kx, ny = c->elc_ky, np = nx * ny;
  std::vector<double> chi(4 * (size_t)(nx + ny) + 8 * (size_t)np + 1), dip(6);
  std::vector<double> k(2 * (size_t)(nx + ny + np) + 1);
  CAV_HIP(hipMemcpyAsync(chi.data(), c->d_elc_chi, sizeof(double) * chi.size(), hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipMemcpyAsync(dip.data(), c->d_elc_dip, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipMemcpyAsync(k.data(), c->d_elc_k, sizeof(double) * k.size(), hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  const double *kxc = k.data() + nx, *kyc = kxc + nx + ny, *kpc = kyc + ny + np;
  const double *cx = chi.data(), *cy = cx + 4 * nx, *cp = cy + 4 * ny;
  double s = 0.0;
  for (int ip = 0; ip < np; ++ip) {
    const double *q = cp + 8 * ip;
    const double sum_chi = -(q[0] * q[0]) - (q[1] * q[1]) - (q[2] * q[2]) + (q[4] * q[4]) - (q[3] * q[3]) +
                           (q[5] * q[5]) + (q[6] * q[6]) + (q[7] * q[7]);
    s += 2.0 * kpc[ip] * sum_chi;
  }
  for (int ix = 0; ix < nx; ++ix) {
    const double *q = cx + 4 * ix;
    s += kxc[ix] * (-(q[0] * q[0]) - (q[1] * q[1]) + (q[2] * q[2]) + (q[3] * q[3]));
  }
  for (int iy = 0; iy < ny; ++iy) {
    const double *q = cy + 4 * iy;
    s += kyc[iy] * (-(q[0] * q[0]) - (q[1] * q[1]) + (q[2] * q[2]) + (q[3] * q[3]));
  }
  s *= c->elc_ke * c->elc_sum_e_coef;
  *e = s + 0.5 * c->elc_dip_coef * ((dip[0] * dip[0] + dip[1] * dip[1]) + dip[2] * dip[2]);
  return CAV_OK;
}

int point_field_elc(cav_ctx *c, int np, const double *pts, double *field, double *pot)
{
  if (np == 0) return CAV_OK;
  double *d = nullptr;
  CAV_HIP(hipMalloc(&d, sizeof(double) * 7 * (size_t)np));
  CAV_HIP(hipMemcpyAsync(d, pts, sizeof(double) * 3 * np, hipMemcpyHostToDevice, c->stream));
  const ElcArgs a = elc_args(c);
  k_elc_point<<<np, kBlock, 0, c->stream>>>(a, np, d, d + 3 * (size_t)np, c->elc_sum_e_coef * c->elc_ke);
  std::vector<double> out(4 * (size_t)np);
  CAV_HIP(hipGetLastError());
  CAV_HIP(hipMemcpyAsync(out.data(), d + 3 * (size_t)np, sizeof(double) * out.size(), hipMemcpyDeviceToHost,
                         c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  CAV_HIP(hipFree(d));
  for (int k = 0; k < np; ++k) {
    if (field)
      for (int e = 0; e < 3; ++e) field[3 * k + e] = out[4 * k + e];
    if (pot) pot[k] = out[4 * k + 3];
  }
  return CAV_OK;
}

}  // namespace cav
