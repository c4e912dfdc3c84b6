// Kinetic energy, temperature and pressure of the owned atoms, and the
// velocity-rescaling thermostats, as the reference's Atom_data and constraint
// objects compute them at the end of a velocity-Verlet step
// (md_simulator.cpp:619-633):
//   finalize_temperature (kinetic_energy.cpp:162-202): T = 2 KE / (k_b dof),
//     KE = 1/2 sum m v'^2 with v' = v (n_r_df 0) or v - v_cm (1..3), dof =
//     3n - n_r_df.  For 4..6 the reference forms the inertia tensor I about
//     the mass centre and tests matrix_inverse_3d(I, I^-1) != 0
//     (kinetic_energy.cpp:92-105), whose 0 is success: an invertible I leaves
//     KE = 0 (its warning branch), an exactly singular one a zero I^-1, so
//     only v_cm is removed; the same holds here;
//   constraint berendsen (constraint/berendsen.cpp:120-146) and nve
//     (constraint/nve.cpp:134-178): v *= lambda every `step` steps;
//   finalize_pressure (kinetic_energy.cpp:210-292): P = n k_b T / V +
//     virial / (3 V).
// Sums are per-block partials added in a fixed order; the scalar work runs in
// one device thread, so a step stays free of host round trips.
#include <cmath>

#include "cav_internal.hpp"

namespace cav {

// t[] slots of the context's thermo buffer
enum ThermoSlot {
  T_MV = 0,     // sum m v (3) and sum m (t[3])
  T_VCM = 4,    // centre-of-mass velocity (3)
  T_MV2 = 7,    // sum m (v - v_cm)^2
  T_KE = 8,     // kinetic energy
  T_TEMP = 9,   // last finalized temperature
  T_LAMBDA = 10,
  T_PRESS = 11, // last finalized pressure
  T_MR = 12,    // sum m r (3)
  T_PCM = 15,   // mass centre (3)
  T_INERTIA = 18,  // inertia tensor about it: xx yy zz xy yz zx (6)
  T_KEEP = 24,  // n_r_df 4..6: 1 when the sum is kept (I exactly singular)
  T_HEAD = 32   // per-block partials follow
};
constexpr int kThermoPartWidth = 6;  // widest per-block partial (the inertia tensor)

__global__ void k_mv_part(int n, const double *vx, const double *vy, const double *vz,
                          const int *type, const double *mass, double *part)
{
  __shared__ double red[4][kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[4] = {0, 0, 0, 0};
  if (i < n) {
    const double m = mass[type[i]];
    v[0] = vx[i] * m;
    v[1] = vy[i] * m;
    v[2] = vz[i] * m;
    v[3] = m;
  }
  for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 4; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// sum m r (owned_position_cm_mpi_domain, kinetic_energy.cpp:391-411)
__global__ void k_mr_part(int n, const double *x, const double *y, const double *z,
                          const int *type, const double *mass, double *part)
{
  __shared__ double red[3][kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[3] = {0, 0, 0};
  if (i < n) {
    const double m = mass[type[i]];
    v[0] = x[i] * m;
    v[1] = y[i] * m;
    v[2] = z[i] * m;
  }
  for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// The inertia tensor about the mass centre (owned_inertia_tensor_cm_mpi_domain,
// kinetic_energy.cpp:611-640), p = r - p_cm: xx yy zz xy yz zx.
__global__ void k_inertia_part(int n, const double *x, const double *y, const double *z,
                               const int *type, const double *mass, const double *t, double *part)
{
  __shared__ double red[kThermoPartWidth][kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double v[kThermoPartWidth] = {0, 0, 0, 0, 0, 0};
  if (i < n) {
    const double m = mass[type[i]];
    const double px = x[i] - t[T_PCM], py = y[i] - t[T_PCM + 1], pz = z[i] - t[T_PCM + 2];
    v[0] = m * (py * py + pz * pz);
    v[1] = m * (px * px + pz * pz);
    v[2] = m * (px * px + py * py);
    v[3] = -(m * (px * py));
    v[4] = -(m * (py * pz));
    v[5] = -(m * (pz * px));
  }
  for (int k = 0; k < kThermoPartWidth; ++k) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < kThermoPartWidth; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < kThermoPartWidth)
    part[kThermoPartWidth * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// remove: 0 nothing, 1 the centre-of-mass velocity, 2 the same where t[T_KEEP]
// is set and an empty sum otherwise (n_r_df 4..6)
__global__ void k_mv2_part(int n, const double *vx, const double *vy, const double *vz,
                           const int *type, const double *mass, const double *t, int remove,
                           double *part)
{
  __shared__ double red[kBlock];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  double e = 0;
  if (i < n && (remove < 2 || t[T_KEEP] != 0.0)) {
    const double m = mass[type[i]];
    double a = vx[i], b = vy[i], c = vz[i];
    if (remove >= 1) {
      a -= t[T_VCM];
      b -= t[T_VCM + 1];
      c -= t[T_VCM + 2];
    }
    e = m * ((a * a + b * b) + c * c);
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Fixed-order sum of `width` interleaved columns of nblk partials into t[slot..].
__global__ void k_thermo_sum(int nblk, int width, const double *part, double *t, int slot)
{
  __shared__ double red[kBlock];
  for (int k = 0; k < width; ++k) {
    double s = 0;
    for (int b = threadIdx.x; b < nblk; b += kBlock) s += part[(long)b * width + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) t[slot + k] = red[0];
    __syncthreads();
  }
}

__global__ void k_vcm(double *t)
{
  for (int d = 0; d < 3; ++d) t[T_VCM + d] = t[T_MV + d] / t[T_MV + 3];
}

__global__ void k_pcm(double *t)
{
  for (int d = 0; d < 3; ++d) t[T_PCM + d] = t[T_MR + d] / t[T_MV + 3];
}

// matrix_inverse_3d's determinant (utility/common_template_functions.h:168-172);
// the reference keeps the sum only when it is exactly zero (kinetic_energy.cpp:92-105).
__global__ void k_inertia_keep(double *t)
{
  const double *q = t + T_INERTIA;
  double A[3][3];
  A[0][0] = q[0];
  A[1][1] = q[1];
  A[2][2] = q[2];
  A[0][1] = A[1][0] = q[3];
  A[1][2] = A[2][1] = q[4];
  A[2][0] = A[0][2] = q[5];
  const double det = (A[0][0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2])) -
                     (A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])) +
                     (A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]));
  t[T_KEEP] = det == 0.0 ? 1.0 : 0.0;
}

// What the scalar thread computes after the sums: the kinetic energy, then
// (mode) the temperature, a Berendsen or NVE lambda, or the pressure.
enum ThermoOp { OP_KE = 0, OP_TEMP = 1, OP_BERENDSEN = 2, OP_NVE = 3, OP_PRESS = 4 };

struct ThermoScalars {
  double k_b, dof, n, volume;
  double ber_t0, ber_ratio, ber_lmin, ber_lmax;  // ratio = dt / coupling
  double nve_e_dof, nve_e_tot, nve_kbt, nve_kb, nve_t;
};

__global__ void k_thermo_scalar(double *t, int op, ThermoScalars p, const double *virial)
{
  if (op == OP_KE) {
    t[T_KE] = t[T_MV2] * 0.5;
  } else if (op == OP_TEMP) {
    t[T_TEMP] = 2.0 * t[T_KE] / (p.k_b * p.dof);
  } else if (op == OP_BERENDSEN) {
    double l = sqrt(1.0 + p.ber_ratio * ((p.ber_t0 / t[T_TEMP]) - 1.0));
    if (l < p.ber_lmin) l = p.ber_lmin;
    if (l > p.ber_lmax) l = p.ber_lmax;
    t[T_LAMBDA] = l;
  } else if (op == OP_NVE) {
    const double e = t[T_KE];
    double l = 1.0;
    if (p.nve_e_dof > 0.0)
      l = sqrt(p.nve_e_dof * p.dof / e);
    else if (p.nve_e_tot > 0.0)
      l = sqrt(p.nve_e_tot / e);
    else if (p.nve_kbt > 0.0)
      l = sqrt(p.dof * p.nve_kbt * 0.5 / e);
    else if (p.nve_t > 0.0)
      l = sqrt(p.dof * p.nve_kb * p.nve_t * 0.5 / e);
    t[T_LAMBDA] = l;
  } else {
    const double p2 = (0.0 + 0 + *virial) / (3.0 * p.volume);
    const double p1 = (p.n * p.k_b * t[T_TEMP]) / p.volume;
    t[T_PRESS] = p1 + p2;
  }
}

__global__ void k_scale_vel(int n, double *vx, double *vy, double *vz, const double *t)
{
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double l = t[T_LAMBDA];
  vx[i] *= l;
  vy[i] *= l;
  vz[i] *= l;
}

// get_n_r_df (atom_data.cpp:322-350) without stochastic forces.
static int n_r_df(const cav_ctx *c)
{
  if (c->n_r_df >= 0) return c->n_r_df;
  const int bc_sum = c->bc[0] + c->bc[1] + c->bc[2];
  return bc_sum == 0 ? 6 : (bc_sum == 3 ? 3 : 6 - bc_sum);
}

static ThermoScalars scalars(const cav_ctx *c)
{
  ThermoScalars p;
  p.k_b = c->k_b;
  p.dof = (double)(3 * c->n - n_r_df(c));
  p.n = (double)c->n;
  p.volume = c->L[0] * c->L[1] * c->L[2];
  p.ber_t0 = c->ber_t0;
  p.ber_ratio = c->ber_dt / c->ber_coupling;
  p.ber_lmin = c->ber_lmin;
  p.ber_lmax = c->ber_lmax;
  p.nve_e_dof = c->nve_e_dof;
  p.nve_e_tot = c->nve_e_tot;
  p.nve_kbt = c->nve_kbt;
  p.nve_kb = c->nve_kb;
  p.nve_t = c->nve_t;
  return p;
}

int thermo_alloc(cav_ctx *c)
{
  if (c->d_thermo) (void)hipFree(c->d_thermo);
  c->d_thermo = nullptr;
  const size_t count = T_HEAD + kThermoPartWidth * (size_t)grid_for(std::max(c->n, 1));
  CAV_HIP(hipMalloc(&c->d_thermo, sizeof(double) * count));
  CAV_HIP(hipMemsetAsync(c->d_thermo, 0, sizeof(double) * count, c->stream));
  return CAV_OK;
}

// kinetic_energy_mpi_domain (kinetic_energy.cpp:30-138) into t[T_KE].
static int kinetic_energy(cav_ctx *c)
{
  double *t = c->d_thermo, *part = c->d_thermo + T_HEAD;
  const int nb = grid_for(c->n);
  const int nr = n_r_df(c);
  const int remove = nr == 0 ? 0 : (nr <= 3 ? 1 : 2);
  if (remove >= 1) {
    k_mv_part<<<nb, kBlock, 0, c->stream>>>(c->n, c->vx, c->vy, c->vz, c->type, c->d_mass, part);
    k_thermo_sum<<<1, kBlock, 0, c->stream>>>(nb, 4, part, t, T_MV);
    k_vcm<<<1, 1, 0, c->stream>>>(t);
  }
  if (remove == 2) {
    k_mr_part<<<nb, kBlock, 0, c->stream>>>(c->n, c->x, c->y, c->z, c->type, c->d_mass, part);
    k_thermo_sum<<<1, kBlock, 0, c->stream>>>(nb, 3, part, t, T_MR);
    k_pcm<<<1, 1, 0, c->stream>>>(t);
    k_inertia_part<<<nb, kBlock, 0, c->stream>>>(c->n, c->x, c->y, c->z, c->type, c->d_mass, t,
                                                part);
    k_thermo_sum<<<1, kBlock, 0, c->stream>>>(nb, kThermoPartWidth, part, t, T_INERTIA);
    k_inertia_keep<<<1, 1, 0, c->stream>>>(t);
  }
  k_mv2_part<<<nb, kBlock, 0, c->stream>>>(c->n, c->vx, c->vy, c->vz, c->type, c->d_mass, t,
                                          remove, part);
  k_thermo_sum<<<1, kBlock, 0, c->stream>>>(nb, 1, part, t, T_MV2);
  k_thermo_scalar<<<1, 1, 0, c->stream>>>(t, OP_KE, scalars(c), c->scal);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// finalize_temperature: only with temperature_process (kinetic_energy.cpp:165-166).
int thermo_temperature(cav_ctx *c)
{
  if (!c->temperature_process || c->n == 0) return CAV_OK;
  if (c->k_b < 0) return fail(c, CAV_ERR_STATE, "k_b is not set.");
  CAV_TRY(kinetic_energy(c));
  k_thermo_scalar<<<1, 1, 0, c->stream>>>(c->d_thermo, OP_TEMP, scalars(c), c->scal);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// The end of a step at Md_simulator::current_step `ts` (md_simulator.cpp:619-633).
int thermo_end_of_step(cav_ctx *c, long ts)
{
  if (!c->thermo_on || c->n == 0) return CAV_OK;
  CAV_TRY(thermo_temperature(c));
  bool scaled = false;
  const int nb = grid_for(c->n);
  if (c->ber_on && ts % c->ber_step == 0) {
    k_thermo_scalar<<<1, 1, 0, c->stream>>>(c->d_thermo, OP_BERENDSEN, scalars(c), c->scal);
    k_scale_vel<<<nb, kBlock, 0, c->stream>>>(c->n, c->vx, c->vy, c->vz, c->d_thermo);
    scaled = true;
  }
  if (c->nve_on && ts % c->nve_step == 0) {
    CAV_TRY(kinetic_energy(c));
    k_thermo_scalar<<<1, 1, 0, c->stream>>>(c->d_thermo, OP_NVE, scalars(c), c->scal);
    k_scale_vel<<<nb, kBlock, 0, c->stream>>>(c->n, c->vx, c->vy, c->vz, c->d_thermo);
    scaled = true;
  }
  if (scaled) CAV_TRY(thermo_temperature(c));
  if (c->pressure_process)
    k_thermo_scalar<<<1, 1, 0, c->stream>>>(c->d_thermo, OP_PRESS, scalars(c), c->scal);
  CAV_HIP(hipGetLastError());
  return CAV_OK;
}

// The kinetic energy now, and the temperature and pressure last finalized.
int thermo_read(cav_ctx *c, double *ke, double *temp, double *press)
{
  if (ke) CAV_TRY(kinetic_energy(c));
  double t[T_HEAD];
  CAV_HIP(hipMemcpyAsync(t, c->d_thermo, sizeof(t), hipMemcpyDeviceToHost, c->stream));
  CAV_HIP(hipStreamSynchronize(c->stream));
  if (ke) *ke = t[T_KE];
  if (temp) *temp = t[T_TEMP];
  if (press) *press = t[T_PRESS];
  return CAV_OK;
}

}  // namespace cav
