// Fixed-order reductions (energies, virial, kinetic energy, displacement) and the standalone NVT integrator of the engine:
// velocity Verlet with a Nose-Hoover chain thermostat, the scheme of LAMMPS fix nvt (FixNH::nhc_temp_integrate, tchain 3, tloop 1),
// in LAMMPS "real" units. Inside LAMMPS the plugin never integrates; this is the MD loop of the standalone engine and of bench.py.
#include "common.hpp"
#include "launch.hpp"

namespace rxa {
namespace {

constexpr int kRT = 256;
constexpr int kRBlocks = 240;
constexpr int kRed = kTerms + 6;

__device__ inline void block_sum(double* v, int K, double* sh) {
  const int tid = threadIdx.x;
  for (int k = 0; k < K; ++k) sh[k * kRT + tid] = v[k];
  __syncthreads();
  for (int s = kRT / 2; s > 0; s >>= 1) {
    if (tid < s)
      for (int k = 0; k < K; ++k) sh[k * kRT + tid] += sh[k * kRT + tid + s];
    __syncthreads();
  }
  for (int k = 0; k < K; ++k) v[k] = sh[k * kRT];
  __syncthreads();
}

__global__ void __launch_bounds__(kRT) k_energy_partial(State st, double* part) {
  __shared__ double sh[kRed * kRT];
  double v[kRed];
  for (int k = 0; k < kRed; ++k) v[k] = 0;
  const int n = st.n;
  for (int i = blockIdx.x * kRT + threadIdx.x; i < n; i += gridDim.x * kRT) {
    for (int k = 0; k < kTerms; ++k) v[k] += st.eatom[k * n + i];
    for (int k = 0; k < 6; ++k) v[kTerms + k] += st.vatom[k * n + i];
  }
  block_sum(v, kRed, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < kRed; ++k) part[blockIdx.x * kRed + k] = v[k];
}

__global__ void __launch_bounds__(kRT) k_energy_final(const double* part, double* out) {
  __shared__ double sh[kRed * kRT];
  double v[kRed];
  for (int k = 0; k < kRed; ++k) v[k] = threadIdx.x < kRBlocks ? part[threadIdx.x * kRed + k] : 0.0;
  block_sum(v, kRed, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < kRed; ++k) out[k] = v[k];
}

// ---- integrator -----------------------------------------------------------------------------------------------
// nh layout: [0..2] eta, [3..6] eta_dot (last stays 0), [7..9] eta_dotdot, [10..12] eta_mass, [13] t_target, [14] t_freq,
// [15] dt, [16] dof, [17] boltz, [18] mvv2e, [19] last scale factor, [20] current temperature, [21] kinetic energy, [22] 1 = NVE
__global__ void k_half_kick(State st, MdBuffers md, double dtf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  const double f = dtf / md.mass[i];
  md.vx[i] -= f * st.gx[i];   // force = -gradient
  md.vy[i] -= f * st.gy[i];
  md.vz[i] -= f * st.gz[i];
}

__global__ void k_drift(State st, MdBuffers md, double dt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  const double x = st.x[i] + dt * md.vx[i], y = st.y[i] + dt * md.vy[i], z = st.z[i] + dt * md.vz[i];
  st.x[i] = x;
  st.y[i] = y;
  st.z[i] = z;
  st.pos[i] = make_double4(x, y, z, pos_w(st, i));
}

// one Nose-Hoover chain half step (LAMMPS FixNH::nhc_temp_integrate, mtchain 3, nc_tchain 1, no drag); writes the velocity factor
__global__ void __launch_bounds__(kRT) k_nhc(MdBuffers md) {
  __shared__ double sh[kRT];
  double v = threadIdx.x < kRBlocks ? md.partial[threadIdx.x] : 0.0;
  block_sum(&v, 1, sh);
  if (threadIdx.x != 0) return;
  double* nh = md.nh;
  const double t_target = nh[13], t_freq = nh[14], dt = nh[15], dof = nh[16], boltz = nh[17], mvv2e = nh[18];
  const double ke2 = v * mvv2e;   // sum m v^2 in energy units = 2 KE
  double t_current = ke2 / (dof * boltz);
  if (nh[22] != 0.0) {   // NVE: temperature and kinetic energy only, velocities untouched
    nh[19] = 1.0;
    nh[20] = t_current;
    nh[21] = 0.5 * ke2;
    return;
  }
  double* eta = nh;
  double* eta_dot = nh + 3;
  double* eta_dotdot = nh + 7;
  double* eta_mass = nh + 10;
  const int M = 3;
  const double dt4 = 0.25 * dt, dt8 = 0.125 * dt, dthalf = 0.5 * dt;
  const double ke_target = dof * boltz * t_target;
  eta_mass[0] = dof * boltz * t_target / (t_freq * t_freq);
  for (int c = 1; c < M; ++c) eta_mass[c] = boltz * t_target / (t_freq * t_freq);
  double kecurrent = dof * boltz * t_current;
  eta_dotdot[0] = eta_mass[0] > 0 ? (kecurrent - ke_target) / eta_mass[0] : 0.0;
  for (int c = M - 1; c > 0; --c) {
    const double e = exp(-dt8 * eta_dot[c + 1]);
    eta_dot[c] *= e;
    eta_dot[c] += eta_dotdot[c] * dt4;
    eta_dot[c] *= e;
  }
  double e = exp(-dt8 * eta_dot[1]);
  eta_dot[0] *= e;
  eta_dot[0] += eta_dotdot[0] * dt4;
  eta_dot[0] *= e;
  const double factor = exp(-dthalf * eta_dot[0]);
  t_current *= factor * factor;
  kecurrent = dof * boltz * t_current;
  eta_dotdot[0] = eta_mass[0] > 0 ? (kecurrent - ke_target) / eta_mass[0] : 0.0;
  for (int c = 0; c < M; ++c) eta[c] += dthalf * eta_dot[c];
  eta_dot[0] *= e;
  eta_dot[0] += eta_dotdot[0] * dt4;
  eta_dot[0] *= e;
  for (int c = 1; c < M; ++c) {
    e = exp(-dt8 * eta_dot[c + 1]);
    eta_dot[c] *= e;
    eta_dotdot[c] = (eta_mass[c - 1] * eta_dot[c - 1] * eta_dot[c - 1] - boltz * t_target) / eta_mass[c];
    eta_dot[c] += eta_dotdot[c] * dt4;
    eta_dot[c] *= e;
  }
  nh[19] = factor;
  nh[20] = t_current;
  nh[21] = 0.5 * kecurrent;
}

__global__ void k_scale_v(State st, MdBuffers md) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  const double f = md.nh[19];
  md.vx[i] *= f;
  md.vy[i] *= f;
  md.vz[i] *= f;
}

__global__ void k_ref_positions(State st, MdBuffers md) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  md.x0[i] = st.x[i];
  md.y0[i] = st.y[i];
  md.z0[i] = st.z[i];
}

// Fused forms of the integrator's element-wise kernels for the step loop (kRBlocks blocks, grid-stride partials):
//   k_scale_ke:   v *= nh[19] (the end-of-step thermostat factor still to apply, if `apply`), then the partials of sum m v^2;
//   k_kick_drift: v *= nh[19] (if `scale`), v -= dtf g / m, x += dt v, packed positions, partials of the largest displacement.
__global__ void __launch_bounds__(kRT) k_scale_ke(State st, MdBuffers md, int apply) {
  __shared__ double sh[kRT];
  const double f = md.nh[19];
  double v = 0;
  for (int i = blockIdx.x * kRT + threadIdx.x; i < st.n; i += gridDim.x * kRT) {
    double vx = md.vx[i], vy = md.vy[i], vz = md.vz[i];
    if (apply) {
      vx *= f; vy *= f; vz *= f;
      md.vx[i] = vx; md.vy[i] = vy; md.vz[i] = vz;
    }
    v += md.mass[i] * (vx * vx + vy * vy + vz * vz);
  }
  block_sum(&v, 1, sh);
  if (threadIdx.x == 0) md.partial[blockIdx.x] = v;
}

__global__ void __launch_bounds__(kRT) k_kick_drift(State st, MdBuffers md, double dtf, double dt, int scale) {
  __shared__ double sh[kRT];
  const double fs = md.nh[19];
  double m = 0;
  for (int i = blockIdx.x * kRT + threadIdx.x; i < st.n; i += gridDim.x * kRT) {
    double vx = md.vx[i], vy = md.vy[i], vz = md.vz[i];
    if (scale) { vx *= fs; vy *= fs; vz *= fs; }
    const double f = dtf / md.mass[i];
    vx -= f * st.gx[i];
    vy -= f * st.gy[i];
    vz -= f * st.gz[i];
    md.vx[i] = vx; md.vy[i] = vy; md.vz[i] = vz;
    const double x = st.x[i] + dt * vx, y = st.y[i] + dt * vy, z = st.z[i] + dt * vz;
    st.x[i] = x;
    st.y[i] = y;
    st.z[i] = z;
    st.pos[i] = make_double4(x, y, z, pos_w(st, i));
    const double dx = x - md.x0[i], dy = y - md.y0[i], dz = z - md.z0[i];
    m = fmax(m, dx * dx + dy * dy + dz * dz);
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = kRT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) md.partial[blockIdx.x] = sh[0];
}

// the gather (gather_atom) and the velocity half step in one pass: v -= dtf g / m, then the partials of sum m v^2
__global__ void __launch_bounds__(kRT) k_gather_kick_ke(State st, MdBuffers md, double dtf) {
  __shared__ double sh[kRT];
  double v = 0;
  for (int i = blockIdx.x * kRT + threadIdx.x; i < st.n; i += gridDim.x * kRT) {
    double gx, gy, gz;
    gather_atom(st, i, gx, gy, gz);
    const double f = dtf / md.mass[i];
    const double vx = md.vx[i] - f * gx, vy = md.vy[i] - f * gy, vz = md.vz[i] - f * gz;
    md.vx[i] = vx; md.vy[i] = vy; md.vz[i] = vz;
    v += md.mass[i] * (vx * vx + vy * vy + vz * vz);
  }
  block_sum(&v, 1, sh);
  if (threadIdx.x == 0) md.partial[blockIdx.x] = v;
}

inline int nblk(int n, int per) { return (n + per - 1) / per; }

}  // namespace

// step start: thermostat half step (the pending end-of-step factor applied first), velocity half step, drift, displacement
void launch_md_step_start(const State& st, MdBuffers& md, double dtf, double dt, bool thermostat, bool pending, hipStream_t s) {
  if (thermostat) {
    hipLaunchKernelGGL(k_scale_ke, dim3(kRBlocks), dim3(kRT), 0, s, st, md, pending ? 1 : 0);
    hipLaunchKernelGGL(k_nhc, dim3(1), dim3(kRT), 0, s, md);
  } else if (pending) {
    hipLaunchKernelGGL(k_scale_v, dim3(nblk(st.n, 256)), dim3(256), 0, s, st, md);
  }
  hipLaunchKernelGGL(k_kick_drift, dim3(kRBlocks), dim3(kRT), 0, s, st, md, dtf, dt, thermostat ? 1 : 0);
}
int md_disp_partials() { return kRBlocks; }

// step end fused with the force gather: gather + velocity half step + kinetic-energy partials, then (if `thermostat`) the
// thermostat half step, whose velocity factor stays pending
void launch_md_gather_step_end(const State& st, MdBuffers& md, double dtf, bool thermostat, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_kick_ke, dim3(kRBlocks), dim3(kRT), 0, s, st, md, dtf);
  if (thermostat) hipLaunchKernelGGL(k_nhc, dim3(1), dim3(kRT), 0, s, md);
}
void launch_md_scale(const State& st, MdBuffers& md, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_v, dim3(nblk(st.n, 256)), dim3(256), 0, s, st, md);
}

void launch_energy_reduce(const State& st, double* out, hipStream_t s) {
  // out[kTerms + 6] receives the sums; out + kRed .. holds kRBlocks * kRed partials (scratch)
  double* part = out + kRed;
  hipLaunchKernelGGL(k_energy_partial, dim3(kRBlocks), dim3(kRT), 0, s, st, part);
  hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kRT), 0, s, part, out);
}

void launch_md_half_kick(const State& st, MdBuffers& md, double dtf, hipStream_t s) {
  hipLaunchKernelGGL(k_half_kick, dim3(nblk(st.n, 256)), dim3(256), 0, s, st, md, dtf);
}
void launch_md_drift(const State& st, MdBuffers& md, double dt, hipStream_t s) {
  hipLaunchKernelGGL(k_drift, dim3(nblk(st.n, 256)), dim3(256), 0, s, st, md, dt);
}
void launch_md_ref_positions(const State& st, MdBuffers& md, hipStream_t s) {
  hipLaunchKernelGGL(k_ref_positions, dim3(nblk(st.n, 256)), dim3(256), 0, s, st, md);
}

// plain streaming copy (16-byte elements, grid-stride): a kernel of known traffic for calibrating memory counters
__global__ void k_stream_copy(const double2* __restrict__ src, double2* __restrict__ dst, size_t m) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (size_t)gridDim.x * blockDim.x) dst[k] = src[k];
}

void launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const size_t m = bytes / 16;
  hipLaunchKernelGGL(k_stream_copy, dim3(4096), dim3(256), 0, s, (const double2*)src, (double2*)dst, m);
}

}  // namespace rxa
