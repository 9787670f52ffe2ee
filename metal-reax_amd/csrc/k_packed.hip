// Several independent systems packed in one engine: the per-system parts of the neighbour search, the MD integrator and the
// energy sums. Every other kernel is per atom or per list entry and serves any number of systems unchanged: lists only ever
// hold partners of the same system, an atom's system travels in pos.w, and its lattice shifts come from its system's table.
//
// Every per-system sum is taken by one wave over the system's own rows (lane = row offset within the system, then a fixed
// butterfly), and a system's neighbour list is rebuilt only when its own atoms moved: a system's results therefore do not
// depend on the systems packed beside it, nor on where in the pack it sits.
//
// Reference: the reference has no batching (capabilities.cpp:112); the per-system arithmetic is that of the single-system
// kernels (thermostat: k_md.hip after LAMMPS FixNH; neighbour semantics: src/neighbor/neighbor.cpp:180-212 over all images).
#include "common.hpp"
#include "launch.hpp"

namespace rxa {

namespace {

inline int nblk(long long items, int per) { return (int)((items + per - 1) / per); }

#pragma clang fp contract(off)
// lattice shift vectors of every code in [-smx, smx]^3 of every system, accumulated exactly as shift_vec does
__global__ void k_pk_shift_table(const Cell* cells, int S, double4* shv, int smx0, int smx1, int smx2) {
  const int d0 = 2 * smx0 + 1, d1 = 2 * smx1 + 1, d2 = 2 * smx2 + 1, T = d0 * d1 * d2;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)S * T) return;
  const int s = (int)(k / T), r = (int)(k % T);
  const int s0 = r % d0 - smx0, s1 = (r / d0) % d1 - smx1, s2 = r / (d0 * d1) - smx2;
  double sx, sy, sz;
  shift_vec(cells[s], shift_code(s0, s1, s2), sx, sy, sz);
  shv[k] = make_double4(sx, sy, sz, 0.0);
}

// positions of the systems being rebuilt wrapped into their cells
__global__ void k_pk_wrap(State st, const Cell* cells, const int* reb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  const int s = st.sysid[i];
  if (!reb[s]) return;
  const Cell& c = cells[s];
  double r[3] = {st.x[i], st.y[i], st.z[i]};
  double k[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < 3; ++d)
    if (c.per[d]) k[d] = floor(frac_coord(c, d, r[0], r[1], r[2]));
  for (int d = 0; d < 3; ++d)
    if (k[d] != 0.0)
      for (int e = 0; e < 3; ++e) r[e] -= k[d] * c.v[d][e];
  st.x[i] = r[0];
  st.y[i] = r[1];
  st.z[i] = r[2];
  st.pos[i] = make_double4(r[0], r[1], r[2], pos_w(st, i));
}
#pragma clang fp contract(on)

// Skin list of every atom of a system being rebuilt, by direct search over the system's atoms: for each candidate j the nearest
// image (from the fractional coordinate difference, in (-1, 1) after the wrap) and the ext[d] images on either side of it along
// each periodic direction (ext[d] = floor(rs / h_d + 1/2), the most that can reach within rs). Entries in (candidate, image)
// order, compacted with a ballot; the entry format of the single-system lists (shift code | kCanon).
__global__ void __launch_bounds__(256) k_pk_nbr(State st, const Cell* cells, const int* sys_start, const int4* ext, const int* reb,
                                               double rs2) {
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= st.n) return;
  const int s = st.sysid[i];
  if (!reb[s]) return;
  const int lane = lane_id();
  const Cell& c = cells[s];
  const int4 e = ext[s];
  const int lo = sys_start[s], hi = sys_start[s + 1];
  const double4 pi = st.pos[i];
  double fi[3];
  for (int d = 0; d < 3; ++d) fi[d] = c.per[d] ? frac_coord(c, d, pi.x, pi.y, pi.z) : 0.0;
  int count = 0;
  const size_t base = (size_t)i * st.maxnb;
  for (int j0 = lo; j0 < hi; j0 += 64) {
    const int j = j0 + lane;
    const bool valid = j < hi;
    int k0[3] = {0, 0, 0};
    if (valid) {
      const double4 pj = st.pos[j];
      for (int d = 0; d < 3; ++d)
        if (c.per[d]) k0[d] = -(int)rint(frac_coord(c, d, pj.x, pj.y, pj.z) - fi[d]);
    }
    for (int o2 = -e.z; o2 <= e.z; ++o2)
      for (int o1 = -e.y; o1 <= e.y; ++o1)
        for (int o0 = -e.x; o0 <= e.x; ++o0) {
          const int s0 = k0[0] + o0, s1 = k0[1] + o1, s2 = k0[2] + o2;
          const int sc = shift_code(s0, s1, s2);
          // the table holds |s_d| <= smx_d; the wrap keeps the nearest image within one cell, so this only drops images that
          // cannot reach within rs
          const bool in_table = abs(s0) <= st.smx0 && abs(s1) <= st.smx1 && abs(s2) <= st.smx2;
          bool keep = false;
          int ent = 0;
          if (valid && in_table && !(j == i && sc == 0)) {
            ent = canonical_first(i, j, sc, st.tag) ? (sc | kCanon) : sc;
            double ddx, ddy, ddz;
            keep = pair_dvec_list(st, pi, j, ent, ddx, ddy, ddz) <= rs2;
          }
          const unsigned long long m = __ballot(keep);
          if (keep) {
            const int slot = count + lanes_below(m);
            if (slot < st.maxnb) { st.nb_j[base + slot] = j; st.nb_s[base + slot] = ent; }
          }
          count += __popcll(m);
        }
  }
  if (lane == 0) {
    st.nb_cnt[i] = count;
    if (count > st.maxnb) atomicMax(&st.flags[0], count);
  }
}

__global__ void k_pk_ref(State st, MdBuffers md, const int* reb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n || !reb[st.sysid[i]]) return;
  md.x0[i] = st.x[i];
  md.y0[i] = st.y[i];
  md.z0[i] = st.z[i];
}

// ---- one block per system (grid = the number of systems) -----------------------------------------------------------------
// Thread t takes the system's atoms t, t + kSysThreads, ... in order; the block combines its values by a xor butterfly per
// wave and then the waves in order, so a system's sums depend only on its own atoms (not on the systems packed with it).
constexpr int kSysThreads = 256;
__device__ inline double sys_block_sum(double v, double* sh) {   // every thread gets the total
  v = wave_sum(v);
  __syncthreads();
  if (lane_id() == 0) sh[threadIdx.x / 64] = v;
  __syncthreads();
  double s = sh[0];
  for (int w = 1; w < kSysThreads / 64; ++w) s += sh[w];
  return s;
}
__device__ inline double sys_block_max(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if (lane_id() == 0) sh[threadIdx.x / 64] = v;
  __syncthreads();
  double m = sh[0];
  for (int w = 1; w < kSysThreads / 64; ++w) m = fmax(m, sh[w]);
  return m;
}

// largest displacement since the system's last rebuild; reb[s] = 1 where it exceeds half the skin (any[0] |= 1)
__global__ void __launch_bounds__(kSysThreads) k_pk_disp(State st, MdBuffers md, const int* sys_start, int S, double lim2, int* reb,
                                                        int* any, double* disp) {
  __shared__ double sh[kSysThreads / 64];
  const int s = blockIdx.x;
  if (s >= S) return;
  double m = 0;
  for (int i = sys_start[s] + threadIdx.x; i < sys_start[s + 1]; i += kSysThreads) {
    const double dx = st.x[i] - md.x0[i], dy = st.y[i] - md.y0[i], dz = st.z[i] - md.z0[i];
    m = fmax(m, dx * dx + dy * dy + dz * dz);
  }
  m = sys_block_max(m, sh);
  if (threadIdx.x == 0) {
    const int r = m > lim2 ? 1 : 0;
    reb[s] = r;
    disp[s] = sqrt(m);
    if (r) atomicOr(any, 1);
  }
}

// One Nose-Hoover chain half step per system (the single-system k_nhc, state nh + 32 s, same layout), from the system's
// sum of m v^2; NVE systems only record temperature and kinetic energy.
__global__ void __launch_bounds__(kSysThreads) k_pk_nhc(State st, MdBuffers md, const int* sys_start, int S, double* nh_all) {
  __shared__ double sh[kSysThreads / 64];
  const int s = blockIdx.x;
  if (s >= S) return;
  double v = 0;
  for (int i = sys_start[s] + threadIdx.x; i < sys_start[s + 1]; i += kSysThreads)
    v += md.mass[i] * (md.vx[i] * md.vx[i] + md.vy[i] * md.vy[i] + md.vz[i] * md.vz[i]);
  v = sys_block_sum(v, sh);
  if (threadIdx.x != 0) return;
  double* nh = nh_all + 32 * (size_t)s;
  const double t_target = nh[13], t_freq = nh[14], dt = nh[15], dof = nh[16], boltz = nh[17], mvv2e = nh[18];
  const double ke2 = v * mvv2e;
  double t_current = dof > 0 ? ke2 / (dof * boltz) : 0.0;
  if (nh[22] != 0.0) {
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

__global__ void k_pk_scale_v(State st, MdBuffers md, const double* nh_all) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  const double f = nh_all[32 * (size_t)st.sysid[i] + 19];
  md.vx[i] *= f;
  md.vy[i] *= f;
  md.vz[i] *= f;
}

// per-system energy terms and virial: out[s * (kTerms + 6) + k]
__global__ void __launch_bounds__(kSysThreads) k_pk_energy(State st, const int* sys_start, int S, double* out) {
  __shared__ double sh[kSysThreads / 64];
  const int s = blockIdx.x;
  if (s >= S) return;
  const int n = st.n;
  for (int k = 0; k < kTerms + 6; ++k) {
    const double* a = k < kTerms ? st.eatom + (size_t)k * n : st.vatom + (size_t)(k - kTerms) * n;
    double v = 0;
    for (int i = sys_start[s] + threadIdx.x; i < sys_start[s + 1]; i += kSysThreads) v += a[i];
    v = sys_block_sum(v, sh);
    if (threadIdx.x == 0) out[(size_t)s * (kTerms + 6) + k] = v;
  }
}

}  // namespace

void launch_pk_shift_table(const Cell* cells, int S, double4* shv, const int smx[3], hipStream_t s) {
  const long long T = (2LL * smx[0] + 1) * (2LL * smx[1] + 1) * (2LL * smx[2] + 1);
  hipLaunchKernelGGL(k_pk_shift_table, dim3(nblk(S * T, 256)), dim3(256), 0, s, cells, S, shv, smx[0], smx[1], smx[2]);
}
void launch_pk_wrap(const State& st, const Cell* cells, const int* reb, hipStream_t s) {
  hipLaunchKernelGGL(k_pk_wrap, dim3(nblk(st.n, 256)), dim3(256), 0, s, st, cells, reb);
}
void launch_pk_nbr(const State& st, const Cell* cells, const int* sys_start, const int4* ext, const int* reb, double rs, hipStream_t s) {
  hipLaunchKernelGGL(k_pk_nbr, dim3(nblk((long long)st.n * 64, 256)), dim3(256), 0, s, st, cells, sys_start, ext, reb, rs * rs);
}
void launch_pk_ref(const State& st, MdBuffers& md, const int* reb, hipStream_t s) {
  hipLaunchKernelGGL(k_pk_ref, dim3(nblk(st.n, 256)), dim3(256), 0, s, st, md, reb);
}
void launch_pk_disp(const State& st, MdBuffers& md, const int* sys_start, int S, double lim, int* reb, int* any, double* disp,
                    hipStream_t s) {
  hipLaunchKernelGGL(k_pk_disp, dim3(S), dim3(kSysThreads), 0, s, st, md, sys_start, S, lim * lim, reb, any, disp);
}
void launch_pk_thermostat(const State& st, MdBuffers& md, const int* sys_start, int S, double* nh_all, hipStream_t s) {
  hipLaunchKernelGGL(k_pk_nhc, dim3(S), dim3(kSysThreads), 0, s, st, md, sys_start, S, nh_all);
  hipLaunchKernelGGL(k_pk_scale_v, dim3(nblk(st.n, 256)), dim3(256), 0, s, st, md, (const double*)nh_all);
}
void launch_pk_energy(const State& st, const int* sys_start, int S, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pk_energy, dim3(S), dim3(kSysThreads), 0, s, st, sys_start, S, out);
}

}  // namespace rxa
