// Charge equilibration on the device: the stock fix qeq/reaxff solve (LAMMPS fix_qeq_reaxff.cpp: init_matvec history extrapolation,
// CG, calculate_Q; the reference keeps the CG line for line in plugin/adapter/fix_qeq_reaxff_metal.cpp:90-123), with both systems
// (s: H s = -chi, t: H t = -1) advanced together so every pass over H serves two vectors. Each system keeps its own scalars and
// stops exactly where the stock loop would (i < imax && sqrt(r.M^-1 r)/|b| > tol), so iteration counts are those of two separate
// stock solves. Vectors are interleaved per atom ([2 i + system]).
//
// Determinism: all dot products are reduced in a fixed order (per-thread grid-stride partials, a fixed LDS tree per block, and the
// per-block partials reduced again in the same fixed order by every block that needs the scalar), so every block computes the same
// alpha/beta bitwise and results repeat run to run. Scalars of iteration `it` live in slot `it` of device arrays, so the blocks of one
// launch never read a value another block of the same launch writes.
#include "common.hpp"
#include "launch.hpp"

namespace rxa {
namespace {

constexpr int kThreads = 256;
constexpr int kMvRowsPerWave = 2;
constexpr int kMvRowsPerBlock = kMvRowsPerWave * (kThreads / 64);
inline __host__ __device__ int mv_blocks(int n) { return (n + kMvRowsPerBlock - 1) / kMvRowsPerBlock; }
// scal layout: [0,1] |b| of s and t, [2,3] u terms, then sig[it][sys] at 8 + 2 it + sys
__device__ inline double& SIG(double* scal, int it, int sys) { return scal[8 + 2 * it + sys]; }

template <int K>
__device__ inline void block_reduce(double (&v)[K], double* sh) {
  const int tid = threadIdx.x;
  for (int k = 0; k < K; ++k) sh[k * kThreads + tid] = v[k];
  __syncthreads();
  for (int stride = kThreads / 2; stride > 0; stride >>= 1) {
    if (tid < stride)
      for (int k = 0; k < K; ++k) sh[k * kThreads + tid] += sh[k * kThreads + tid + stride];
    __syncthreads();
  }
  for (int k = 0; k < K; ++k) v[k] = sh[k * kThreads];
  __syncthreads();
}

// every block re-reduces the `np` partials of a previous launch (K values per block) in the same fixed order: thread t sums
// partials t, t + kThreads, ... in sequence, then the fixed LDS tree
template <int K>
__device__ inline void reduce_partials(const double* part, double (&v)[K], double* sh, int np = kCgBlocks) {
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < np; b += kThreads)
    for (int k = 0; k < K; ++k) v[k] += part[b * K + k];
  block_reduce<K>(v, sh);
}

// guess mode: 0 zero start, 1 the stock history extrapolation, 2 keep x and b as the caller wrote them
__global__ void __launch_bounds__(kThreads) k_cg_guess(int n, QeqBuffers qb, int use_hist) {
  if (use_hist == 2) return;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const double* sh = qb.s_hist;
    const double* th = qb.t_hist;
    double s = 0, t = 0;
    if (use_hist) {
      s = 4 * (sh[0 * n + i] + sh[2 * n + i]) - (6 * sh[1 * n + i] + sh[3 * n + i]);
      t = th[2 * n + i] + 3 * (th[0 * n + i] - th[1 * n + i]);
    }
    qb.x[2 * i] = s;
    qb.x[2 * i + 1] = t;
    qb.b[2 * i] = -qb.chi[i];
    qb.b[2 * i + 1] = -1.0;
  }
}

// w = (diag(eta) + H) v for both systems; optional per-block partial sums of v.w (launches skip entirely once both systems
// stopped). Each wave takes kMvRowsPerWave consecutive rows and walks a row with its 64 lanes, two entries per lane in flight;
// the pair (v_s, v_t) of a column is one 16-byte load. The matrix values are FP64 or FP32 (TV); products and sums are FP64.
template <class TV>
__global__ void __launch_bounds__(kThreads) k_spmv2(int n, int maxnb, const int* __restrict__ cnt, const int* __restrict__ col,
                                                   const TV* __restrict__ val, const double* __restrict__ eta, const double* v,
                                                   double* w, double* part, const int* active) {
  __shared__ double sh[2 * kThreads];
  const int a0 = active ? active[0] : 1, a1 = active ? active[1] : 1;
  if (!a0 && !a1) return;
  const int lane = lane_id();
  const double2* __restrict__ v2 = reinterpret_cast<const double2*>(v);
  double2* w2 = reinterpret_cast<double2*>(w);
  double dot0 = 0, dot1 = 0;
  const int row0 = blockIdx.x * kMvRowsPerBlock + (threadIdx.x >> 6) * kMvRowsPerWave;
  for (int r = 0; r < kMvRowsPerWave; ++r) {
    const int i = row0 + r;
    if (i >= n) break;
    const size_t base = (size_t)i * maxnb;
    const int c = cnt[i];
    double y0 = 0, y1 = 0, z0 = 0, z1 = 0;
    int k = lane;
    for (; k + 64 < c; k += 128) {
      const int ja = col[base + k], jb = col[base + k + 64];
      const double ha = (double)val[base + k], hb = (double)val[base + k + 64];
      const double2 xa = v2[ja], xb = v2[jb];
      y0 += ha * xa.x; y1 += ha * xa.y;
      z0 += hb * xb.x; z1 += hb * xb.y;
    }
    if (k < c) {
      const int ja = col[base + k];
      const double ha = (double)val[base + k];
      const double2 xa = v2[ja];
      y0 += ha * xa.x; y1 += ha * xa.y;
    }
    y0 = wave_sum(y0 + z0);
    y1 = wave_sum(y1 + z1);
    const double2 xi = v2[i];
    const double ei = eta[i];
    y0 += ei * xi.x;
    y1 += ei * xi.y;
    if (lane == 0) {
      w2[i] = make_double2(y0, y1);
      dot0 += xi.x * y0;
      dot1 += xi.y * y1;
    }
  }
  if (part) {
    double vv[2] = {dot0, dot1};
    block_reduce<2>(vv, sh);
    if (threadIdx.x == 0) { part[blockIdx.x * 2] = vv[0]; part[blockIdx.x * 2 + 1] = vv[1]; }
  }
}

// single-vector matvec (the fix's sparse_matvec), owned rows
__global__ void __launch_bounds__(kThreads) k_spmv1(int n, int maxnb, QeqBuffers qb, const double* v, double* w) {
  const int lane = lane_id();
  const int i = (blockIdx.x * kThreads + threadIdx.x) >> 6;
  if (i >= n) return;
  const size_t base = (size_t)i * maxnb;
  double y = 0;
  for (int k = lane; k < qb.cnt[i]; k += 64) y += qb.val[base + k] * v[qb.col[base + k]];
  y = wave_sum(y);
  if (lane == 0) w[i] = y + qb.eta[i] * v[i];
}

__global__ void __launch_bounds__(kThreads) k_cg_init(int n, QeqBuffers qb) {
  __shared__ double sh[4 * kThreads];
  double acc[4] = {0, 0, 0, 0};
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const double hi = qb.hinv[i];
    for (int s = 0; s < 2; ++s) {
      const double r = qb.b[2 * i + s] - qb.q[2 * i + s];
      const double d = r * hi;
      qb.r[2 * i + s] = r;
      qb.d[2 * i + s] = d;
      acc[s] += qb.b[2 * i + s] * qb.b[2 * i + s];
      acc[2 + s] += r * d;
    }
  }
  block_reduce<4>(acc, sh);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) qb.partial[blockIdx.x * 4 + k] = acc[k];
}

__global__ void __launch_bounds__(kThreads) k_cg_init_fin(QeqBuffers qb, int imax, double tol) {
  __shared__ double sh[4 * kThreads];
  double v[4];
  reduce_partials<4>(qb.partial, v, sh);
  if (threadIdx.x == 0) {
    for (int s = 0; s < 2; ++s) {
      const double bn = sqrt(v[s]);
      qb.scal[s] = bn;
      SIG(qb.scal, 0, s) = v[2 + s];
      qb.iscal[s] = (1 < imax && sqrt(v[2 + s]) / bn > tol) ? 1 : 0;
    }
  }
}

// x += alpha d, r -= alpha q, p = r / eta, partial r.p   (iteration it >= 1)
__global__ void __launch_bounds__(kThreads) k_cg_update(int n, QeqBuffers qb, int it) {
  __shared__ double sh[2 * kThreads];
  const int* act = qb.iscal + 2 * (it - 1);
  const int a[2] = {act[0], act[1]};
  if (!a[0] && !a[1]) return;
  double dq[2];
  reduce_partials<2>(qb.partial_mv, dq, sh, mv_blocks(n));
  double alpha[2];
  for (int s = 0; s < 2; ++s) alpha[s] = a[s] ? SIG(qb.scal, it - 1, s) / dq[s] : 0.0;
  double acc[2] = {0, 0};
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const double hi = qb.hinv[i];
    for (int s = 0; s < 2; ++s) {
      if (!a[s]) continue;
      const int k = 2 * i + s;
      qb.x[k] += alpha[s] * qb.d[k];
      qb.r[k] -= alpha[s] * qb.q[k];
      const double p = qb.r[k] * hi;
      qb.p[k] = p;
      acc[s] += qb.r[k] * p;
    }
  }
  block_reduce<2>(acc, sh);
  if (threadIdx.x == 0) { qb.partial[2 * kCgBlocks + blockIdx.x * 2] = acc[0]; qb.partial[2 * kCgBlocks + blockIdx.x * 2 + 1] = acc[1]; }
}

// beta, d = p + beta d, next convergence decision
__global__ void __launch_bounds__(kThreads) k_cg_dir(int n, QeqBuffers qb, int it, int imax, double tol) {
  __shared__ double sh[2 * kThreads];
  const int* act = qb.iscal + 2 * (it - 1);
  const int a[2] = {act[0], act[1]};
  if (!a[0] && !a[1]) {
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int s = 0; s < 2; ++s) { SIG(qb.scal, it, s) = SIG(qb.scal, it - 1, s); qb.iscal[2 * it + s] = 0; }
    return;
  }
  double sn[2];
  reduce_partials<2>(qb.partial + 2 * kCgBlocks, sn, sh);
  double beta[2];
  for (int s = 0; s < 2; ++s) beta[s] = a[s] ? sn[s] / SIG(qb.scal, it - 1, s) : 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
    for (int s = 0; s < 2; ++s) {
      if (!a[s]) continue;
      const int k = 2 * i + s;
      qb.d[k] = qb.p[k] + beta[s] * qb.d[k];
    }
  if (blockIdx.x == 0 && threadIdx.x == 0)
    for (int s = 0; s < 2; ++s) {
      if (a[s]) {
        SIG(qb.scal, it, s) = sn[s];
        qb.iscal[2 * it + s] = (it + 1 < imax && sqrt(sn[s]) / qb.scal[s] > tol) ? 1 : 0;
      } else {
        SIG(qb.scal, it, s) = SIG(qb.scal, it - 1, s);
        qb.iscal[2 * it + s] = 0;
      }
    }
}

__global__ void __launch_bounds__(kThreads) k_cg_sums(int n, QeqBuffers qb) {
  __shared__ double sh[2 * kThreads];
  double acc[2] = {0, 0};
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    acc[0] += qb.x[2 * i];
    acc[1] += qb.x[2 * i + 1];
  }
  block_reduce<2>(acc, sh);
  if (threadIdx.x == 0) { qb.partial[blockIdx.x * 2] = acc[0]; qb.partial[blockIdx.x * 2 + 1] = acc[1]; }
}

// q = s - (sum s / sum t) t, then the history shift of calculate_Q
__global__ void __launch_bounds__(kThreads) k_cg_charges(int n, QeqBuffers qb, double* q) {
  __shared__ double sh[2 * kThreads];
  double v[2];
  reduce_partials<2>(qb.partial, v, sh);
  const double u = v[0] / v[1];
  if (blockIdx.x == 0 && threadIdx.x == 0) { qb.scal[2] = v[0]; qb.scal[3] = v[1]; }
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    const double s = qb.x[2 * i], t = qb.x[2 * i + 1];
    q[i] = s - u * t;
    for (int k = 3; k > 0; --k) {
      qb.s_hist[k * n + i] = qb.s_hist[(k - 1) * n + i];
      qb.t_hist[k * n + i] = qb.t_hist[(k - 1) * n + i];
    }
    qb.s_hist[i] = s;
    qb.t_hist[i] = t;
  }
}

}  // namespace

void launch_qeq_matvec1(const State& st, const QeqBuffers& qb, const double* x, double* y, hipStream_t s) {
  hipLaunchKernelGGL(k_spmv1, dim3((st.n * 64 + kThreads - 1) / kThreads), dim3(kThreads), 0, s, st.n, st.maxnb, qb, x, y);
}

void launch_qeq_init(const State& st, QeqBuffers& qb, int imax, double tol, int use_hist, hipStream_t s) {
  const int n = st.n;
  hipLaunchKernelGGL(k_cg_guess, dim3(kCgBlocks), dim3(kThreads), 0, s, n, qb, use_hist);
  if (qb.fp32)
    hipLaunchKernelGGL(k_spmv2<float>, dim3(mv_blocks(n)), dim3(kThreads), 0, s, n, st.maxnb, qb.cnt, qb.col, qb.val32, qb.eta, qb.x, qb.q,
                       (double*)nullptr, (const int*)nullptr);
  else
    hipLaunchKernelGGL(k_spmv2<double>, dim3(mv_blocks(n)), dim3(kThreads), 0, s, n, st.maxnb, qb.cnt, qb.col, qb.val, qb.eta, qb.x, qb.q,
                       (double*)nullptr, (const int*)nullptr);
  hipLaunchKernelGGL(k_cg_init, dim3(kCgBlocks), dim3(kThreads), 0, s, n, qb);
  hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(kThreads), 0, s, qb, imax, tol);
}

void launch_qeq_spmv(const State& st, QeqBuffers& qb, int it, hipStream_t s) {
  const int n = st.n;
  if (qb.fp32)
    hipLaunchKernelGGL(k_spmv2<float>, dim3(mv_blocks(n)), dim3(kThreads), 0, s, n, st.maxnb, qb.cnt, qb.col, qb.val32, qb.eta, qb.d, qb.q,
                       qb.partial_mv, qb.iscal + 2 * (it - 1));
  else
    hipLaunchKernelGGL(k_spmv2<double>, dim3(mv_blocks(n)), dim3(kThreads), 0, s, n, st.maxnb, qb.cnt, qb.col, qb.val, qb.eta, qb.d, qb.q,
                       qb.partial_mv, qb.iscal + 2 * (it - 1));
}

int qeq_mv_partials(int n) { return mv_blocks(n); }

void launch_qeq_update(const State& st, QeqBuffers& qb, int it, int imax, double tol, hipStream_t s) {
  const int n = st.n;
  hipLaunchKernelGGL(k_cg_update, dim3(kCgBlocks), dim3(kThreads), 0, s, n, qb, it);
  hipLaunchKernelGGL(k_cg_dir, dim3(kCgBlocks), dim3(kThreads), 0, s, n, qb, it, imax, tol);
}

void launch_qeq_finish(const State& st, QeqBuffers& qb, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_sums, dim3(kCgBlocks), dim3(kThreads), 0, s, st.n, qb);
  hipLaunchKernelGGL(k_cg_charges, dim3(kCgBlocks), dim3(kThreads), 0, s, st.n, qb, st.q);
}

}  // namespace rxa
