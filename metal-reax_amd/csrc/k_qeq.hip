// Charge equilibration on the device: the stock fix qeq/reaxff solve (LAMMPS fix_qeq_reaxff.cpp: init_matvec history extrapolation,
// CG, calculate_Q; the reference keeps the CG line for line in plugin/adapter/fix_qeq_reaxff_metal.cpp:90-123), with both systems
// (s: H s = -chi, t: H t = -1) advanced together so every pass over H serves two vectors. Each system keeps its own scalars and
// stops exactly where the stock loop would (i < imax && sqrt(r.M^-1 r)/|b| > tol), so iteration counts are those of two separate
// stock solves. Vectors are interleaved per atom ([2 i + system]).
//
// FP64 iterative refinement (qb.refine: mixed mode, and FP32 mode with the fix's 'resident' keyword; the reference's resident
// solve, src/metal/metal_backend.mm:435-530): the CG itself multiplies with the FP32 matrix values, but its initial residual and
// a verification residual after the iterations stop are formed with the two-float matrix hi + lo (about 48 bits). A system
// whose true residual still fails the stock condition restarts from it (beta 0), and its iterations count on, bounded by imax
// as one budget, so the charges solve the FP64 matrix to the requested tolerance.
//
// Determinism: all dot products are reduced in a fixed order (per-thread grid-stride partials, a fixed per-block order, and the
// per-block partials reduced again in the same fixed order by every block that needs the scalar), so every block computes the same
// alpha/beta bitwise and results repeat run to run. Scalars of iteration `it` live in slot `it` of device arrays, so the blocks of one
// launch never read a value another block of the same launch writes.
#include <type_traits>

#include "common.hpp"
#include "launch.hpp"

namespace rxa {
namespace {

constexpr int kThreads = 256;
constexpr int kUpdBlocks = 240, kUpdThreads = 256;   // CG update kernel
constexpr int kMvThreads = 256;             // product kernels: 4 waves per block,
constexpr int kMvRowsPerBlock = kMvThreads / 16;   // 16 lanes per row (kRowLanes below)
inline __host__ __device__ int mv_blocks(int n) { return (n + kMvRowsPerBlock - 1) / kMvRowsPerBlock; }
// scal layout: [0,1] |b| of s and t, [2,3] u terms, [4] refinement checks whose product ran, [5] restarts (running counts),
// then kScalSlot doubles per iteration index k: sig of the two systems at
// 8 + 4 k + sys, the two activity flags (ints) in 8 + 4 k + 2 and the two restart records of a refinement check (ints: -1 no
// check ran, else the restart flags) in 8 + 4 k + 3, so one copy of the first 8 + 4 k doubles brings them all to the host
__device__ inline double& SIG(double* scal, int k, int sys) { return scal[8 + kScalSlot * k + sys]; }
__device__ inline int* ACT(double* scal, int k) { return reinterpret_cast<int*>(scal + 8 + kScalSlot * k + 2); }
__device__ inline int* RST(double* scal, int k) { return reinterpret_cast<int*>(scal + 8 + kScalSlot * k + 3); }

template <int K, int NT = kThreads>
__device__ inline void block_reduce(double (&v)[K], double* sh) {
  // xor butterfly inside each wave, then every thread sums the per-wave values in wave order: one barrier instead of a
  // log2(NT)-level LDS tree, and the same fixed order in every block
  constexpr int kWaves = NT / 64;
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (lane_id() == 0)
    for (int k = 0; k < K; ++k) sh[k * kWaves + threadIdx.x / 64] = v[k];
  __syncthreads();
  for (int k = 0; k < K; ++k) {
    double s = sh[k * kWaves];
    for (int w = 1; w < kWaves; ++w) s += sh[k * kWaves + w];
    v[k] = s;
  }
  __syncthreads();
}

// every block re-reduces the `np` partials of a previous launch (K values per block) in the same fixed order: thread t sums
// partials t, t + NT, ... in sequence, then block_reduce's fixed order
template <int K, int NT = kThreads>
__device__ inline void reduce_partials(const double* part, double (&v)[K], double* sh, int np = kCgBlocks) {
  for (int k = 0; k < K; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < np; b += NT)
    for (int k = 0; k < K; ++k) v[k] += part[b * K + k];
  block_reduce<K, NT>(v, sh);
}

// guess mode: 0 zero start, 1 the stock history extrapolation, 2 keep x and b as the caller wrote them
// The start is also kept in x0: the refinement's bound measures the correction x - x0 the FP32-valued iterations made.
__global__ void __launch_bounds__(kThreads) k_cg_guess(int n, QeqBuffers qb, int use_hist) {
  if (use_hist == 2) {
    for (int k = blockIdx.x * kThreads + threadIdx.x; k < 2 * n; k += gridDim.x * kThreads) qb.x0[k] = qb.x[k];
    return;
  }
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
    qb.x0[2 * i] = s;
    qb.x0[2 * i + 1] = t;
    qb.b[2 * i] = -qb.chi[i];
    qb.b[2 * i + 1] = -1.0;
  }
}

// Product layout: kRowLanes lanes per row (64 / kRowLanes rows per wave side by side), each lane keeping kInFlight entries of its
// row in flight (independent loads) and striding the row by kRowLanes; the row sum is a xor butterfly inside the lane group.
// Measured on the 24,000-atom water matrix (422 entries per row): 16 lanes x 8 entries is 1.3x faster than one wave per row.
constexpr int kRowLanes = 16;
constexpr int kInFlight = 8;

// Matrix value accessors: FP64 values, FP32 values, or the two-float pair hi + lo (FP64 to about 48 bits) of the refinement's
// residuals. Products and sums are FP64 in every case.
struct ValD {
  const double* v;
  __device__ double operator()(size_t k) const { return v[k]; }
};
struct ValF {
  const float* v;
  __device__ double operator()(size_t k) const { return (double)v[k]; }
};
struct ValDS {
  const float* hi;
  const float* lo;
  __device__ double operator()(size_t k) const { return (double)hi[k] + (double)lo[k]; }
};

// (H v)_row for both systems (v interleaved), off-diagonal part, summed over the row's lane group (all its lanes get the sum);
// rows >= n give zero. Columns 32 or 16 bits (TC).
template <class VA, class TC>
__device__ inline double2 row_product(int row, int n, int maxnb, const int* __restrict__ cnt, const TC* __restrict__ col,
                                      const VA val, const double2* __restrict__ v2) {
  const int li = lane_id() % kRowLanes;
  double y0 = 0, y1 = 0;
  if (row < n) {
    const size_t base = (size_t)row * maxnb;
    const int c = cnt[row];
    int k = li;
    for (; k + (kInFlight - 1) * kRowLanes < c; k += kInFlight * kRowLanes) {
      int j[kInFlight];
      double h[kInFlight];
      double2 x[kInFlight];
#pragma unroll
      for (int u = 0; u < kInFlight; ++u) { j[u] = col[base + k + u * kRowLanes]; h[u] = val(base + k + u * kRowLanes); }
#pragma unroll
      for (int u = 0; u < kInFlight; ++u) x[u] = v2[j[u]];
#pragma unroll
      for (int u = 0; u < kInFlight; ++u) { y0 += h[u] * x[u].x; y1 += h[u] * x[u].y; }
    }
    if (k < c) {   // at most kInFlight - 1 entries left: one more pass with their loads in flight together, same sum order
      int j[kInFlight];
      double h[kInFlight];
      double2 x[kInFlight];
#pragma unroll
      for (int u = 0; u < kInFlight; ++u) {
        const int kk = k + u * kRowLanes;
        j[u] = kk < c ? (int)col[base + kk] : -1;
        h[u] = kk < c ? val(base + kk) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kInFlight; ++u) x[u] = j[u] >= 0 ? v2[j[u]] : make_double2(0.0, 0.0);
#pragma unroll
      for (int u = 0; u < kInFlight; ++u)
        if (j[u] >= 0) { y0 += h[u] * x[u].x; y1 += h[u] * x[u].y; }
    }
  }
  for (int off = kRowLanes / 2; off > 0; off >>= 1) {
    y0 += __shfl_xor(y0, off, 64);
    y1 += __shfl_xor(y1, off, 64);
  }
  return make_double2(y0, y1);
}

// Grid of the product kernels: a multiple of the 8 XCDs. Workgroup b is dispatched to XCD b % 8, so it takes the virtual block
// (b % 8) C + b / 8 (C = grid / 8): each XCD walks one contiguous eighth of the rows, and with atoms in cell order that eighth
// is a compact region whose column gathers stay in that XCD's L2. Virtual blocks past the last row compute nothing.
constexpr int kXcds = 8;
inline __host__ __device__ int mv_grid(int n) { return (mv_blocks(n) + kXcds - 1) / kXcds * kXcds; }
__device__ inline int mv_vblock() { return (int)((blockIdx.x % kXcds) * (gridDim.x / kXcds) + blockIdx.x / kXcds); }
__device__ inline int product_row() { return (int)((mv_vblock() * blockDim.x + threadIdx.x) / kRowLanes); }

// Refinement's bound. The CG starts from the true residual r0 = b - A x0 (two-float matrix A = hi + lo) and iterates with the FP32
// values hi, so its recursive residual is r0 - hi (x - x0) while the true one is b - A x = r0 - A (x - x0): they differ by
// lo (x - x0), whose M^-1-norm is at most max_j |x_j - x0_j| * S with S^2 = sum_i (sum_j |lo_ij|)^2 / eta_i. When the recursion's
// sqrt(r.M^-1 r) plus that bound is within tol |b| (with a margin of 1e-3 for the recursion's own FP64 rounding), the true
// residual passes the stock test and the check's product would change nothing: the check is skipped. k_check_gate leaves per
// block max |x - x0| of both systems and the partial sum of S^2 (only once both systems stopped, like the check itself).
__global__ void __launch_bounds__(kThreads) k_check_gate(int n, QeqBuffers qb, int it) {
  __shared__ double sh[3 * kThreads / 64];
  const int* a = ACT(qb.scal, it - 1);
  if (a[0] != 0 || a[1] != 0) return;   // uniform over the grid
  double m0 = 0.0, m1 = 0.0, s2 = 0.0;
  for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    m0 = fmax(m0, fabs(qb.x[2 * i] - qb.x0[2 * i]));
    m1 = fmax(m1, fabs(qb.x[2 * i + 1] - qb.x0[2 * i + 1]));
    const double l = qb.lorow[i];
    s2 += l * l * qb.hinv[i];
  }
  for (int off = 32; off > 0; off >>= 1) {
    m0 = fmax(m0, __shfl_xor(m0, off, 64));
    m1 = fmax(m1, __shfl_xor(m1, off, 64));
    s2 += __shfl_xor(s2, off, 64);
  }
  constexpr int kW = kThreads / 64;
  if (lane_id() == 0) { sh[threadIdx.x / 64] = m0; sh[kW + threadIdx.x / 64] = m1; sh[2 * kW + threadIdx.x / 64] = s2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double v[3] = {sh[0], sh[kW], sh[2 * kW]};
    for (int w = 1; w < kW; ++w) { v[0] = fmax(v[0], sh[w]); v[1] = fmax(v[1], sh[kW + w]); v[2] += sh[2 * kW + w]; }
    for (int k = 0; k < 3; ++k) qb.partial_gate[blockIdx.x * 3 + k] = v[k];
  }
}
// the check is skipped for both systems (every thread computes the same answer from the gate's partials, in a fixed order)
template <int NT>
__device__ inline bool check_skippable(QeqBuffers qb, int it, double tol, double* sh) {
  double m0 = 0.0, m1 = 0.0, s2 = 0.0;
  for (int b = threadIdx.x; b < kCgBlocks; b += NT) {
    m0 = fmax(m0, qb.partial_gate[3 * b]);
    m1 = fmax(m1, qb.partial_gate[3 * b + 1]);
    s2 += qb.partial_gate[3 * b + 2];
  }
  for (int off = 32; off > 0; off >>= 1) {
    m0 = fmax(m0, __shfl_xor(m0, off, 64));
    m1 = fmax(m1, __shfl_xor(m1, off, 64));
    s2 += __shfl_xor(s2, off, 64);
  }
  constexpr int kW = NT / 64;
  __syncthreads();
  if (lane_id() == 0) { sh[threadIdx.x / 64] = m0; sh[kW + threadIdx.x / 64] = m1; sh[2 * kW + threadIdx.x / 64] = s2; }
  __syncthreads();
  double v[3] = {sh[0], sh[kW], sh[2 * kW]};
  for (int w = 1; w < kW; ++w) { v[0] = fmax(v[0], sh[w]); v[1] = fmax(v[1], sh[kW + w]); v[2] += sh[2 * kW + w]; }
  __syncthreads();
  const double S = sqrt(v[2]);
  bool ok = true;
  for (int s = 0; s < 2; ++s) ok = ok && sqrt(SIG(qb.scal, it - 1, s)) + v[s] * S <= (1.0 - 1e-3) * tol * qb.scal[s];
  return ok;
}

// w = (diag(eta) + H) v for both systems. kMode 0 writes w. kMode 1 (CG start) and 2 (refinement check) fuse the start in: the
// row's owner writes r = b - w and p = r / eta instead of w, and each block leaves the partials of b.b and r.p of its rows
// (1: partial_mv, 2: partial_rs). kMode 2 runs only when neither system is active in iteration `it` (activity slot it - 1): the
// check waits until both systems have stopped and otherwise leaves the iteration's vectors alone.
template <class VA, class TC, int kMode>
__global__ void __launch_bounds__(kMvThreads) k_spmv2(int n, int maxnb, const int* __restrict__ cnt, const TC* __restrict__ col,
                                                     const VA val, const double* __restrict__ eta, const double* v, double* w,
                                                     QeqBuffers qb, int it, double tol) {
  if (kMode == 2) {
    const int* a = ACT(qb.scal, it - 1);
    if (a[0] != 0 || a[1] != 0) return;   // uniform over the grid
    __shared__ double shg[3 * kMvThreads / 64];
    if (check_skippable<kMvThreads>(qb, it, tol, shg)) return;   // uniform: the bound already passes both systems
  }
  const double2* __restrict__ v2 = reinterpret_cast<const double2*>(v);
  double2* w2 = reinterpret_cast<double2*>(w);
  const int i = product_row();
  const double2 y = row_product<VA, TC>(i, n, maxnb, cnt, col, val, v2);
  double acc[4] = {0, 0, 0, 0};
  if (i < n && lane_id() % kRowLanes == 0) {
    const double2 xi = v2[i];
    const double ei = eta[i];
    const double2 wi = make_double2(y.x + ei * xi.x, y.y + ei * xi.y);
    if (kMode == 0) {
      w2[i] = wi;
    } else {
      const double hi = qb.hinv[i];
      const double wv[2] = {wi.x, wi.y};
      for (int s = 0; s < 2; ++s) {
        const double b = qb.b[2 * i + s];
        const double r = b - wv[s];
        const double d = r * hi;
        qb.r[2 * i + s] = r;
        qb.p[2 * i + s] = d;
        acc[s] = b * b;
        acc[2 + s] = r * d;
      }
    }
  }
  if (kMode != 0) {
    __shared__ double sh[4 * kMvThreads / 64];
    block_reduce<4, kMvThreads>(acc, sh);
    if (threadIdx.x == 0 && mv_vblock() < mv_blocks(n)) {
      double* part = kMode == 1 ? qb.partial_mv : qb.partial_rs;
      for (int k = 0; k < 4; ++k) part[mv_vblock() * 4 + k] = acc[k];
    }
  }
}

// One CG iteration's product and direction update (iteration it >= 1), both systems:
//   it >= 2: the sums r.p of iteration it-1 are reduced (every block, same fixed order) to sig; a system stays active while the
//            stock condition holds (it < imax and sqrt(sig)/|b| > tol); beta = sig / sig_prev;
//   then for every row i: w = (diag(eta) + H) p, d_i = p_i + beta d_i, q_i = w_i + beta q_i (= (diag(eta) + H) d_i by
//   linearity, so the direction never has to be complete before the product), and per-block partials of d.q.
// Block 0 records sig (slot it-1) and the activity flags of iteration it for the later kernels and the host.
// chk = 1: a refinement check ran before this launch (after update(it - 1)); if it restarted the solve (its record in slot
// it - 1 is not the marker -1), iteration it starts afresh from the true residual (beta 0, activity from the record), else the
// iteration continues as usual.
template <class VA, class TC>
__global__ void __launch_bounds__(kMvThreads) k_cg_mvdir(int n, int maxnb, const TC* __restrict__ col, const VA val, QeqBuffers qb,
                                                        int it, int imax, double tol, int chk) {
  __shared__ double sh[2 * kMvThreads / 64];
  int act[2];
  double beta[2] = {0.0, 0.0};
  bool fresh = it == 1;
  if (fresh) {
    act[0] = ACT(qb.scal, 0)[0]; act[1] = ACT(qb.scal, 0)[1];
  } else if (chk && RST(qb.scal, it - 1)[0] >= 0) {
    fresh = true;
    act[0] = RST(qb.scal, it - 1)[0]; act[1] = RST(qb.scal, it - 1)[1];
    if (blockIdx.x == 0 && threadIdx.x == 0) { ACT(qb.scal, it - 1)[0] = act[0]; ACT(qb.scal, it - 1)[1] = act[1]; }
  } else {
    const int* pa = ACT(qb.scal, it - 2);
    const int pact[2] = {pa[0], pa[1]};
    if (!pact[0] && !pact[1]) {   // both stopped before: nothing to reduce, the slots of it - 1 carry the last values on
      if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int s = 0; s < 2; ++s) { SIG(qb.scal, it - 1, s) = SIG(qb.scal, it - 2, s); ACT(qb.scal, it - 1)[s] = 0; }
      return;
    }
    double sn[2];
    reduce_partials<2, kMvThreads>(qb.partial + 2 * kCgBlocks, sn, sh, kUpdBlocks);
    for (int s = 0; s < 2; ++s) {
      act[s] = pact[s] && it < imax && sqrt(sn[s]) / qb.scal[s] > tol ? 1 : 0;
      if (pact[s]) beta[s] = sn[s] / SIG(qb.scal, it - 2, s);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (int s = 0; s < 2; ++s) {
        SIG(qb.scal, it - 1, s) = pact[s] ? sn[s] : SIG(qb.scal, it - 2, s);
        ACT(qb.scal, it - 1)[s] = act[s];
      }
  }
  if (!act[0] && !act[1]) return;
  const double2* __restrict__ p2 = reinterpret_cast<const double2*>(qb.p);
  const int i = product_row();
  const double2 y = row_product<VA, TC>(i, n, maxnb, qb.cnt, col, val, p2);
  double dq[2] = {0.0, 0.0};
  if (i < n && lane_id() % kRowLanes == 0) {
    const double2 pi = p2[i];
    const double ei = qb.eta[i];
    const double w[2] = {y.x + ei * pi.x, y.y + ei * pi.y}, pv[2] = {pi.x, pi.y};
    for (int s = 0; s < 2; ++s) {
      if (!act[s]) continue;
      const int k = 2 * i + s;
      if (fresh && it > 1) qb.x0[k] = qb.x[k];   // a restart: the correction is measured from here
      const double d = fresh ? pv[s] : pv[s] + beta[s] * qb.d[k];
      const double q = fresh ? w[s] : w[s] + beta[s] * qb.q[k];
      qb.d[k] = d;
      qb.q[k] = q;
      dq[s] = d * q;
    }
  }
  block_reduce<2, kMvThreads>(dq, sh);
  if (threadIdx.x == 0 && mv_vblock() < mv_blocks(n)) { qb.partial_mv[mv_vblock() * 2] = dq[0]; qb.partial_mv[mv_vblock() * 2 + 1] = dq[1]; }
}

// single-vector matvec (the fix's sparse_matvec), owned rows
template <class VA, class TC>
__global__ void __launch_bounds__(kThreads) k_spmv1(int n, int maxnb, const int* cnt, const TC* col, const VA val, const double* eta,
                                                   const double* v, double* w) {
  const int lane = lane_id();
  const int i = (blockIdx.x * kThreads + threadIdx.x) >> 6;
  if (i >= n) return;
  const size_t base = (size_t)i * maxnb;
  double y = 0;
  for (int k = lane; k < cnt[i]; k += 64) y += val(base + k) * v[col[base + k]];
  y = wave_sum(y);
  if (lane == 0) w[i] = y + eta[i] * v[i];
}

template <int NT>
__global__ void __launch_bounds__(NT) k_cg_init_fin(QeqBuffers qb, const double* part, int nparts, int imax, double tol) {
  __shared__ double sh[4 * NT / 64];
  double v[4];
  reduce_partials<4, NT>(part, v, sh, nparts);
  if (threadIdx.x == 0) {
    for (int s = 0; s < 2; ++s) {
      const double bn = sqrt(v[s]);
      qb.scal[s] = bn;
      SIG(qb.scal, 0, s) = v[2 + s];
      ACT(qb.scal, 0)[s] = (1 < imax && sqrt(v[2 + s]) / bn > tol) ? 1 : 0;
    }
  }
}

// End of a refinement check (after update(it)): if either system was active in iteration it, the marker -1 (the iteration
// continues); else the true residual's r.p gives sig of slot it and each system's restart flag for iteration it + 1 (the stock
// condition with the true residual: it + 1 < imax and sqrt(r.M^-1 r)/|b| > tol).
__global__ void __launch_bounds__(kThreads) k_cg_check_fin(QeqBuffers qb, int nparts, int it, int imax, double tol) {
  __shared__ double sh[4 * kThreads / 64];
  const int* pa = ACT(qb.scal, it - 1);
  int* rs = RST(qb.scal, it);
  if (pa[0] != 0 || pa[1] != 0) {
    if (threadIdx.x == 0) { rs[0] = -1; rs[1] = -1; }
    return;
  }
  if (check_skippable<kThreads>(qb, it, tol, sh)) {   // verified by the bound: the recursion's r.M^-1 r stands
    if (threadIdx.x == 0)
      for (int s = 0; s < 2; ++s) { SIG(qb.scal, it, s) = SIG(qb.scal, it - 1, s); rs[s] = 0; }
    return;
  }
  double v[4];
  reduce_partials<4>(qb.partial_rs, v, sh, nparts);
  if (threadIdx.x == 0) {
    for (int s = 0; s < 2; ++s) {
      const double bn = sqrt(v[s]);
      SIG(qb.scal, it, s) = v[2 + s];
      rs[s] = (it + 1 < imax && sqrt(v[2 + s]) / bn > tol) ? 1 : 0;
    }
    // running counts since the scalars were allocated: checks whose product ran, and restarts
    qb.scal[4] += 1.0;
    if (rs[0] || rs[1]) qb.scal[5] += 1.0;
  }
}

// x += alpha d, r -= alpha q, p = r / eta, partial r.p   (iteration it >= 1); few large blocks, since every block first
// re-reduces all partials of the product
__global__ void __launch_bounds__(kUpdThreads) k_cg_update(int n, QeqBuffers qb, int it) {
  __shared__ double sh[2 * kUpdThreads / 64];
  const int* act = ACT(qb.scal, it - 1);
  const int a[2] = {act[0], act[1]};
  if (!a[0] && !a[1]) return;
  double dq[2];
  reduce_partials<2, kUpdThreads>(qb.partial_mv, dq, sh, mv_blocks(n));
  double alpha[2];
  for (int s = 0; s < 2; ++s) alpha[s] = a[s] ? SIG(qb.scal, it - 1, s) / dq[s] : 0.0;
  double acc[2] = {0, 0};
  for (int i = blockIdx.x * kUpdThreads + threadIdx.x; i < n; i += gridDim.x * kUpdThreads) {
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
  block_reduce<2, kUpdThreads>(acc, sh);
  if (threadIdx.x == 0) { qb.partial[2 * kCgBlocks + blockIdx.x * 2] = acc[0]; qb.partial[2 * kCgBlocks + blockIdx.x * 2 + 1] = acc[1]; }
}

__global__ void __launch_bounds__(kThreads) k_cg_sums(int n, QeqBuffers qb) {
  __shared__ double sh[2 * kThreads / 64];
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
  __shared__ double sh[2 * kThreads / 64];
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

// ---- several independent systems packed in one engine ---------------------------------------------------------------
// The same two CG systems (s, t) per packed system, with that system's own scalars: one block per system sums its rows' values
// (thread t takes the rows t, t + kSysThreads, ... of the system in order, then block_reduce's fixed order), so a system's solve
// does not depend on the systems beside it.
// pk[10 s + 4 k + {0, 1, 2, 3}] = |b|, r.p, alpha, beta of solve k; pk[10 s + 8, 9] = the sums of s and t for the charges;
// pki[4 s + 2 k + {0, 1}] = state, the stock loop index. State 1: active; 0: stopped; 2: stopped and its true residual verified
// by the refinement check (only with refinement). rowv[4 i + {0, 1}] = d.q (b.b at the start), rowv[4 i + {2, 3}] = r.p.
__global__ void k_pk_cg_init(int n, QeqBuffers qb, double* rowv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double hi = qb.hinv[i];
  for (int s = 0; s < 2; ++s) {
    const double r = qb.b[2 * i + s] - qb.q[2 * i + s];
    const double d = r * hi;
    qb.r[2 * i + s] = r;
    qb.p[2 * i + s] = d;
    rowv[4 * i + s] = qb.b[2 * i + s] * qb.b[2 * i + s];
    rowv[4 * i + 2 + s] = r * d;
  }
}

// kind 0: start (|b|, r.p, state for iteration 1); 3: sums of s and t; 4: refinement check after iteration it (stopped solves:
// the true residual's r.p restarts the solve with beta 0 while it exceeds tol, else the solve is verified). The iterations'
// own scalars (alpha, r.p, beta, state) are formed by k_pk_cg_step.
constexpr int kSysThreads = 256;   // threads per system of the packed reductions
__global__ void __launch_bounds__(kSysThreads) k_pk_cg_seg(int kind, int S, const int* sys_start, const double* rowv, const double* x,
                                                          double* pk, int* pki, int it, int imax, double tol, int* any) {
  const int sy = blockIdx.x;   // one block per system (grid = S)
  if (sy >= S) return;
  __shared__ double sh[4 * kSysThreads / 64];
  double v[4] = {0, 0, 0, 0};
  for (int i = sys_start[sy] + threadIdx.x; i < sys_start[sy + 1]; i += kSysThreads) {
    if (kind == 3) { v[0] += x[2 * i]; v[1] += x[2 * i + 1]; }
    else for (int k = 0; k < 4; ++k) v[k] += rowv[4 * i + k];
  }
  block_reduce<4, kSysThreads>(v, sh);
  if (threadIdx.x != 0) return;
  double* P = pk + 10 * (size_t)sy;
  int* A = pki + 4 * (size_t)sy;
  if (kind == 3) { P[8] = v[0]; P[9] = v[1]; return; }
  for (int k = 0; k < 2; ++k) {
    double* q = P + 4 * k;
    int* a = A + 2 * k;
    if (kind == 0) {
      q[0] = sqrt(v[k]);
      q[1] = v[2 + k];
      q[3] = 0.0;
      a[0] = (1 < imax && sqrt(v[2 + k]) / q[0] > tol) ? 1 : 0;
      a[1] = 1;
      if (a[0]) atomicOr(any, 1);
    } else if (a[0] == 0) {   // kind 4
      q[1] = v[2 + k];
      if (a[1] < imax && sqrt(v[2 + k]) / q[0] > tol) {
        a[0] = 1;
        q[3] = 0.0;
        atomicOr(any + it, 1);
      } else {
        a[0] = 2;
      }
    }
  }
}

// w = (diag(eta) + H) p, d = p + beta d, q = w + beta q and d.q per row of the active solves (iteration it)
template <class VA, class TC>
__global__ void __launch_bounds__(kMvThreads) k_pk_cg_mvdir(int n, int maxnb, const TC* __restrict__ col, const VA val, QeqBuffers qb,
                                                           const int* sysid, const double* pk, const int* pki, double* rowv, int it) {
  const int i = product_row();
  int act[2] = {0, 0}, sy = 0;
  if (i < n) {
    sy = sysid[i];
    act[0] = pki[4 * sy] == 1;
    act[1] = pki[4 * sy + 2] == 1;
  }
  if (!act[0] && !act[1]) return;   // uniform over the row's lane group
  const double2* __restrict__ p2 = reinterpret_cast<const double2*>(qb.p);
  const double2 y = row_product<VA, TC>(i, n, maxnb, qb.cnt, col, val, p2);
  if (lane_id() % kRowLanes != 0) return;
  const double2 pi = p2[i];
  const double ei = qb.eta[i];
  const double w[2] = {y.x + ei * pi.x, y.y + ei * pi.y}, pv[2] = {pi.x, pi.y};
  for (int s = 0; s < 2; ++s) {
    if (!act[s]) continue;
    const int k = 2 * i + s;
    const double beta = pk[10 * sy + 4 * s + 3];
    const double d = it == 1 ? pv[s] : pv[s] + beta * qb.d[k];
    const double q = it == 1 ? w[s] : w[s] + beta * qb.q[k];
    qb.d[k] = d;
    qb.q[k] = q;
    rowv[4 * i + s] = d * q;
  }
}

// refinement check, rows of the systems with a stopped solve: the true residual r = b - (diag(eta) + H) x with the two-float
// matrix, p = r / eta and r.p for those solves (the rows of the other solves are left alone)
template <class TC>
__global__ void __launch_bounds__(kMvThreads) k_pk_cg_check(int n, int maxnb, const TC* __restrict__ col, const ValDS val, QeqBuffers qb,
                                                           const int* sysid, const int* pki, double* rowv) {
  const int i = product_row();
  int chk[2] = {0, 0};
  if (i < n) {
    const int sy = sysid[i];
    chk[0] = pki[4 * sy] == 0;
    chk[1] = pki[4 * sy + 2] == 0;
  }
  if (!chk[0] && !chk[1]) return;   // uniform over the row's lane group
  const double2* __restrict__ x2 = reinterpret_cast<const double2*>(qb.x);
  const double2 y = row_product<ValDS, TC>(i, n, maxnb, qb.cnt, col, val, x2);
  if (lane_id() % kRowLanes != 0) return;
  const double2 xi = x2[i];
  const double ei = qb.eta[i], hi = qb.hinv[i];
  const double w[2] = {y.x + ei * xi.x, y.y + ei * xi.y};
  for (int s = 0; s < 2; ++s) {
    if (!chk[s]) continue;
    const int k = 2 * i + s;
    const double r = qb.b[k] - w[s];
    const double p = r * hi;
    qb.r[k] = r;
    qb.p[k] = p;
    rowv[4 * i + 2 + s] = r * p;
  }
}

// One CG iteration's scalar and vector steps after the product (iteration it), one block per system: alpha from the system's
// d.q (the rows' values from k_pk_cg_mvdir), x += alpha d, r -= alpha q, p = r / eta for the system's rows, then the new r.p,
// beta and the stock stopping test, in one launch. Thread t takes the system's rows t, t + kStepThreads, ..., and the sums
// follow block_reduce's fixed order, so a system's solve depends only on its own rows. 1024 threads, since one block updates
// all the system's rows (a lone 9000-atom system: 0.32 ms of CG per step with 256 threads, 0.20 ms with 1024).
constexpr int kStepThreads = 1024;
__global__ void __launch_bounds__(kStepThreads) k_pk_cg_step(QeqBuffers qb, int S, const int* sys_start, double* pk, int* pki,
                                                            const double* rowv, int it, int imax, double tol, int* any) {
  const int sy = blockIdx.x;   // one block per system (grid = S)
  if (sy >= S) return;
  double* P = pk + 10 * (size_t)sy;
  int* A = pki + 4 * (size_t)sy;
  const bool act[2] = {A[0] == 1, A[2] == 1};
  if (!act[0] && !act[1]) return;   // uniform over the block
  __shared__ double sh[2 * kStepThreads / 64];
  const int a = sys_start[sy], b = sys_start[sy + 1];
  double v[2] = {0, 0};
  for (int i = a + threadIdx.x; i < b; i += kStepThreads) { v[0] += rowv[4 * i]; v[1] += rowv[4 * i + 1]; }
  block_reduce<2, kStepThreads>(v, sh);
  const double alpha[2] = {act[0] ? P[1] / v[0] : 0.0, act[1] ? P[4 + 1] / v[1] : 0.0};
  double w[2] = {0, 0};
  for (int i = a + threadIdx.x; i < b; i += kStepThreads) {
    const double hi = qb.hinv[i];
    for (int s = 0; s < 2; ++s) {
      if (!act[s]) continue;
      const int k = 2 * i + s;
      qb.x[k] += alpha[s] * qb.d[k];
      qb.r[k] -= alpha[s] * qb.q[k];
      const double p = qb.r[k] * hi;
      qb.p[k] = p;
      w[s] += qb.r[k] * p;
    }
  }
  block_reduce<2, kStepThreads>(w, sh);
  if (threadIdx.x != 0) return;
  for (int s = 0; s < 2; ++s) {
    if (!act[s]) continue;
    double* q = P + 4 * s;
    int* as = A + 2 * s;
    q[2] = alpha[s];
    q[3] = w[s] / q[1];
    q[1] = w[s];
    as[1] += 1;
    as[0] = (as[1] < imax && sqrt(w[s]) / q[0] > tol) ? 1 : 0;
    if (as[0]) atomicOr(any + it, 1);
  }
}

// q = s - (sum s / sum t) t per system, then the history shift of calculate_Q
__global__ void k_pk_cg_charges(int n, QeqBuffers qb, const int* sysid, const double* pk, double* q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int sy = sysid[i];
  const double u = pk[10 * sy + 8] / pk[10 * sy + 9];
  const double s = qb.x[2 * i], t = qb.x[2 * i + 1];
  q[i] = s - u * t;
  for (int k = 3; k > 0; --k) {
    qb.s_hist[k * n + i] = qb.s_hist[(k - 1) * n + i];
    qb.t_hist[k * n + i] = qb.t_hist[(k - 1) * n + i];
  }
  qb.s_hist[i] = s;
  qb.t_hist[i] = t;
}

inline int pk_blocks(long long items, int per) { return (int)((items + per - 1) / per); }

// Calls f(values, columns) with the matrix in its stored format: FP64 or FP32 values, or (ds) the two-float values of the
// refinement's residuals; 16- or 32-bit columns.
template <class F>
void with_matrix(const QeqBuffers& qb, bool ds, F&& f) {
  if (ds) {
    const ValDS v{qb.val32, qb.lo32};
    if (qb.narrow) f(v, (const unsigned short*)qb.col16); else f(v, (const int*)qb.col);
  } else if (qb.fp32) {
    const ValF v{qb.val32};
    if (qb.narrow) f(v, (const unsigned short*)qb.col16); else f(v, (const int*)qb.col);
  } else {
    const ValD v{qb.val};
    if (qb.narrow) f(v, (const unsigned short*)qb.col16); else f(v, (const int*)qb.col);
  }
}
template <class P>
using col_t = typename std::remove_const<typename std::remove_pointer<P>::type>::type;

template <int kMode>
void spmv2(int n, int maxnb, QeqBuffers& qb, bool ds, const double* x, double* y, int it, hipStream_t s, double tol = 0.0) {
  with_matrix(qb, ds, [&](auto v, auto col) {
    hipLaunchKernelGGL((k_spmv2<decltype(v), col_t<decltype(col)>, kMode>), dim3(mv_grid(n)), dim3(kMvThreads), 0, s, n, maxnb, qb.cnt,
                       col, v, qb.eta, x, y, qb, it, tol);
  });
}

}  // namespace

void launch_qeq_matvec1(const State& st, const QeqBuffers& qb, const double* x, double* y, hipStream_t s) {
  // the fix's sparse_matvec: FP64 values where the engine keeps them (FP64 mode, or the two-float pair under refinement)
  with_matrix(qb, qb.fp32 && qb.refine, [&](auto v, auto col) {
    hipLaunchKernelGGL((k_spmv1<decltype(v), col_t<decltype(col)>>), dim3((st.n * 64 + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                       st.n, st.maxnb, qb.cnt, col, v, qb.eta, x, y);
  });
}

void launch_qeq_init(const State& st, QeqBuffers& qb, int imax, double tol, int use_hist, hipStream_t s) {
  const int n = st.n;
  hipLaunchKernelGGL(k_cg_guess, dim3(kCgBlocks), dim3(kThreads), 0, s, n, qb, use_hist);
  spmv2<1>(n, st.maxnb, qb, qb.fp32 && qb.refine, qb.x, qb.q, 0, s);
  hipLaunchKernelGGL(k_cg_init_fin<kThreads>, dim3(1), dim3(kThreads), 0, s, qb, (const double*)qb.partial_mv, mv_blocks(n), imax, tol);
}

void launch_qeq_guess(const State& st, QeqBuffers& qb, int use_hist, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_guess, dim3(kCgBlocks), dim3(kThreads), 0, s, st.n, qb, use_hist);
}

void launch_qeq_init_fused(const State& st, QeqBuffers& qb, int imax, double tol, hipStream_t s) {
  // the build leaves a partial per 4 rows: one block of 512 threads reduces them (twice the threads of the product's case)
  hipLaunchKernelGGL(k_cg_init_fin<512>, dim3(1), dim3(512), 0, s, qb, (const double*)qb.partial_qb, qeq_build_blocks(st.n), imax,
                     tol);
}

namespace {
void mvdir_launch(const State& st, QeqBuffers& qb, int it, int imax, double tol, int chk, hipStream_t s) {
  const int n = st.n;
  with_matrix(qb, false, [&](auto v, auto col) {
    hipLaunchKernelGGL((k_cg_mvdir<decltype(v), col_t<decltype(col)>>), dim3(mv_grid(n)), dim3(kMvThreads), 0, s, n, st.maxnb, col, v, qb,
                       it, imax, tol, chk);
  });
}
}  // namespace

void launch_qeq_mvdir(const State& st, QeqBuffers& qb, int it, int imax, double tol, hipStream_t s) { mvdir_launch(st, qb, it, imax, tol, 0, s); }
void launch_qeq_mvdir_chk(const State& st, QeqBuffers& qb, int it, int imax, double tol, hipStream_t s) { mvdir_launch(st, qb, it, imax, tol, 1, s); }

void launch_qeq_restart(const State& st, QeqBuffers& qb, int it, int imax, double tol, hipStream_t s) {
  const int n = st.n;
  hipLaunchKernelGGL(k_check_gate, dim3(kCgBlocks), dim3(kThreads), 0, s, n, qb, it);
  spmv2<2>(n, st.maxnb, qb, true, qb.x, qb.q, it, s, tol);
  hipLaunchKernelGGL(k_cg_check_fin, dim3(1), dim3(kThreads), 0, s, qb, mv_blocks(n), it, imax, tol);
}

int qeq_mv_partials(int n) { return mv_blocks(n); }

void launch_qeq_update(const State& st, QeqBuffers& qb, int it, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_update, dim3(kUpdBlocks), dim3(kUpdThreads), 0, s, st.n, qb, it);
}

void launch_qeq_finish(const State& st, QeqBuffers& qb, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_sums, dim3(kCgBlocks), dim3(kThreads), 0, s, st.n, qb);
  hipLaunchKernelGGL(k_cg_charges, dim3(kCgBlocks), dim3(kThreads), 0, s, st.n, qb, st.q);
}

// ---- packed systems ----
void launch_pk_qeq_init(const State& st, QeqBuffers& qb, const PackedQeq& pq, int imax, double tol, int use_hist, hipStream_t s) {
  const int n = st.n;
  hipLaunchKernelGGL(k_cg_guess, dim3(kCgBlocks), dim3(kThreads), 0, s, n, qb, use_hist);
  spmv2<0>(n, st.maxnb, qb, qb.fp32 && qb.refine, qb.x, qb.q, 0, s);
  hipLaunchKernelGGL(k_pk_cg_init, dim3(pk_blocks(n, 256)), dim3(256), 0, s, n, qb, pq.rowv);
  hipLaunchKernelGGL(k_pk_cg_seg, dim3(pq.S), dim3(kSysThreads), 0, s, 0, pq.S, pq.sys_start, pq.rowv, (const double*)qb.x, pq.pk,
                     pq.pki, 0, imax, tol, pq.any);
}

void launch_pk_qeq_iteration(const State& st, QeqBuffers& qb, const PackedQeq& pq, int it, int imax, double tol, hipStream_t s) {
  const int n = st.n;
  with_matrix(qb, false, [&](auto v, auto col) {
    hipLaunchKernelGGL((k_pk_cg_mvdir<decltype(v), col_t<decltype(col)>>), dim3(mv_grid(n)), dim3(kMvThreads), 0, s, n, st.maxnb, col, v, qb,
                       pq.sysid, (const double*)pq.pk, (const int*)pq.pki, pq.rowv, it);
  });
  hipLaunchKernelGGL(k_pk_cg_step, dim3(pq.S), dim3(kStepThreads), 0, s, qb, pq.S, pq.sys_start, pq.pk, pq.pki, (const double*)pq.rowv, it,
                     imax, tol, pq.any);
}

void launch_pk_qeq_restart(const State& st, QeqBuffers& qb, const PackedQeq& pq, int it, int imax, double tol, hipStream_t s) {
  const int n = st.n;
  const ValDS v{qb.val32, qb.lo32};
  if (qb.narrow)
    hipLaunchKernelGGL((k_pk_cg_check<unsigned short>), dim3(mv_grid(n)), dim3(kMvThreads), 0, s, n, st.maxnb, (const unsigned short*)qb.col16, v,
                       qb, pq.sysid, (const int*)pq.pki, pq.rowv);
  else
    hipLaunchKernelGGL((k_pk_cg_check<int>), dim3(mv_grid(n)), dim3(kMvThreads), 0, s, n, st.maxnb, (const int*)qb.col, v, qb, pq.sysid,
                       (const int*)pq.pki, pq.rowv);
  hipLaunchKernelGGL(k_pk_cg_seg, dim3(pq.S), dim3(kSysThreads), 0, s, 4, pq.S, pq.sys_start, pq.rowv, (const double*)qb.x, pq.pk,
                     pq.pki, it, imax, tol, pq.any);
}

void launch_pk_qeq_finish(const State& st, QeqBuffers& qb, const PackedQeq& pq, hipStream_t s) {
  hipLaunchKernelGGL(k_pk_cg_seg, dim3(pq.S), dim3(kSysThreads), 0, s, 3, pq.S, pq.sys_start, pq.rowv, (const double*)qb.x, pq.pk,
                     pq.pki, 0, 0, 0.0, pq.any);
  hipLaunchKernelGGL(k_pk_cg_charges, dim3(pk_blocks(st.n, 256)), dim3(256), 0, s, st.n, qb, pq.sysid, (const double*)pq.pk, st.q);
}

}  // namespace rxa
