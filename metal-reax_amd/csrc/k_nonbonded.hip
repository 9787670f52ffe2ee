// Tapered van der Waals + shielded Coulomb + polarization, and the charge-equilibration matrix H.
//
// One 64-lane wave per owned atom walks that atom's full skin list; every pair (i, j, s) with r <= nonb_cut is evaluated from both
// ends, each end keeping only its own gradient (-CE dvec on i) and half of the pair energy, so no atom writes another atom's force
// (atomic-free, bitwise repeatable: fixed lane assignment + fixed butterfly reduction). The distance is the canonical one, so both
// ends agree on the cut-off decision and on r. Reference: src/cpu/nonbonded.cpp:18-82 (pair loop, owner-computes counting: a
// physical pair counted once = half from each end here), include/reaxmetal/terms.hpp:262-325 (formulas), terms_host.hpp:17-39
// (taper). FP64 mode uses the reference's Horner taper and its exponents (pow(x, 0.33333333333333) for the cube root), each
// pow evaluated as exp(y log x); FP32 terms use the algebraically equal t-form of the taper (no cancellation of O(10)
// coefficients in float) and the float cube root. Both precisions run the same two passes: over the skin list with the pairs
// within the cutoff compacted into full waves, and over the charge-matrix rows when those are current.
//
// H (stock fix qeq/reaxff compute_H / calculate_H, LAMMPS fix_qeq_reaxff.cpp): H_ij = 14.4 Tap_q(r) / (r^3 + (g_i g_j)^-1.5)^(1/3) for
// r^2 <= swb^2, stored as full rows (both directions, every image): FP64 values in FP64 mode, FP32 values otherwise; 16-bit
// column indices when the system has at most 65535 atoms.
#include "common.hpp"
#include "launch.hpp"

namespace rxa {
namespace {

struct NbAcc { double gx = 0, gy = 0, gz = 0, ev = 0, ee = 0, v[6] = {0, 0, 0, 0, 0, 0}; };

// this end's gradient -CE dvec, half of the pair energies and half of the pair virial r_ij (x) f_j = -CE dvec (x) dvec
__device__ inline void nb_acc(NbAcc& a, double ce, double evdw, double eel, double dx, double dy, double dz) {
  a.gx += -ce * dx; a.gy += -ce * dy; a.gz += -ce * dz;
  a.ev += 0.5 * evdw;
  a.ee += 0.5 * eel;
  const double hv = -0.5 * ce;
  a.v[0] += hv * dx * dx; a.v[1] += hv * dy * dy; a.v[2] += hv * dz * dz; a.v[3] += hv * dx * dy; a.v[4] += hv * dx * dz; a.v[5] += hv * dy * dz;
}

// FP64 terms of one pair (i, j) within nonb_cut at distance r, dvec = x_j - x_i: the reference's formulas in double with its
// Horner taper in r and its exponents (0.33333333333333 for the shielded Coulomb's cube root as written), pow(x, y) evaluated
// as exp(y log x) (a few ulp), the type pair's (1 / gamma_w)^p_vdW1 and the reciprocals of r_vdW and rcore from the table,
// e1 = e2^2 for the two Morse exponentials, and one division for 1 / r
__device__ inline void nb_pair_d(const Globals& g, const DevPair* __restrict__ row, double qi, double qj, int tj, double r,
                                 double dx, double dy, double dz, NbAcc& a) {
  const DevPair& P = row[tj];
  const double p_vdW1 = g.p_vdW1, p1i = 1.0 / g.p_vdW1;
  const int vt = g.vdw_type;
  const double ir = 1.0 / r;
  const double* c = g.tap;
  double Tap = c[7] * r + c[6];
  Tap = Tap * r + c[5]; Tap = Tap * r + c[4]; Tap = Tap * r + c[3]; Tap = Tap * r + c[2]; Tap = Tap * r + c[1]; Tap = Tap * r + c[0];
  double dTap = 7 * c[7] * r + 6 * c[6];
  dTap = dTap * r + 5 * c[5]; dTap = dTap * r + 4 * c[4]; dTap = dTap * r + 3 * c[3]; dTap = dTap * r + 2 * c[2];
  dTap += c[1] * ir;
  const double alpha = P.alpha, D = P.D, a_r = alpha * P.irvdw;
  double evdw, CEvd;
  if (vt == 1 || vt == 3) {
    const double pr = exp(p_vdW1 * log(r)), S = pr + P.pgw;
    const double fn13 = exp(p1i * log(S));
    const double e2 = exp(0.5 * alpha * (1.0 - fn13 * P.irvdw)), e1 = e2 * e2;
    const double e0 = D * (e1 - 2.0 * e2);
    evdw = Tap * e0;
    const double dfn13 = fn13 / S * pr * (ir * ir);
    CEvd = dTap * e0 - Tap * D * a_r * (e1 - e2) * dfn13;
  } else {
    const double e2 = exp(0.5 * alpha * (1.0 - r * P.irvdw)), e1 = e2 * e2;
    const double e0 = D * (e1 - 2.0 * e2);
    evdw = Tap * e0;
    CEvd = dTap * e0 - Tap * D * a_r * (e1 - e2) * ir;
  }
  if (vt == 2 || vt == 3) {
    const double ec = P.ecore * exp(P.acore * (1.0 - r * P.ircore));
    evdw += Tap * ec;
    CEvd += dTap * ec + Tap * (-(P.acore * P.ircore) * ec) * ir;
    if (g.lg) {
      const double r6 = (r * r * r) * (r * r * r), idn = 1.0 / (r6 + P.re6);
      const double elg = -P.lgcij * idn;
      evdw += Tap * elg;
      CEvd += dTap * elg + Tap * (-6.0 * elg * (r6 * ir) * idn) * ir;
    }
  }
  const double d1 = r * r * r + P.gamma, id3 = 1.0 / exp(0.33333333333333 * log(d1));
  const double qq = 332.06371 * qi * qj;
  const double eel = qq * Tap * id3;
  const double CE = CEvd + qq * (dTap - Tap * r / d1) * id3;
  nb_acc(a, CE, evdw, eel, dx, dy, dz);
}

// Single-precision terms (mixed and FP32 modes): the same pair loop with the hardware base-2 exponential and logarithm
// (v_exp_f32 / v_log_f32, about 1 ulp) for exp, pow and the cube root, e1 = e2^2 for the two Morse exponentials, a pair table
// in float with (1/gamma_w)^p_vdW1 precomputed, and a squared-distance pre-test before the FP64 square root of the cut-off test.
__device__ inline float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ inline float flog2(float x) { return __builtin_amdgcn_logf(x); }

// one pair (i, j) within nonb_cut at distance rd, dvec = x_j - x_i: this end's gradient, half of the pair energy and virial
__device__ inline void nb_pair_f(const Globals& g, const NbPairF* __restrict__ row, float qi, float qj, int tj, double rd, double dx,
                             double dy, double dz, NbAcc& a) {
  // divisions by constants are products with reciprocals (table or loop-invariant), divisions by pair quantities use the
  // hardware reciprocal (1 ulp): single-precision terms, within their own rounding
  const float p_vdW1 = (float)g.p_vdW1, p1i = (float)(1.0 / g.p_vdW1);
  const int vt = g.vdw_type, lg = g.lg;
  const float swa = (float)g.nonb_low, itw = (float)(1.0 / (g.nonb_cut - g.nonb_low));
  const float L2E = 1.4426950408889634f;
  const NbPairF P = row[tj];
  const float r = (float)rd, ir = __builtin_amdgcn_rcpf(r);
  const float x = (r - swa) * itw, x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
  const float Tap = 1.0f - x4 * (35.0f - 84.0f * x + 70.0f * x2 - 20.0f * x3);
  const float dTap = (x3 * (-140.0f + 420.0f * x - 420.0f * x2 + 140.0f * x3)) * (itw * ir);
  float evdw, CEvd;
  if (vt == 1 || vt == 3) {
    const float pr = fexp2(p_vdW1 * flog2(r)), S = pr + P.pgw;
    const float fn13 = fexp2(p1i * flog2(S));
    const float e2 = fexp2((0.5f * L2E) * P.alpha * (1.0f - fn13 * P.irvdw)), e1 = e2 * e2;
    const float e0 = P.D * (e1 - 2.0f * e2);
    evdw = Tap * e0;
    const float dfn13 = fn13 * __builtin_amdgcn_rcpf(S) * pr * (ir * ir);
    CEvd = dTap * e0 - Tap * P.D * P.a_r * (e1 - e2) * dfn13;
  } else {
    const float e2 = fexp2((0.5f * L2E) * P.alpha * (1.0f - r * P.irvdw)), e1 = e2 * e2;
    const float e0 = P.D * (e1 - 2.0f * e2);
    evdw = Tap * e0;
    CEvd = dTap * e0 - Tap * P.D * P.a_r * (e1 - e2) * ir;
  }
  if (vt == 2 || vt == 3) {
    const float ec = P.ecore * fexp2(L2E * P.acore * (1.0f - r * P.ircore));
    evdw += Tap * ec;
    CEvd += dTap * ec + Tap * (-P.ac_r * ec) * ir;
    if (lg) {
      const float r3 = r * r * r, r6 = r3 * r3, idn = __builtin_amdgcn_rcpf(r6 + P.re6);
      const float elg = -P.lgcij * idn;
      evdw += Tap * elg;
      CEvd += dTap * elg + Tap * (-6.0f * elg * (r6 * ir) * idn) * ir;
    }
  }
  const float d1 = r * r * r + P.gamma, rc = fexp2(flog2(d1) * (-1.0f / 3.0f));   // d1^(-1/3)
  const float qq = 332.06371f * qi * qj;
  const float eel = qq * Tap * rc;
  const float CE = CEvd + qq * (dTap - Tap * r * __builtin_amdgcn_rcpf(d1)) * rc;
  const double ce = (double)CE;
  a.gx += -ce * dx; a.gy += -ce * dy; a.gz += -ce * dz;
  a.ev += 0.5 * (double)evdw;
  a.ee += 0.5 * (double)eel;
  const double hv = -0.5 * ce;
  a.v[0] += hv * dx * dx; a.v[1] += hv * dy * dy; a.v[2] += hv * dz * dz; a.v[3] += hv * dx * dy; a.v[4] += hv * dx * dz; a.v[5] += hv * dy * dz;
}

// the per-precision parts of the non-bonded kernels below: the charges' type, the pair table row, the pair evaluation and the
// polarization energy
template <bool D> struct NbPrec;
template <> struct NbPrec<false> {
  using Q = float;
  using Row = NbPairF;
  static __device__ const NbPairF* row(const Globals& g, const Tables& t, int ti) { return t.nbf + ti * g.ntypes; }
  static __device__ void pair(const Globals& g, const NbPairF* row, float qi, float qj, int tj, double rd, double dx, double dy,
                              double dz, NbAcc& a) {
    nb_pair_f(g, row, qi, qj, tj, rd, dx, dy, dz, a);
  }
  static __device__ double pol(const DevSpecies& sp, double qd) {
    const float q = (float)qd;
    return (double)(23.02f * ((float)sp.chi * q + ((float)sp.eta / 2.0f) * (q * q)));
  }
};
template <> struct NbPrec<true> {
  using Q = double;
  using Row = DevPair;
  static __device__ const DevPair* row(const Globals& g, const Tables& t, int ti) { return t.pair + ti * g.ntypes; }
  static __device__ void pair(const Globals& g, const DevPair* row, double qi, double qj, int tj, double rd, double dx, double dy,
                              double dz, NbAcc& a) {
    nb_pair_d(g, row, qi, qj, tj, rd, dx, dy, dz, a);
  }
  static __device__ double pol(const DevSpecies& sp, double q) { return 23.02 * (sp.chi * q + (sp.eta / 2.) * (q * q)); }
};

// wave sums in a fixed butterfly order, then lane 0 stores the atom's gradient, energies and virial
template <bool D>
__device__ inline void nb_finish(const State& st, const Tables& t, int i, int ti, int lane, const NbAcc& a) {
  double gx = wave_sum(a.gx), gy = wave_sum(a.gy), gz = wave_sum(a.gz), ev = wave_sum(a.ev), ee = wave_sum(a.ee);
  double v0 = wave_sum(a.v[0]), v1 = wave_sum(a.v[1]), v2 = wave_sum(a.v[2]), v3 = wave_sum(a.v[3]), v4 = wave_sum(a.v[4]),
         v5 = wave_sum(a.v[5]);
  if (lane == 0) {
    const int n = st.n;
    st.gnb_x[i] = gx; st.gnb_y[i] = gy; st.gnb_z[i] = gz;
    st.eatom[T_VDW * n + i] = ev;
    st.eatom[T_ELE * n + i] = ee;
    st.eatom[T_POL * n + i] = ti >= 0 ? NbPrec<D>::pol(t.sp[ti], st.q[i]) : 0.0;
    st.vatom[0 * n + i] = v0; st.vatom[1 * n + i] = v1; st.vatom[2 * n + i] = v2;
    st.vatom[3 * n + i] = v3; st.vatom[4 * n + i] = v4; st.vatom[5 * n + i] = v5;
  }
}

// Wave per atom over the skin list: the distance test runs on all lanes, pairs within nonb_cut are compacted (list order) into
// a per-wave LDS ring, and the term arithmetic runs on full waves of 64 pairs.
constexpr int kNbWaves = 4;
constexpr int kNbRing = 128;

template <bool D>
__device__ inline void nb_list_wave(const State& st, const Globals& g, const Tables& t, int i, int* rj, int* rt, double* rx,
                                    double* ry, double* rz, double* rr) {
  using Q = typename NbPrec<D>::Q;
  const int lane = lane_id();
  const int ti = st.type[i];
  NbAcc a;
  if (ti >= 0) {
    const size_t base = (size_t)i * st.maxnb;
    const int cnt = min(st.nb_cnt[i], st.maxnb);
    const Q qi = (Q)st.q[i];
    const double4 pi = st.pos[i];
    const double cut = g.nonb_cut, cut2 = cut * cut * (1.0 + 1e-12);
    const typename NbPrec<D>::Row* __restrict__ row = NbPrec<D>::row(g, t, ti);
    int head = 0, fill = 0;
    auto batch = [&](int m) {
      if (lane < m) {
        const int e = (head + lane) & (kNbRing - 1);
        const int j = rj[e];
        NbPrec<D>::pair(g, row, qi, (Q)st.q[j], rt[e], rr[e], rx[e], ry[e], rz[e], a);
      }
    };
    for (int k0 = 0; k0 < cnt; k0 += 64) {
      const int k = k0 + lane;
      bool keep = false;
      int j = 0, tj = -1;
      double dx = 0, dy = 0, dz = 0, rd = 0;
      if (k < cnt) {
        j = st.nb_j[base + k];
        const double r2 = pair_dvec_list(st, pi, j, st.nb_s[base + k], dx, dy, dz, tj);
        if (tj >= 0 && r2 <= cut2) {
          rd = sqrt(r2);
          keep = rd <= cut;
        }
      }
      const unsigned long long mk = __ballot(keep);
      if (keep) {
        const int e = (head + fill + lanes_below(mk)) & (kNbRing - 1);
        rj[e] = j; rt[e] = tj; rx[e] = dx; ry[e] = dy; rz[e] = dz; rr[e] = rd;
      }
      fill += __popcll(mk);
      __builtin_amdgcn_wave_barrier();
      if (fill >= 64) {
        batch(64);
        head = (head + 64) & (kNbRing - 1);
        fill -= 64;
        __builtin_amdgcn_wave_barrier();
      }
    }
    if (fill > 0) batch(fill);
  }
  nb_finish<D>(st, t, i, ti, lane, a);
}

__global__ void __launch_bounds__(64 * kNbWaves) k_nonbonded_f(State st, Globals g, Tables t) {
  __shared__ int rj_[kNbWaves][kNbRing], rt_[kNbWaves][kNbRing];
  __shared__ double rx_[kNbWaves][kNbRing], ry_[kNbWaves][kNbRing], rz_[kNbWaves][kNbRing], rr_[kNbWaves][kNbRing];
  const int w = threadIdx.x >> 6;
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= st.n) return;
  nb_list_wave<false>(st, g, t, i, rj_[w], rt_[w], rx_[w], ry_[w], rz_[w], rr_[w]);
}

__global__ void __launch_bounds__(64 * kNbWaves) k_nonbonded_d(State st, Globals g, Tables t) {
  __shared__ int rj_[kNbWaves][kNbRing], rt_[kNbWaves][kNbRing];
  __shared__ double rx_[kNbWaves][kNbRing], ry_[kNbWaves][kNbRing], rz_[kNbWaves][kNbRing], rr_[kNbWaves][kNbRing];
  const int w = threadIdx.x >> 6;
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= st.n) return;
  nb_list_wave<true>(st, g, t, i, rj_[w], rt_[w], rx_[w], ry_[w], rz_[w], rr_[w]);
}

// pair_dvec_list for a charge-matrix entry: the same operands in the same order, the shift vector found through the entry's key
#pragma clang fp contract(off)
__device__ inline double pair_dvec_key(const State& st, const double4 pi, const double4 pj, int key, double& dx, double& dy, double& dz) {
  const bool canon = (key & 0x8000) != 0;
  const double4 S = st.shv[pos_sys(pi.w) * st.shv_stride + (key & 0x7FFF)];
  const double ax = canon ? pj.x : pi.x, ay = canon ? pj.y : pi.y, az = canon ? pj.z : pi.z;
  const double bx = canon ? pi.x : pj.x, by = canon ? pi.y : pj.y, bz = canon ? pi.z : pj.z;
  const double ex = (ax + S.x) - bx, ey = (ay + S.y) - by, ez = (az + S.z) - bz;
  dx = canon ? ex : -ex;
  dy = canon ? ey : -ey;
  dz = canon ? ez : -ez;
  return dx * dx + dy * dy + dz * dz;
}
#pragma clang fp contract(on)

// The non-bonded terms over the charge-matrix rows: the rows hold, in skin-list order, exactly the pairs within the matrix
// cutoff, so every lane carries a pair (no distance pre-pass, no compaction) and a row costs its column and key (4 bytes per
// pair with 16-bit columns) instead of the 8-byte skin entries of all pairs out to the list radius. Lane l takes entries l,
// l + 64, ..., the assignment of the skin-list kernel's compacted batches, so the sums are the same.
template <bool D, class TC>
__device__ inline void nb_rows_wave(const State& st, const Globals& g, const Tables& t, const TC* __restrict__ col,
                                    const unsigned short* __restrict__ key, const int* __restrict__ rcnt, int i) {
  using Q = typename NbPrec<D>::Q;
  const int lane = lane_id();
  const int ti = st.type[i];
  NbAcc a;
  if (ti >= 0) {
    const size_t base = (size_t)i * st.maxnb;
    const int cnt = rcnt[i];
    const Q qi = (Q)st.q[i];
    const double4 pi = st.pos[i];
    const double cut = g.nonb_cut, cut2 = cut * cut * (1.0 + 1e-12);
    const typename NbPrec<D>::Row* __restrict__ row = NbPrec<D>::row(g, t, ti);
    // two-deep software pipeline as in the matrix build: columns/keys two strides ahead, positions and charges one ahead
    // (lanes past the row's end carry the row atom itself and key 0, valid indices never used)
    int jB = i, kyB = 0, jC = i, kyC = 0;
    if (lane < cnt) { jC = (int)col[base + lane]; kyC = key[base + lane]; }
    double4 pC = st.pos[jC];
    Q qC = (Q)st.q[jC];
    if (lane + 64 < cnt) { jB = (int)col[base + lane + 64]; kyB = key[base + lane + 64]; }
    for (int k = lane; k < cnt; k += 64) {
      const double4 pj = pC;
      const Q qj = qC;
      const int ky = kyC;
      jC = jB; kyC = kyB;
      pC = st.pos[jC];
      qC = (Q)st.q[jC];
      jB = i; kyB = 0;
      if (k + 128 < cnt) { jB = (int)col[base + k + 128]; kyB = key[base + k + 128]; }
      double dx, dy, dz;
      const double r2 = pair_dvec_key(st, pi, pj, ky, dx, dy, dz);
      if (r2 <= cut2) {
        const double rd = sqrt(r2);
        if (rd <= cut) NbPrec<D>::pair(g, row, qi, qj, pos_type(pj.w), rd, dx, dy, dz, a);
      }
    }
  }
  nb_finish<D>(st, t, i, ti, lane, a);
}

// (5 waves per SIMD: 96 VGPRs without spills; the compiler's choice, 106 VGPRs and 4 waves, is slower)
template <class TC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_nonbonded_rows(State st, Globals g, Tables t, const TC* __restrict__ col,
                                                        const unsigned short* __restrict__ key, const int* __restrict__ rcnt) {
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= st.n) return;
  nb_rows_wave<false, TC>(st, g, t, col, key, rcnt, i);
}

template <class TC>
__global__ void __launch_bounds__(256) k_nonbonded_rows_d(State st, Globals g, Tables t, const TC* __restrict__ col,
                                                          const unsigned short* __restrict__ key, const int* __restrict__ rcnt) {
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= st.n) return;
  nb_rows_wave<true, TC>(st, g, t, col, key, rcnt, i);
}

constexpr int kQbWaves = 4, kQbRing = 128;

// One wave per row atom over its skin list: a filter pass (distance only, list entries two chunks ahead and positions one chunk
// ahead in flight) pushes the pairs within the matrix cutoff, in list order, into a per-wave LDS ring; every 64 of them are
// finished together (taper, shielded 1/r, stores), so the costly part runs on full waves. Row entries are kept in skin-list
// order: slot m of the row is the m-th pair of the list within the cutoff. The 16-bit key is the canonical flag and the
// shift-table index the distance used.
// fuse: the row also forms the charge solve's start from the guess in x (k_cg_guess ran before): y = H x with the values as
// built (FP64, or the FP32 values a non-refined FP32 solve iterates with), r = b - (y + eta x), p = r / eta, and the row's
// b.b and r.p for the block partials; this replaces a separate product pass over the rows.
__device__ inline void qeq_build_row(const State& st, const Globals& g, const Tables& t, const QeqBuffers& qb, int i, int fuse,
                                     int* rj, int* rk, double* rr, double (&acc)[4]) {
  const int lane = lane_id();
  const int ti = st.type[i];
  const size_t base = (size_t)i * st.maxnb;
  const int cnt = ti >= 0 ? min(st.nb_cnt[i], st.maxnb) : 0;
  const double swb2 = g.q_swb * g.q_swb;
  const double* T = g.qtap;
  const double4 pi = st.pos[i];
  const double4* __restrict__ shv = st.shv + pos_sys(pi.w) * st.shv_stride;
  const double* __restrict__ shld = t.qeq_shld + max(ti, 0) * g.ntypes;
  int m = 0, head = 0, fill = 0;
  double lo_sum = 0.0;   // this lane's part of sum_j |lo_ij| (refinement)
  double y0 = 0.0, y1 = 0.0;   // fuse: this lane's part of (H x) of both systems
  const double2* __restrict__ x2 = reinterpret_cast<const double2*>(qb.x);
  auto batch = [&](int count) {
    if (lane < count) {
      const int e = (head + lane) & (kQbRing - 1);
      const int j = rj[e], kt = rk[e];
      const double r = sqrt(rr[e]);
      double Tap = T[7] * r + T[6];
      Tap = Tap * r + T[5]; Tap = Tap * r + T[4]; Tap = Tap * r + T[3]; Tap = Tap * r + T[2]; Tap = Tap * r + T[1]; Tap = Tap * r + T[0];
      const double h = (Tap * 14.4) * rcbrt_f64(r * r * r + shld[kt >> 16]);
      const size_t o = base + m + lane;
      if (qb.narrow) qb.col16[o] = (unsigned short)j; else qb.col[o] = j;
      double hv = h;   // the value the start's product uses
      if (qb.fp32) {
        const float hf = (float)h;
        qb.val32[o] = hf;
        // h - hf is exact in FP64; rounded to FP32 it leaves hf + lo within about 2^-48 (relative) of h
        if (qb.refine) {
          const float lo = (float)(h - (double)hf);
          qb.lo32[o] = lo;
          lo_sum += fabs((double)lo);
        } else {
          hv = (double)hf;
        }
      } else {
        qb.val[o] = h;
      }
      if (fuse) {
        const double2 xj = x2[j];
        y0 += hv * xj.x;
        y1 += hv * xj.y;
      }
      if (qb.keyed) qb.key[o] = (unsigned short)(kt & 0xFFFF);
    }
    m += count;
  };
  const int* __restrict__ lj = st.nb_j + base;
  const int* __restrict__ ls = st.nb_s + base;
  int jB = i, eB = 0, jC = i, eC = 0;
  if (lane < cnt) { jC = lj[lane]; eC = ls[lane]; }
  double4 pC = st.pos[jC];
  if (64 + lane < cnt) { jB = lj[64 + lane]; eB = ls[64 + lane]; }
  for (int k0 = 0; k0 < cnt; k0 += 64) {
    const int k = k0 + lane;
    const int j = jC, e = eC;
    const double4 pj = pC;
    jC = jB; eC = eB;
    if (k0 + 64 < cnt) pC = st.pos[jC];
    jB = i; eB = 0;
    if (k + 128 < cnt) { jB = lj[k + 128]; eB = ls[k + 128]; }
    bool keep = false;
    double r2 = 0;
    int kt = 0;
    if (k < cnt) {
      const bool canon = (e & kCanon) != 0;
      const int sidx = entry_shift_index(st, e);
      double dx, dy, dz;
      r2 = pair_dvec_shift(pi, pj, canon, shv[sidx], dx, dy, dz);
      const int tj = pos_type(pj.w);
      keep = tj >= 0 && r2 <= swb2;
      kt = (qb.keyed ? (canon ? 0x8000 : 0) | sidx : 0) | (tj << 16);   // a keyed table has at most 32768 entries
    }
    const unsigned long long msk = __ballot(keep);
    if (keep) {
      const int e2 = (head + fill + lanes_below(msk)) & (kQbRing - 1);
      rj[e2] = j; rk[e2] = kt; rr[e2] = r2;
    }
    fill += __popcll(msk);
    __builtin_amdgcn_wave_barrier();
    if (fill >= 64) {
      batch(64);
      head = (head + 64) & (kQbRing - 1);
      fill -= 64;
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (fill > 0) batch(fill);
  if (qb.refine) lo_sum = wave_sum(lo_sum);
  if (fuse) { y0 = wave_sum(y0); y1 = wave_sum(y1); }
  if (lane == 0) {
    qb.cnt[i] = m;
    if (qb.refine) qb.lorow[i] = lo_sum;
    const double eta = ti >= 0 ? t.qeq_eta[ti] : 1.0;
    const double hinv = 1.0 / eta;
    const double chi = ti >= 0 ? t.qeq_chi[ti] : 0.0;
    qb.eta[i] = eta;
    qb.hinv[i] = hinv;
    qb.chi[i] = chi;
    if (fuse) {   // the start of k_spmv2's kMode 1 for this row (b from this build's chi: the guess kernel ran before it)
      const double2 xi = x2[i];
      const double wv[2] = {y0 + eta * xi.x, y1 + eta * xi.y};
      for (int s = 0; s < 2; ++s) {
        const double b = s == 0 ? -chi : -1.0;
        qb.b[2 * i + s] = b;
        const double r = b - wv[s];
        const double d = r * hinv;
        qb.r[2 * i + s] = r;
        qb.p[2 * i + s] = d;
        acc[s] = b * b;
        acc[2 + s] = r * d;
      }
    }
  }
}

__global__ void __launch_bounds__(256) k_qeq_build(State st, Globals g, Tables t, QeqBuffers qb, int fuse) {
  __shared__ int ring_j[kQbWaves][kQbRing];
  __shared__ int ring_k[kQbWaves][kQbRing];   // key | neighbour type << 16
  __shared__ double ring_r2[kQbWaves][kQbRing];
  __shared__ double sh[4 * kQbWaves];
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int wv = threadIdx.x >> 6;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (i < st.n) qeq_build_row(st, g, t, qb, i, fuse, ring_j[wv], ring_k[wv], ring_r2[wv], acc);
  if (fuse) {   // the block's partials of b.b and r.p, its rows in wave order (a fixed order)
    if (lane_id() == 0)
      for (int k = 0; k < 4; ++k) sh[k * kQbWaves + wv] = acc[k];
    __syncthreads();
    if (threadIdx.x == 0)
      for (int k = 0; k < 4; ++k) {
        double v = sh[k * kQbWaves];
        for (int w = 1; w < kQbWaves; ++w) v += sh[k * kQbWaves + w];
        qb.partial_qb[blockIdx.x * 4 + k] = v;
      }
  }
}

// Census of the non-bonded cut-off decision: per atom, the pairs its skin list holds at the counted (canonical) end within
// nonb_cut. The test is the one every non-bonded kernel takes (the FP64 square root of the canonical squared distance <=
// nonb_cut, in all precisions), on the same list entries, so the count is the number of pairs those kernels evaluate there.
__global__ void __launch_bounds__(256) k_census_nb(State st, Globals g) {
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= st.n) return;
  const int lane = lane_id();
  const int ti = st.type[i];
  int c = 0;
  if (ti >= 0) {
    const size_t base = (size_t)i * st.maxnb;
    const int cnt = min(st.nb_cnt[i], st.maxnb);
    const double4 pi = st.pos[i];
    for (int k = lane; k < cnt; k += 64) {
      const int e = st.nb_s[base + k];
      if (!(e & kCanon)) continue;
      double dx, dy, dz;
      int tj;
      const double r2 = pair_dvec_list(st, pi, st.nb_j[base + k], e, dx, dy, dz, tj);
      if (tj >= 0 && sqrt(r2) <= g.nonb_cut) ++c;
    }
  }
  c = wave_sum_int(c);
  if (lane == 0) st.cen[kCenNonbonded * st.n + i] = c;
}

inline int nblk(long long items, int per) { return (int)((items + per - 1) / per); }

}  // namespace

void launch_census_nb(const State& st, const Globals& g, hipStream_t s) {
  hipLaunchKernelGGL(k_census_nb, dim3(nblk((long long)st.n * 64, 256)), dim3(256), 0, s, st, g);
}

void launch_nonbonded(Precision p, const State& st, const Cell& c, const Globals& g, const Tables& t, hipStream_t s) {
  const dim3 grid(nblk((long long)st.n * 64, 256)), blk(256);
  (void)c;
  if (p == Precision::FP64) hipLaunchKernelGGL(k_nonbonded_d, grid, blk, 0, s, st, g, t);
  else hipLaunchKernelGGL(k_nonbonded_f, grid, blk, 0, s, st, g, t);
}

void launch_nonbonded_rows(Precision p, const State& st, const Globals& g, const Tables& t, const QeqBuffers& qb, hipStream_t s) {
  const dim3 grid(nblk((long long)st.n * 64, 256)), blk(256);
  if (p == Precision::FP64) {
    if (qb.narrow) hipLaunchKernelGGL(k_nonbonded_rows_d<unsigned short>, grid, blk, 0, s, st, g, t, qb.col16, qb.key, qb.cnt);
    else hipLaunchKernelGGL(k_nonbonded_rows_d<int>, grid, blk, 0, s, st, g, t, qb.col, qb.key, qb.cnt);
  } else {
    if (qb.narrow) hipLaunchKernelGGL(k_nonbonded_rows<unsigned short>, grid, blk, 0, s, st, g, t, qb.col16, qb.key, qb.cnt);
    else hipLaunchKernelGGL(k_nonbonded_rows<int>, grid, blk, 0, s, st, g, t, qb.col, qb.key, qb.cnt);
  }
}

void launch_qeq_build(const State& st, const Cell& c, const Globals& g, const Tables& t, QeqBuffers& qb, int fuse, hipStream_t s) {
  (void)c;
  hipLaunchKernelGGL(k_qeq_build, dim3(qeq_build_blocks(st.n)), dim3(256), 0, s, st, g, t, qb, fuse);
}

}  // namespace rxa
