// Bond-order chain and the ten bonded ReaxFF terms on the device.
//
// Pipeline per step (one launch each, in stream order):
//   bond_build  (wave per atom)   bond list from the skin list: d <= bond_cut and BO' >= bo_cut, uncorrected BO', dBO'/dx, Delta'
//                                 [reference src/cpu/bonded.cpp:109-162, terms.hpp:50-56]; H-bond partner lists [bonded.cpp:120-124]
//   bo_correct  (thread per atom) corrected BO and its 11 derivative coefficients, always from the pair's canonical end so both ends
//                                 hold bitwise-identical values [bonded.cpp:164-204, terms.hpp:66-164]; index of the mirror slot of
//                                 every bond and the step's largest bond count
//   atom_terms  (thread per atom) bond energy + triple-bond stabilisation, lone pair + C2, over/under coordination [bonded.cpp:219-318]
//   valence     (thread per atom) valence angle, penalty, 3-body conjugation at the centre [bonded.cpp:321-474]
//   torsion     (thread per atom) torsion + 4-body conjugation, evaluated at BOTH atoms of the central bond; each end keeps the force
//                                 and derivative contributions of its own half (its own atom, its outer partner, its outer bond) and
//                                 the end the reference counts (tag order) adds the energy and the central-bond terms [bonded.cpp:476-652]
//   hbond       (wave per atom)   donor hydrogens evaluate their H-bond triples once and leave the acceptor's force in the
//                                 candidate's slot; acceptors collect those slots through a mirror index [bonded.cpp:654-715]
//   cdelta, dbond, gather         the reference's Add_dBond_to_Forces [bonded.cpp:719-766] rewritten as gathers: every BO-derivative
//                                 contribution is a central pair force along a bond, so each atom sums its own force from its own
//                                 bond slots and the mirror slots of its partners.
// No atomics on any force or energy path: every array element has exactly one writer, and every sum has a fixed order, so results
// are bitwise repeatable. Contributions an atom makes to a bonded partner (angle/torsion/H-bond forces, CdDelta) go into the
// partner's slot of the writer's own bond list (fpx/fpy/fpz/cdp) and are collected by the partner through the mirror index.
//
// Precision: TB is the type of the bond-order chain (BO', corrections, Delta quantities and their threshold tests); TR the type of
// the term arithmetic. FP64: <double, double>; mixed: <double, float>; FP32: <float, float>. The terms' own discrete decisions
// (three- and four-body gates, the H-bond BO gate, the triple-bond and C2 gates, the SBO branch) are taken in TS, the type of the
// chain whose values they test: double in FP64 and mixed modes, so they match the FP64 engine outside its rounding, float in
// FP32 mode. Energies, CdDelta and forces are
// accumulated in double in every mode.
#include <type_traits>

#include "common.hpp"
#include "launch.hpp"

namespace rxa {
namespace {

template <class T> __device__ inline T rexp(T x);
template <> __device__ inline double rexp(double x) { return exp(x); }
template <> __device__ inline float rexp(float x) { return expf(x); }
template <class T> __device__ inline T rexpm1(T x);
template <> __device__ inline double rexpm1(double x) { return expm1(x); }
template <> __device__ inline float rexpm1(float x) { return expm1f(x); }
template <class T> __device__ inline T rpow(T x, T y);
template <> __device__ inline double rpow(double x, double y) { return pow(x, y); }
template <> __device__ inline float rpow(float x, float y) { return powf(x, y); }
template <class T> __device__ inline T rlog(T x);
template <> __device__ inline double rlog(double x) { return log(x); }
template <> __device__ inline float rlog(float x) { return logf(x); }
template <class T> __device__ inline T rsqrt_(T x);
template <> __device__ inline double rsqrt_(double x) { return sqrt(x); }
template <> __device__ inline float rsqrt_(float x) { return sqrtf(x); }
template <class T> __device__ inline T racos(T x);
template <> __device__ inline double racos(double x) { return acos(x); }
template <> __device__ inline float racos(float x) { return acosf(x); }
template <class T> __device__ inline T rsin(T x);
template <> __device__ inline double rsin(double x) { return sin(x); }
template <> __device__ inline float rsin(float x) { return sinf(x); }
template <class T> __device__ inline T rcos(T x);
template <> __device__ inline double rcos(double x) { return cos(x); }
template <> __device__ inline float rcos(float x) { return cosf(x); }
template <class T> __device__ inline T ratan2(T y, T x);
template <> __device__ inline double ratan2(double y, double x) { return atan2(y, x); }
template <> __device__ inline float ratan2(float y, float x) { return atan2f(y, x); }
// exp() capped from above (float: e^40, double: e^700); only saturating ratios use it (reference terms.hpp:30-34)
template <class T> __device__ inline T safe_exp(T x) {
  const T cap = sizeof(T) == 4 ? T(40) : T(700);
  return rexp<T>(x > cap ? cap : x);
}

// d/dD of the switching ratio (2 + a) / (1 + a + b), where a' = ka a and b' = kb b and inv = 1 / (1 + a + b): the f8, f9 and
// f11 functions of the valence, penalty and torsion terms. The quotient rule as usually written, [a' (1 + a + b) - (2 + a)(a' + b')]
// / (1 + a + b)^2, holds a^2 terms that cancel; in FP32 they overflow once a passes ~1e19 (exp(10 * 4.5) in f11 at two
// under-coordinated centres) and the derivative, which is ~1/a, comes out infinite. With the cancellation done by hand,
// [-ka a - 2 kb b + (ka - kb) a b] inv inv, and a, b from safe_exp, every intermediate stays finite.
template <class T> __device__ inline T ratio2_deriv(T a, T b, T ka, T kb, T inv) {
  return ((-ka * a - T(2) * kb * b + (ka - kb) * (a * b)) * inv) * inv;
}

template <class T> struct V3 { T x, y, z; };
template <class T> __device__ inline T dot3(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> __device__ inline V3<T> cross3(const V3<T>& a, const V3<T>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// angle between two bond vectors from the same centre, its sine, and d cos(theta) / d x of (end a, centre, end b) (reference
// bonded.cpp:58-75). FP64 follows the reference: theta = acos of the clamped cosine, sin(theta) of that. FP32 takes
// theta = atan2(|a x b|, a.b) and sin = |a x b| / (|a||b|) instead: acosf of a cosine within an ulp of -1 or 1 loses half the
// digits, and sinf(acosf(-1)) is -8.7e-8 (the float nearest pi lies above pi), a negative sine that flips the sign of the
// torsion's 1/tan guard. The cosine stays the clamped dot product in both.
// kTheta = false (the torsion, which needs only the cosine and the sine): FP64 leaves theta unset and takes
// sin = sqrt((1 - c)(1 + c)), the reference's sin(acos(c)) to within a few ulp without the two trigonometric calls (1 - c and
// 1 + c are exact where they are small, so the sine keeps its relative accuracy near 0 and pi); FP32 is unchanged.
template <class T, bool kTheta = true>
__device__ inline void angle_and_dcos(const V3<T>& a, T da, const V3<T>& b, T db, T& theta, T& cost, T& sint, V3<T>& dca, V3<T>& dcj,
                                      V3<T>& dcb) {
  cost = dot3(a, b) / (da * db);
  if (cost > T(1)) cost = T(1);
  if (cost < T(-1)) cost = T(-1);
  if (sizeof(T) == 8 && !kTheta) {
    theta = T(0);
    sint = rsqrt_<T>((T(1) - cost) * (T(1) + cost));
  } else if (sizeof(T) == 8) {
    theta = racos<T>(cost);
    sint = rsin<T>(theta);
  } else {
    const V3<T> c = cross3(a, b);
    const T sn = rsqrt_<T>(dot3(c, c));
    theta = ratan2<T>(sn, dot3(a, b));
    sint = sn / (da * db);
  }
  const T da2 = da * da, db2 = db * db, inv = T(1) / (da * db), inv3 = inv * inv * inv, cd = dot3(a, b) * inv3;
  dca = {b.x * inv - cd * db2 * a.x, b.y * inv - cd * db2 * a.y, b.z * inv - cd * db2 * a.z};
  dcj = {-(b.x + a.x) * inv + cd * (db2 * a.x + da2 * b.x), -(b.y + a.y) * inv + cd * (db2 * a.y + da2 * b.y),
         -(b.z + a.z) * inv + cd * (db2 * a.z + da2 * b.z)};
  dcb = {a.x * inv - cd * da2 * b.x, a.y * inv - cd * da2 * b.y, a.z * inv - cd * da2 * b.z};
}

// ================================================================================================================
// bond list + uncorrected bond orders
// ================================================================================================================
// One wave per atom walks the skin list in chunks of 64. The distance pass runs on every lane; entries within bond_cut are
// compacted (ballot order = skin-list order) into a per-wave LDS ring and the FP64 BO' evaluation runs on full waves of 64
// candidates, so the transcendental work is not repeated for the ~90% of skin entries that are out of bond range.
constexpr int kBuildWaves = 4;
constexpr int kRing = 128;

template <class TB>
__device__ inline void bo_prime_batch(const State& st, const Globals& g, const Tables& t, int i, int ti, int m, const int* rj,
                                      const int* rs, const double* rd, const double* rx, const double* ry, const double* rz,
                                      int head, int& nb, TB& sum_bo, TB& ddx, TB& ddy, TB& ddz) {
  const int lane = lane_id(), n = st.n;
  const TB bo_cut = (TB)g.bo_cut;
  bool bkeep = false;
  int j = 0, s = 0;
  double d = 0, dx = 0, dy = 0, dz = 0;
  TB bs = 0, bp = 0, bpp = 0, cln = 0, clp = 0, clpp = 0;
  if (lane < m) {
    const int e = (head + lane) & (kRing - 1);
    j = rj[e]; s = rs[e]; d = rd[e]; dx = rx[e]; dy = ry[e]; dz = rz[e];
    const DevPair& P = t.pair[ti * g.ntypes + st.type[j]];
    const TB dd = (TB)d;
    TB C12 = 0, C34 = 0, C56 = 0;
    if constexpr (std::is_same<TB, double>::value) {
      // (d / r)^p as exp(p (log d - log r)): one logarithm for the three kinds (log r from the pair table) instead of three
      // general powers; the exponent's rounding error is a few 1e-16, so BO' moves by about 1e-15 relative
      const double ld = log(dd);
      if (P.s_ok) { C12 = P.p_bo1 * exp(P.p_bo2 * (ld - P.ln_rs)); bs = (1.0 + bo_cut) * exp(C12); }
      if (P.pi_ok) { C34 = P.p_bo3 * exp(P.p_bo4 * (ld - P.ln_rp)); bp = exp(C34); }
      if (P.pp_ok) { C56 = P.p_bo5 * exp(P.p_bo6 * (ld - P.ln_rpp)); bpp = exp(C56); }
    } else {
      if (P.s_ok) { C12 = (TB)P.p_bo1 * rpow<TB>(dd / (TB)P.r_s, (TB)P.p_bo2); bs = (TB(1) + bo_cut) * rexp<TB>(C12); }
      if (P.pi_ok) { C34 = (TB)P.p_bo3 * rpow<TB>(dd / (TB)P.r_p, (TB)P.p_bo4); bp = rexp<TB>(C34); }
      if (P.pp_ok) { C56 = (TB)P.p_bo5 * rpow<TB>(dd / (TB)P.r_pp, (TB)P.p_bo6); bpp = rexp<TB>(C56); }
    }
    bkeep = bs + bp + bpp >= bo_cut;
    const TB rr2 = TB(1) / (dd * dd);
    cln = (TB)P.p_bo2 * C12 * rr2;
    clp = (TB)P.p_bo4 * C34 * rr2;
    clpp = (TB)P.p_bo6 * C56 * rr2;
  }
  const unsigned long long mb = __ballot(bkeep);
  if (bkeep) {
    const int p = nb + lanes_below(mb);
    if (p < st.maxb) {
      const size_t q = sl(p, i, n);
      st.b_j[q] = j; st.b_s[q] = s;
      st.b_d[q] = d; st.b_dx[q] = dx; st.b_dy[q] = dy; st.b_dz[q] = dz;
      const TB tot = (bs + bp + bpp) - bo_cut;
      st.bop_s[q] = (double)(bs - bo_cut);
      st.bop_pi[q] = (double)bp;
      st.bop_pi2[q] = (double)bpp;
      st.bo[q] = (double)tot;   // uncorrected total until bo_correct overwrites it
      const TB dc = -(bs * cln + bp * clp + bpp * clpp);
      st.dbop_c[q] = (double)dc;
      st.dlnpi_c[q] = (double)(-bp * clp);
      st.dlnpi2_c[q] = (double)(-bpp * clpp);
      sum_bo += tot;
      ddx += dc * (TB)dx; ddy += dc * (TB)dy; ddz += dc * (TB)dz;
    }
  }
  nb += __popcll(mb);
}

// 4 waves per SIMD (108 VGPRs without spills; 129 and 3 waves otherwise)
template <class TB>
__global__ void __attribute__((amdgpu_waves_per_eu(4))) __launch_bounds__(64 * kBuildWaves) k_bond_build(State st, Cell c, Globals g, Tables t) {
  __shared__ int rj_[kBuildWaves][kRing], rs_[kBuildWaves][kRing];
  __shared__ double rd_[kBuildWaves][kRing], rx_[kBuildWaves][kRing], ry_[kBuildWaves][kRing], rz_[kBuildWaves][kRing];
  const int w = threadIdx.x >> 6;
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= st.n) return;
  int* rj = rj_[w]; int* rs = rs_[w];
  double *rd = rd_[w], *rx = rx_[w], *ry = ry_[w], *rz = rz_[w];
  const int lane = lane_id();
  const int ti = st.type[i];
  int nb = 0;
  TB sum_bo = 0, ddx = 0, ddy = 0, ddz = 0;
  if (ti >= 0) {
    const double4 pi = st.pos[i];
    const size_t base = (size_t)i * st.maxbc;
    const int cnt = min(st.bc_cnt[i], st.maxbc);
    int head = 0, fill = 0;   // LDS ring of bond candidates (wave-uniform)
    for (int k0 = 0; k0 < cnt; k0 += 64) {
      const int k = k0 + lane;
      bool near = false;
      int j = 0, s = 0;
      double dx = 0, dy = 0, dz = 0, d = 0;
      if (k < cnt) {
        j = st.bc_j[base + k];
        const int e = st.bc_s[base + k];
        s = entry_code(e);
        d = sqrt(pair_dvec_list(st, pi, j, e, dx, dy, dz));
        near = d <= g.bond_cut;
      }
      const unsigned long long mn = __ballot(near);
      if (near) {
        const int e = (head + fill + lanes_below(mn)) & (kRing - 1);
        rj[e] = j; rs[e] = s; rd[e] = d; rx[e] = dx; ry[e] = dy; rz[e] = dz;
      }
      fill += __popcll(mn);
      __builtin_amdgcn_wave_barrier();
      if (fill >= 64) {
        bo_prime_batch<TB>(st, g, t, i, ti, 64, rj, rs, rd, rx, ry, rz, head, nb, sum_bo, ddx, ddy, ddz);
        head = (head + 64) & (kRing - 1);
        fill -= 64;
        __builtin_amdgcn_wave_barrier();
      }
    }
    if (fill > 0) bo_prime_batch<TB>(st, g, t, i, ti, fill, rj, rs, rd, rx, ry, rz, head, nb, sum_bo, ddx, ddy, ddz);
  }
  // fixed-order wave reductions
  for (int off = 32; off > 0; off >>= 1) {
    sum_bo += __shfl_xor(sum_bo, off, 64);
    ddx += __shfl_xor(ddx, off, 64);
    ddy += __shfl_xor(ddy, off, 64);
    ddz += __shfl_xor(ddz, off, 64);
  }
  if (lane == 0) {
    st.b_cnt[i] = nb;
    if (nb > st.maxb) atomicMax(&st.flags[1], nb);
    const double val = ti >= 0 ? t.sp[ti].valency : 0.0, valv = ti >= 0 ? t.sp[ti].valency_val : 0.0;
    st.Dp[i] = (double)(sum_bo - (TB)val);
    st.Dpb[i] = (double)(sum_bo - (TB)valv);
    st.dDps_x[i] = (double)ddx; st.dDps_y[i] = (double)ddy; st.dDps_z[i] = (double)ddz;
  }
}

// ---- atom groups: G lanes per atom (G a power of two >= the largest bond count), lane p owns bond slot p of its atom ----------
// Sums over a group use xor shuffles with offsets below G, so they stay inside the group and run in a fixed order.
template <int G>
__device__ inline int gsum_int(int v) {
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
template <int G>
__device__ inline double gsum(double v) {
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline int wave_max(int v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

// ================================================================================================================
// corrected bond orders (always from the canonical end) and atom quantities
// ================================================================================================================
// Also finds each slot's mirror on the other atom (b_sym) and the step's largest bond count (flags[7]); a slot without a mirror
// sets flags[3]. L lanes per atom, lane p taking slots p, p + L, ...: the slots are independent (the correction reads the
// uncorrected sums of both ends), so a lane's chain of dependent loads is one slot's instead of the atom's; the total bond order
// is summed in slot order through the group (the order of a serial loop, so the sums are bitwise those of one thread per atom).

// the mirror slot, the corrected bond orders and the derivative coefficients of slot p of atom i; returns the corrected BO
template <class TB>
__device__ inline TB bo_correct_slot(const State& st, const Globals& g, const Tables& t, int i, int p) {
  const int n = st.n;
  const size_t q = sl(p, i, n);
  const int j = st.b_j[q];
  {   // mirror slot of this bond on atom j
    const int want = shift_neg(st.b_s[q]), nbj = min(st.b_cnt[j], st.maxb);
    int found = -1;
    for (int r = 0; r < nbj; ++r)
      if (st.b_j[sl(r, j, n)] == i && st.b_s[sl(r, j, n)] == want) { found = r; break; }
    st.b_sym[q] = found < 0 ? 0 : found;
    if (found < 0) st.flags[3] = 1;
  }
  const bool first = canonical_first(i, j, st.b_s[q], st.tag);
  const int a = first ? i : j, b = first ? j : i;
  const int ta = st.type[a], tb = st.type[b];
  const DevPair& P = t.pair[ta * g.ntypes + tb];
  const TB vi = (TB)t.sp[ta].valency, vj = (TB)t.sp[tb].valency;
  const TB Dpi = (TB)st.Dp[a], Dpj = (TB)st.Dp[b], Dbi = (TB)st.Dpb[a], Dbj = (TB)st.Dpb[b];
  const TB BO = (TB)st.bo[q], BOs = (TB)st.bop_s[q], BOpi = (TB)st.bop_pi[q], BOpi2 = (TB)st.bop_pi2[q];
  TB oBO, oBOs, oBOpi, oBOpi2, C1dbo = 0, C2dbo = 0, C3dbo = 0, C1pi = 0, C2pi = 0, C3pi = 0, C4pi = 0, C1pi2 = 0, C2pi2 = 0, C3pi2 = 0, C4pi2 = 0;
  if ((TB)P.ovc < TB(0.001) && (TB)P.v13cor < TB(0.001)) {
    oBO = BO; oBOs = BOs; oBOpi = BOpi; oBOpi2 = BOpi2;
    C1dbo = 1; C1pi = 1; C1pi2 = 1;
  } else {
    const TB boc1 = (TB)g.p_boc1, boc2 = (TB)g.p_boc2;
    TB f1 = 1, Cf1ij = 0, Cf1ji = 0;
    if ((TB)P.ovc >= TB(0.001)) {
      const TB e1i = safe_exp<TB>(-boc1 * Dpi), e2i = safe_exp<TB>(-boc2 * Dpi), e1j = safe_exp<TB>(-boc1 * Dpj), e2j = safe_exp<TB>(-boc2 * Dpj);
      const TB f2 = e1i + e1j, f3 = -TB(1) / boc2 * rlog<TB>(TB(0.5) * (e2i + e2j));
      f1 = TB(0.5) * ((vi + f2) / (vi + f2 + f3) + (vj + f2) / (vj + f2 + f3));
      const TB ui = vi + (f2 + f3), uj = vj + (f2 + f3), S2 = e2i + e2j;
      Cf1ij = TB(0.5) * (-boc1 * e1i / ui - ((vi + f2) / (ui * ui)) * (-boc1 * e1i + e2i / S2) + -boc1 * e1i / uj -
                         ((vj + f2) / (uj * uj)) * (-boc1 * e1i + e2i / S2));
      const TB A = TB(0.5) * f3 * (TB(1) / (ui * ui) + TB(1) / (uj * uj));
      const TB B = -TB(0.5) * ((ui - f3) / (ui * ui) + (uj - f3) / (uj * uj));
      Cf1ji = -A * boc1 * e1j + B * e2j / S2;
    }
    TB f4 = 1, f5 = 1, Cf45ij = 0, Cf45ji = 0;
    const TB boc3 = (TB)P.p_boc3, boc4 = (TB)P.p_boc4, boc5 = (TB)P.p_boc5;
    if ((TB)P.v13cor >= TB(0.001)) {
      const TB ef4 = safe_exp<TB>(-(boc4 * (BO * BO) - Dbi) * boc3 + boc5);
      const TB ef5 = safe_exp<TB>(-(boc4 * (BO * BO) - Dbj) * boc3 + boc5);
      f4 = TB(1) / (TB(1) + ef4);
      f5 = TB(1) / (TB(1) + ef5);
      Cf45ij = -f4 * ef4;
      Cf45ji = -f5 * ef5;
    }
    const TB A0 = f1 * (f4 * f5);
    const TB A1 = -2 * boc3 * boc4 * BO * (Cf45ij + Cf45ji);
    const TB A2ij = Cf1ij / f1 + boc3 * Cf45ij, A2ji = Cf1ji / f1 + boc3 * Cf45ji;
    const TB A3ij = A2ij + Cf1ij / f1, A3ji = A2ji + Cf1ji / f1;
    oBO = BO * A0;
    oBOpi = BOpi * A0 * f1;
    oBOpi2 = BOpi2 * A0 * f1;
    oBOs = oBO - (oBOpi + oBOpi2);
    C1dbo = A0 + oBO * A1; C2dbo = oBO * A2ij; C3dbo = oBO * A2ji;
    C1pi = f1 * f1 * f4 * f5; C2pi = oBOpi * A1; C3pi = oBOpi * A3ij; C4pi = oBOpi * A3ji;
    C1pi2 = f1 * f1 * f4 * f5; C2pi2 = oBOpi2 * A1; C3pi2 = oBOpi2 * A3ij; C4pi2 = oBOpi2 * A3ji;
  }
  if (oBO < TB(1e-10)) oBO = 0;
  if (oBOs < TB(1e-10)) oBOs = 0;
  if (oBOpi < TB(1e-10)) oBOpi = 0;
  if (oBOpi2 < TB(1e-10)) oBOpi2 = 0;
  st.bo[q] = (double)oBO; st.bo_s[q] = (double)oBOs; st.bo_pi[q] = (double)oBOpi; st.bo_pi2[q] = (double)oBOpi2;
  st.C[0][q] = (double)C1dbo; st.C[1][q] = (double)C2pi; st.C[2][q] = (double)C2pi2; st.C[3][q] = (double)C1pi; st.C[4][q] = (double)C1pi2;
  st.C[5][q] = (double)(first ? C2dbo : C3dbo);
  st.C[6][q] = (double)(first ? C3pi : C4pi);
  st.C[7][q] = (double)(first ? C3pi2 : C4pi2);
  st.Cdbo[q] = 0; st.Cdbopi[q] = 0; st.Cdbopi2[q] = 0; st.fpx[q] = 0; st.fpy[q] = 0; st.fpz[q] = 0; st.cdp[q] = 0;
  return oBO;
}

template <class TB, int L>
__global__ void __launch_bounds__(256) k_bo_correct(State st, Globals g, Tables t) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gid / L, p0 = gid % L, n = st.n;
  const bool live = i < n;
  const int nb = live ? min(st.b_cnt[i], st.maxb) : 0;
  {   // the largest bond count of this step (flags[7]), one atomic per wave
    int c = live && p0 == 0 ? st.b_cnt[i] : 0;
    for (int off = 32; off > 0; off >>= 1) c = max(c, __shfl_xor(c, off, 64));
    if (lane_id() == 0 && c > 0) atomicMax(&st.flags[7], c);
  }
  const int nbw = wave_max(nb);
  const int base = lane_id() - p0;   // the group's first lane
  TB tbo = 0;
  for (int k0 = 0; k0 < nbw; k0 += L) {   // wave-uniform rounds of L slots
    const int p = k0 + p0;
    const TB o = p < nb ? bo_correct_slot<TB>(st, g, t, i, p) : TB(0);
    for (int l = 0; l < L; ++l) tbo += __shfl(o, base + l, 64);   // slot order; absent slots add exact zeros
  }
  if (!live || p0 != 0) return;
  const int ti = st.type[i];
  st.tbo[i] = (double)tbo;
  st.CdD_self[i] = 0; st.fs_x[i] = 0; st.fs_y[i] = 0; st.fs_z[i] = 0;
  for (int k = 0; k < 6; ++k) st.vb[(size_t)k * n + i] = 0;
  if (ti < 0) return;
  const DevSpecies& s = t.sp[ti];
  // atom quantities (reference terms.hpp:141-164); trunc toward zero
  const TB Delta = tbo - (TB)s.valency, De = tbo - (TB)s.valency_e;
  const TB ht = (TB)(int)(De / TB(2));
  const TB vlpex = De - TB(2) * ht;
  const TB ex = rexp<TB>(-(TB)g.p_lp1 * ((TB(2) + vlpex) * (TB(2) + vlpex)));
  const TB nlp = ex - ht;
  const TB Dlp = (TB)s.nlp_opt - nlp;
  st.Delta[i] = (double)Delta;
  st.Delta_e[i] = (double)De;
  st.Delta_boc[i] = (double)(tbo - (TB)s.valency_boc);
  st.Delta_val[i] = (double)(tbo - (TB)s.valency_val);
  st.vlpex[i] = (double)vlpex;
  st.nlp[i] = (double)nlp;
  st.Delta_lp[i] = (double)Dlp;
  st.dDelta_lp[i] = (double)(TB(2) * (TB)g.p_lp1 * ex * (TB(2) + vlpex));
  st.Delta_lp_temp[i] = s.heavy ? (double)((TB)s.nlp_opt - TB(0.5) * ((TB)s.valency_e - (TB)s.valency)) : (double)Dlp;
}

// ================================================================================================================
// bond energy, triple-bond stabilisation, lone pair (+ C2), over/under coordination
// ================================================================================================================
// L lanes per atom, lane p taking slots p, p + L, ...; the per-atom sums run in slot order through the group (each lane of the
// group forms the same sums, in the order of a serial loop over the slots, so they are bitwise those of one thread per atom,
// whatever L), and the atom's own terms are evaluated by every lane of the group alike.
template <class TR, class TS, int L>
__global__ void __launch_bounds__(256) k_atom_terms(State st, Globals g, Tables t) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gid / L, p0 = gid % L, n = st.n;
  const bool live = i < n;
  const int ti = live ? st.type[i] : -1;
  const int nb = ti >= 0 ? min(st.b_cnt[i], st.maxb) : 0;
  const int nbw = wave_max(nb);
  const int base = lane_id() - p0;
  const size_t nsl = (size_t)st.maxb * n;
  if (live && p0 == 0 && st.peratom)   // the reference-split virial shares, accumulated by the angle kernels that follow
    for (int c = 0; c < 6; ++c) st.vref[(size_t)c * n + i] = 0.0;
  double e_bond = 0, e_lp = 0, e_ov = 0, e_un = 0, cdself = 0, e_c2 = 0;
  // bond energy of the bonds whose canonical end is this atom (the reference's tag order, bonded.cpp:222-252)
  for (int k0 = 0; k0 < nbw; k0 += L) {
    const int p = k0 + p0;
    double eb = 0, et = 0, cdt = 0;
    if (p < nb) {
      const size_t q = sl(p, i, n);
      if (st.peratom) {
        st.eslot[q] = 0.0;
        for (int c = 0; c < 6; ++c) st.vslot[c * nsl + q] = 0.0;
      }
      const int j = st.b_j[q];
      if (canonical_first(i, j, st.b_s[q], st.tag)) {
        const int tj = st.type[j];
        const DevPair& P = t.pair[ti * g.ntypes + tj];
        const TR BOs = (TR)st.bo_s[q], BOpi = (TR)st.bo_pi[q], BOpi2 = (TR)st.bo_pi2[q];
        const TR pw = BOs == TR(0) ? TR(0) : rpow<TR>(BOs, (TR)P.p_be2);
        const TR ex = rexp<TR>((TR)P.p_be1 * (TR(1) - pw));
        const TR CEbo = -(TR)P.De_s * ex * (TR(1) - (TR)P.p_be1 * (TR)P.p_be2 * pw);
        eb = (double)(-(TR)P.De_s * BOs * ex - (TR)P.De_p * BOpi - (TR)P.De_pp * BOpi2);
        if (st.peratom) st.eslot[q] += 0.5 * eb;
        double cdbo = (double)CEbo;
        st.Cdbopi[q] += -(double)(CEbo + (TR)P.De_p);
        st.Cdbopi2[q] += -(double)(CEbo + (TR)P.De_pp);
        const TR BO = (TR)st.bo[q];
        if ((TS)st.bo[q] >= TS(1.00) && P.tbs) {
          const TR gp3 = (TR)g.gp3, gp4 = (TR)g.gp4, gp7 = (TR)g.gp7, gp10 = (TR)g.gp10;
          const TR exphu = rexp<TR>(-gp7 * ((BO - TR(2.50)) * (BO - TR(2.50))));
          const TR ea = rexp<TR>(-gp3 * ((TR)st.tbo[i] - BO)), ebb = rexp<TR>(-gp3 * ((TR)st.tbo[j] - BO));
          const TR eov = safe_exp<TR>(gp4 * ((TR)st.Delta[i] + (TR)st.Delta[j]));
          const TR hov = TR(1) / (TR(1) + TR(25.0) * eov);
          et = (double)(gp10 * exphu * hov * (ea + ebb));
          if (st.peratom) st.eslot[q] += 0.5 * et;
          cdbo += (double)(gp10 * exphu * hov * (ea + ebb) * (gp3 - TR(2.0) * gp7 * (BO - TR(2.50))));
          cdt = (double)(-gp10 * exphu * hov * (gp3 * ea + TR(25.0) * gp4 * eov * hov * (ea + ebb)));
          st.cdp[q] += (double)(-gp10 * exphu * hov * (gp3 * ebb + TR(25.0) * gp4 * eov * hov * (ea + ebb)));
        }
        st.Cdbo[q] += cdbo;
      }
    }
    for (int l = 0; l < L; ++l) {   // slot order; slots without these terms add exact zeros
      e_bond += __shfl(eb, base + l, 64);
      e_bond += __shfl(et, base + l, 64);
      cdself += __shfl(cdt, base + l, 64);
    }
  }
  const DevSpecies& si = t.sp[ti >= 0 ? ti : 0];
  // lone pair
  if (ti >= 0 && (nb > 0 || g.enobonds)) {
    const TR Dlp = (TR)st.Delta_lp[i], p_lp2 = (TR)si.p_lp2;
    const TR e2 = safe_exp<TR>(TR(-75) * Dlp), inv = TR(1) / (TR(1) + e2);
    e_lp += (double)(p_lp2 * Dlp * inv);
    cdself += (double)((p_lp2 * inv + TR(75) * p_lp2 * Dlp * e2 * (inv * inv)) * (TR)st.dDelta_lp[i]);
  }
  // C2 correction, every directed C-C bond (reference bonded.cpp:268-280)
  const bool c2 = ti >= 0 && g.c2_active && si.c2;
  if (__any(c2)) {
    const TR p_lp3 = (TR)g.p_lp3, Di = (TR)(c2 ? st.Delta[i] : 0.0);
    for (int k0 = 0; k0 < nbw; k0 += L) {
      const int p = k0 + p0;
      double ec = 0, cdc = 0;
      if (c2 && p < nb) {
        const size_t q = sl(p, i, n);
        if (t.sp[st.type[st.b_j[q]]].c2) {
          const TR vov3 = (TR)st.bo[q] - Di - TR(0.040) * rpow<TR>(Di, TR(4.));
          const TS Ds = (TS)st.Delta[i];
          if ((TS)st.bo[q] - Ds - TS(0.040) * rpow<TS>(Ds, TS(4.)) > TS(3.)) {
            ec = (double)(p_lp3 * ((vov3 - TR(3.0)) * (vov3 - TR(3.0))));
            if (st.peratom) st.eslot[q] += 0.5 * ec;
            st.Cdbo[q] += (double)(TR(2.) * p_lp3 * (vov3 - TR(3.)));
            cdc = (double)(TR(2.) * p_lp3 * (vov3 - TR(3.)) * (-TR(1.) - TR(0.16) * rpow<TR>(Di, TR(3.))));
          }
        }
      }
      for (int l = 0; l < L; ++l) {
        const double e = __shfl(ec, base + l, 64);
        e_lp += e;
        e_c2 += e;
        cdself += __shfl(cdc, base + l, 64);
      }
    }
  }
  // over / under coordination (reference bonded.cpp:282-318, terms.hpp:231-260)
  {
    const TR dfvl = si.heavy ? TR(0) : TR(1);
    TR sum1 = 0, sum2 = 0;
    for (int k0 = 0; k0 < nbw; k0 += L) {
      const int p = k0 + p0;
      TR a1 = 0, a2 = 0;
      if (p < nb) {
        const size_t q = sl(p, i, n);
        const int j = st.b_j[q];
        const DevPair& P = t.pair[ti * g.ntypes + st.type[j]];
        a1 = (TR)P.p_ovun1 * (TR)P.De_s * (TR)st.bo[q];
        a2 = ((TR)st.Delta[j] - dfvl * (TR)st.Delta_lp_temp[j]) * ((TR)st.bo_pi[q] + (TR)st.bo_pi2[q]);
      }
      for (int l = 0; l < L; ++l) {
        sum1 += __shfl(a1, base + l, 64);
        sum2 += __shfl(a2, base + l, 64);
      }
    }
    if (ti >= 0) {
      const bool under = nb > 0 || g.enobonds;
      const TR ovun3 = (TR)g.p_ovun3, ovun4 = (TR)g.p_ovun4, ovun6 = (TR)g.p_ovun6, ovun7 = (TR)g.p_ovun7, ovun8 = (TR)g.p_ovun8;
      const TR p_ovun2 = (TR)si.p_ovun2, p_ovun5 = (TR)si.p_ovun5;
      const TR Dlpt = (TR)st.Delta_lp_temp[i], dDlp = (TR)st.dDelta_lp[i];
      const TR eo1 = ovun3 * safe_exp<TR>(ovun4 * sum2), ieo1 = TR(1) / (TR(1) + eo1);
      const TR Dlc = (TR)st.Delta[i] - (dfvl * Dlpt) * ieo1;
      const TR eo2 = safe_exp<TR>(p_ovun2 * Dlc), ieo2 = TR(1) / (TR(1) + eo2);
      const TR DlpVi = TR(1) / (Dlc + (TR)si.valency + TR(1e-8));
      const TR CEover1 = Dlc * DlpVi * ieo2;
      e_ov += (double)(sum1 * CEover1);
      const TR CEover2 = sum1 * DlpVi * ieo2 * (TR(1) - Dlc * (DlpVi + p_ovun2 * eo2 * ieo2));
      const TR CEover3 = CEover2 * (TR(1) - dfvl * dDlp * ieo1);
      const TR CEover4 = CEover2 * (dfvl * Dlpt) * ovun4 * eo1 * (ieo1 * ieo1);
      const TR eo2n = safe_exp<TR>(-p_ovun2 * Dlc), eo6 = safe_exp<TR>(ovun6 * Dlc), eo8 = ovun7 * safe_exp<TR>(ovun8 * sum2);
      const TR ieo2n = TR(1) / (TR(1) + eo2n), ieo8 = TR(1) / (TR(1) + eo8);
      const TR eun = under ? -p_ovun5 * (TR(1) - eo6) * ieo2n * ieo8 : TR(0);
      const TR CEunder1 = ieo2n * (p_ovun5 * ovun6 * eo6 * ieo8 + p_ovun2 * eun * eo2n);
      const TR CEunder2 = -eun * ovun8 * eo8 * ieo8;
      const TR CEunder3 = CEunder1 * (TR(1) - dfvl * dDlp * ieo1);
      const TR CEunder4 = CEunder1 * (dfvl * Dlpt) * ovun4 * eo1 * (ieo1 * ieo1) + CEunder2;
      if (under) e_un += (double)eun;
      cdself += (double)CEover3;
      if (under) cdself += (double)CEunder3;
      for (int p = p0; p < nb; p += L) {
        const size_t q = sl(p, i, n);
        const int j = st.b_j[q];
        const DevPair& P = t.pair[ti * g.ntypes + st.type[j]];
        st.Cdbo[q] += (double)(CEover1 * (TR)P.p_ovun1 * (TR)P.De_s);
        const TR pis = (TR)st.bo_pi[q] + (TR)st.bo_pi2[q];
        const TR dj = (TR)st.Delta[j] - dfvl * (TR)st.Delta_lp_temp[j];
        const TR w = TR(1) - dfvl * (TR)st.dDelta_lp[j];
        st.cdp[q] += (double)(CEover4 * w * pis) + (double)(CEunder4 * w * pis);
        st.Cdbopi[q] += (double)(CEover4 * dj) + (double)(CEunder4 * dj);
        st.Cdbopi2[q] += (double)(CEover4 * dj) + (double)(CEunder4 * dj);
      }
    }
  }
  if (!live || p0 != 0) return;
  st.CdD_self[i] += cdself;
  st.eatom[T_BOND * n + i] = e_bond;
  st.eatom[T_LP * n + i] = e_lp;
  st.eatom[T_OV * n + i] = e_ov;
  st.eatom[T_UN * n + i] = e_un;
  if (st.peratom) st.eown[i] = 0.5 * e_bond + (e_lp - 0.5 * e_c2) + e_ov + e_un;   // halves of the pair-split terms stay here
}

// ================================================================================================================
// valence angle, penalty, 3-body conjugation (centre j)
// ================================================================================================================

// Valence angle, penalty and three-body conjugation at the centre j [reference bonded.cpp:321-474]. Lane p of j's group evaluates
// the pairs (p, k) with p < k of j's bonds, k running over a wave-uniform loop: the contributions to slot p stay in the lane, the
// contributions to slot k are summed over the group and kept by lane k, and the centre's are summed over the group at the end.
// PA: also the reference's per-atom virial split (bonded.cpp:462-466): each angle's (x_i - x_j) (x) F_i + (x_k - x_j) (x) F_k,
// a third to each of i, j and k (the ends through their slots).
template <class TR, class TS, int G, bool PA>
__device__ inline void valence_body(const State& st, const Globals& g, const Tables& t) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = gid / G, p = gid % G, n = st.n;
  const bool live = j < n;
  const int tj = live ? st.type[j] : -1;
  const int nb = live ? min(st.b_cnt[j], st.maxb) : 0;
  const int nbw = wave_max(nb);
  const bool act = tj >= 0 && nb >= 2;
  // the census also wants the SBO branch of atoms with fewer than two bonds (no angle of theirs is evaluated)
  const bool sbo_on = act || (st.census && tj >= 0);
  int n_ang = 0, sbo_region = 0;   // census: bond pairs through the three-body gates, SBO branch
  double e_ang = 0, e_pen = 0, e_coa = 0, cdself = 0, fjx = 0, fjy = 0, fjz = 0, sCE5 = 0, sCE6 = 0;
  double own_cdbo = 0, own_cdp = 0, own_fx = 0, own_fy = 0, own_fz = 0;
  // centre quantities (every lane of the group computes them)
  TR SBO2 = 0, CSBO2 = 0, dSBO1 = 0, dSBO2 = 0, ev6 = 0, f9 = 0, Cf9 = 0, ec2 = 0, Dboc = 0, p_val3 = 0, p_val5 = 0;
  const TR thb_cut = (TR)g.thb_cut;
  const TS thb_cut_s = (TS)g.thb_cut, thb_cutsq_s = (TS)g.thb_cutsq;
  const TR p_val6 = (TR)g.p_val6, p_val10 = (TR)g.p_val10;
  const TR p_pen2 = (TR)g.p_pen2, p_pen3 = (TR)g.p_pen3, p_pen4 = (TR)g.p_pen4;
  const TR p_coa2 = (TR)g.p_coa2, p_coa3 = (TR)g.p_coa3, p_coa4 = (TR)g.p_coa4;
  if (sbo_on) {
    const DevSpecies& sj = t.sp[tj];
    p_val3 = (TR)sj.p_val3; p_val5 = (TR)sj.p_val5;
    // SBO and its branch in TS. omp = 1 - prod(exp(-BO^8)): in double the reference's product; in float from expm1 of the
    // exponent sum, since a bond with BO^8 below FP32's epsilon leaves the float product at exactly 1 and would drop out of
    // 1 - prod, which decides the sign of SBO at under-coordinated centres
    TS SBOp = 0, prod = 1, omp;
    {
      TS sw = 0;
      for (int b = 0; b < nb; ++b) {
        const size_t q = sl(b, j, n);
        const TS bo = (TS)st.bo[q];
        SBOp += (TS)st.bo_pi[q] + (TS)st.bo_pi2[q];
        TS w = bo * bo;
        w *= w;
        w *= w;
        if (sizeof(TS) == 8) prod *= rexp<TS>(-w); else sw += w;
      }
      if (sizeof(TS) == 8) omp = TS(1) - prod;
      else { omp = -rexpm1<TS>(-sw); prod = TS(1) - omp; }
    }
    const TS p_val8s = (TS)g.p_val8, dDlp = (TS)st.dDelta_lp[j];
    TS vlpadj, dSBO2s;
    if ((TS)st.vlpex[j] >= TS(0)) { vlpadj = 0; dSBO2s = -omp; }
    else { vlpadj = (TS)st.nlp[j]; dSBO2s = -omp * (TS(1) - p_val8s * dDlp); }
    const TS Dbocs = (TS)st.Delta_boc[j];
    Dboc = (TR)Dbocs;
    const TS SBO = SBOp + omp * (-Dbocs - p_val8s * vlpadj);
    dSBO1 = (TR)(-TS(8) * prod * (Dbocs + p_val8s * vlpadj));
    dSBO2 = (TR)dSBO2s;
    const TS pv9 = (TS)g.p_val9;
    if (SBO <= TS(0)) { SBO2 = 0; CSBO2 = 0; sbo_region = 0; }
    else if (SBO <= TS(1)) { SBO2 = (TR)rpow<TS>(SBO, pv9); CSBO2 = (TR)(pv9 * rpow<TS>(SBO, pv9 - TS(1))); sbo_region = 1; }
    else if (SBO < TS(2)) { SBO2 = (TR)(TS(2) - rpow<TS>(TS(2) - SBO, pv9)); CSBO2 = (TR)(pv9 * rpow<TS>(TS(2) - SBO, pv9 - TS(1))); sbo_region = 2; }
    else { SBO2 = 2; CSBO2 = 0; sbo_region = 3; }
    ev6 = safe_exp<TR>(p_val6 * Dboc);
    const TR Delta = (TR)st.Delta[j], Dval = (TR)st.Delta_val[j];
    const TR ep3 = safe_exp<TR>(-p_pen3 * Delta), ep4 = safe_exp<TR>(p_pen4 * Delta), itrm = TR(1) / (TR(1) + ep3 + ep4);
    f9 = (TR(2) + ep3) * itrm;
    Cf9 = ratio2_deriv<TR>(ep3, ep4, -p_pen3, p_pen4, itrm);
    ec2 = safe_exp<TR>(p_coa2 * Dval);
  }
  // this lane's bond (j, i)
  const bool mine = act && p < nb;
  size_t qi = 0;
  int i = 0, ti = 0;
  TR BOij = 0, BOA_ij = 0, dij = 1, tboi = 0;
  TS BOij_s = 0;
  V3<TR> vij{0, 0, 0};
  if (mine) {
    qi = sl(p, j, n);
    BOij_s = (TS)st.bo[qi];
    BOij = (TR)st.bo[qi];
    BOA_ij = BOij - thb_cut;
    i = st.b_j[qi]; ti = st.type[i];
    vij = V3<TR>{(TR)st.b_dx[qi], (TR)st.b_dy[qi], (TR)st.b_dz[qi]};
    dij = (TR)st.b_d[qi];
    tboi = (TR)st.tbo[i];
  }
  double vA[6] = {0, 0, 0, 0, 0, 0}, vC[6] = {0, 0, 0, 0, 0, 0};   // PA: slot p's partner, the centre (this lane's angles)
  for (int pk = 1; pk < nbw; ++pk) {
    double to_cdbo = 0, to_cdp = 0, to_fx = 0, to_fy = 0, to_fz = 0;
    double vto[6] = {0, 0, 0, 0, 0, 0};
    if (mine && p < pk && pk < nb && BOij_s - thb_cut_s > TS(0)) {
      const size_t qk = sl(pk, j, n);
      const TS BOjk_s = (TS)st.bo[qk];
      const TR BOjk = (TR)st.bo[qk];
      const TR BOA_jk = BOjk - thb_cut;
      const int k = st.b_j[qk], tk = st.type[k];
      const int2 ai = t.angle_idx[(ti * g.ntypes + tj) * g.ntypes + tk];
      const bool gate = BOjk_s - thb_cut_s > TS(0) && BOij_s > thb_cut_s && BOjk_s > thb_cut_s && BOij_s * BOjk_s > thb_cutsq_s;
      n_ang += gate ? 1 : 0;
      if (gate && ai.y != 0) {
        const V3<TR> vjk{(TR)st.b_dx[qk], (TR)st.b_dy[qk], (TR)st.b_dz[qk]};
        const TR djk = (TR)st.b_d[qk];
        TR th, ct, sth;
        V3<TR> dci, dcj, dck;
        angle_and_dcos<TR>(vij, dij, vjk, djk, th, ct, sth, dci, dcj, dck);
        if (sth < TR(1.0e-5)) sth = TR(1.0e-5);
        const TR tbok = (TR)st.tbo[k];
        for (int a = 0; a < ai.y; ++a) {
          const DevAngle A = t.angle[ai.x + a];
          const TR p_val1 = (TR)A.p_val1;
          if (!(fabs((double)p_val1) > 0.001)) continue;
          const TR p_val2 = (TR)A.p_val2, p_val4 = (TR)A.p_val4, p_val7 = (TR)A.p_val7, th00 = (TR)A.theta_00;
          const TR e3ij = rexp<TR>(-p_val3 * rpow<TR>(BOA_ij, p_val4)), f7ij = TR(1) - e3ij;
          const TR Cf7ij = p_val3 * p_val4 * rpow<TR>(BOA_ij, p_val4 - TR(1)) * e3ij;
          const TR e3jk = rexp<TR>(-p_val3 * rpow<TR>(BOA_jk, p_val4)), f7jk = TR(1) - e3jk;
          const TR Cf7jk = p_val3 * p_val4 * rpow<TR>(BOA_jk, p_val4 - TR(1)) * e3jk;
          const TR ev7 = safe_exp<TR>(-p_val7 * Dboc), itrm8 = TR(1) / (TR(1) + ev6 + ev7);
          const TR f8 = p_val5 - (p_val5 - TR(1)) * ((TR(2) + ev6) * itrm8);
          const TR Cf8 = (TR(1) - p_val5) * ratio2_deriv<TR>(ev6, ev7, p_val6, -p_val7, itrm8);
          const TR th0 = (TR(180.0) - th00 * (TR(1) - rexp<TR>(-p_val10 * (TR(2) - SBO2)))) * TR(3.14159265) / TR(180.0);
          const TR e2t = rexp<TR>(-p_val2 * ((th0 - th) * (th0 - th)));
          const TR e12t = p_val1 >= TR(0) ? p_val1 * (TR(1) - e2t) : p_val1 * -e2t;
          const TR CE1 = Cf7ij * f7jk * f8 * e12t, CE2 = Cf7jk * f7ij * f8 * e12t, CE3 = Cf8 * f7ij * f7jk * e12t;
          const TR CE4 = -TR(2) * p_val1 * p_val2 * f7ij * f7jk * f8 * e2t * (th0 - th);
          const TR Cth0 = p_val10 * (th00 * TR(3.14159265) / TR(180.0)) * rexp<TR>(-p_val10 * (TR(2) - SBO2));
          const TR CE5 = -CE4 * Cth0 * CSBO2, CE6 = CE5 * dSBO1, CE7 = CE5 * dSBO2, CE8 = -CE4 / sth;
          e_ang += (double)(f7ij * f7jk * f8 * e12t);
          const TR p_pen1 = (TR)A.p_pen1;
          const TR ep2ij = rexp<TR>(-p_pen2 * ((BOA_ij - TR(2)) * (BOA_ij - TR(2)))), ep2jk = rexp<TR>(-p_pen2 * ((BOA_jk - TR(2)) * (BOA_jk - TR(2))));
          const TR epen = p_pen1 * f9 * ep2ij * ep2jk;
          e_pen += (double)epen;
          const TR CEp1 = epen * Cf9 / f9, tp = -TR(2) * p_pen2 * epen, CEp2 = tp * (BOA_ij - TR(2)), CEp3 = tp * (BOA_jk - TR(2));
          const TR p_coa1 = (TR)A.p_coa1;
          const TR dti = tboi - BOA_ij, dtk = tbok - BOA_jk;
          const TR ecoa = p_coa1 / (TR(1) + ec2) * rexp<TR>(-p_coa3 * (dti * dti)) * rexp<TR>(-p_coa3 * (dtk * dtk)) *
                          rexp<TR>(-p_coa4 * ((BOA_ij - TR(1.5)) * (BOA_ij - TR(1.5)))) * rexp<TR>(-p_coa4 * ((BOA_jk - TR(1.5)) * (BOA_jk - TR(1.5))));
          e_coa += (double)ecoa;
          const TR CEc1 = -TR(2) * p_coa4 * (BOA_ij - TR(1.5)) * ecoa, CEc2 = -TR(2) * p_coa4 * (BOA_jk - TR(1.5)) * ecoa;
          const TR CEc3 = -p_coa2 * (ec2 / (TR(1) + ec2)) * ecoa, CEc4 = -TR(2) * p_coa3 * dti * ecoa, CEc5 = -TR(2) * p_coa3 * dtk * ecoa;
          own_cdbo += (double)(CE1 + CEp2 + (CEc1 - CEc4));
          to_cdbo += (double)(CE2 + CEp3 + (CEc2 - CEc5));
          cdself += (double)((CE3 + CE7) + CEp1 + CEc3);
          own_cdp += (double)CEc4;
          to_cdp += (double)CEc5;
          sCE5 += (double)CE5;
          sCE6 += (double)CE6;
          fjx += (double)(CE8 * dcj.x); fjy += (double)(CE8 * dcj.y); fjz += (double)(CE8 * dcj.z);
          own_fx += (double)(CE8 * dci.x); own_fy += (double)(CE8 * dci.y); own_fz += (double)(CE8 * dci.z);
          to_fx += (double)(CE8 * dck.x); to_fy += (double)(CE8 * dck.y); to_fz += (double)(CE8 * dck.z);
          if (PA) {   // F_i = -CE8 dcos/dx_i, F_k = -CE8 dcos/dx_k
            vir_add(vto, -1.0, (double)vij.x, (double)vij.y, (double)vij.z, (double)(CE8 * dci.x), (double)(CE8 * dci.y),
                    (double)(CE8 * dci.z));
            vir_add(vto, -1.0, (double)vjk.x, (double)vjk.y, (double)vjk.z, (double)(CE8 * dck.x), (double)(CE8 * dck.y),
                    (double)(CE8 * dck.z));
          }
        }
      }
    }
    to_cdbo = gsum<G>(to_cdbo); to_cdp = gsum<G>(to_cdp); to_fx = gsum<G>(to_fx); to_fy = gsum<G>(to_fy); to_fz = gsum<G>(to_fz);
    if (p == pk) { own_cdbo += to_cdbo; own_cdp += to_cdp; own_fx += to_fx; own_fy += to_fy; own_fz += to_fz; }
    if (PA) {
      for (int c = 0; c < 6; ++c) {
        vA[c] += vto[c];
        vC[c] += vto[c];
        const double tk = gsum<G>(vto[c]);
        if (p == pk) vA[c] += tk;
      }
    }
  }
  e_ang = gsum<G>(e_ang); e_pen = gsum<G>(e_pen); e_coa = gsum<G>(e_coa); cdself = gsum<G>(cdself);
  fjx = gsum<G>(fjx); fjy = gsum<G>(fjy); fjz = gsum<G>(fjz); sCE5 = gsum<G>(sCE5); sCE6 = gsum<G>(sCE6);
  // virial about the centre j: - sum over j's slots of (x_slot - x_j) (x) gradient of the slot atom
  double w[6] = {0, 0, 0, 0, 0, 0};
  if (mine) vir_add(w, -1.0, st.b_dx[qi], st.b_dy[qi], st.b_dz[qi], own_fx, own_fy, own_fz);
  for (int c = 0; c < 6; ++c) w[c] = gsum<G>(w[c]);
  if (mine) {
    const TR bo = BOij, c3 = bo * bo * bo;
    st.Cdbo[qi] += own_cdbo + (double)((TR)sCE6 * (c3 * c3 * bo));
    st.Cdbopi[qi] += (double)(TR)sCE5;
    st.Cdbopi2[qi] += (double)(TR)sCE5;
    st.cdp[qi] += own_cdp;
    st.fpx[qi] += own_fx; st.fpy[qi] += own_fy; st.fpz[qi] += own_fz;
  }
  if (live && p == 0) {
    st.CdD_self[j] += cdself;
    st.fs_x[j] += fjx; st.fs_y[j] += fjy; st.fs_z[j] += fjz;
    for (int c = 0; c < 6; ++c) st.vb[(size_t)c * n + j] += w[c];
    st.eatom[T_ANG * n + j] = e_ang;
    st.eatom[T_PEN * n + j] = e_pen;
    st.eatom[T_COA * n + j] = e_coa;
  }
  if (PA) {
    for (int c = 0; c < 6; ++c) vC[c] = gsum<G>(vC[c]);
    const size_t nsl = (size_t)st.maxb * n;
    if (mine)
      for (int c = 0; c < 6; ++c) st.vslot[c * nsl + qi] += vA[c] * (1.0 / 3.0);
    if (live && p == 0)
      for (int c = 0; c < 6; ++c) st.vref[(size_t)c * n + j] += vC[c] * (1.0 / 3.0);
  }
  if (st.census) {
    n_ang = gsum_int<G>(n_ang);
    if (live && p == 0) { st.cen[kCenAngles * n + j] = n_ang; st.cen[kCenSbo * n + j] = sbo_region; }
  }
}

// FP32 terms at the compiler's choice of registers; FP64 terms held to 256 VGPRs, 2 waves per SIMD (the compiler's choice, 260
// VGPRs, leaves one): the bound spills two 8-byte values before the partner loop, reloaded once per partner k outside the
// angle arithmetic, and gains 2-4% of the FP64 step
template <class TS, int G, bool PA>
__global__ void __launch_bounds__(256) k_valence(State st, Globals g, Tables t) {
  valence_body<float, TS, G, PA>(st, g, t);
}
template <int G, bool PA>
__global__ void __attribute__((amdgpu_waves_per_eu(2))) __launch_bounds__(256) k_valence_d(State st, Globals g, Tables t) {
  valence_body<double, double, G, PA>(st, g, t);
}

template <class TR>
__device__ inline TR tan_inv(TR s, TR c) {
  const TR M = TR(1e-10);
  if (s >= TR(0) && s <= M) return c / M;
  if (s <= TR(0) && s >= -M) return c / -M;
  return c / s;
}

// Torsion + four-body conjugation [reference bonded.cpp:476-652], evaluated at BOTH atoms of the central bond: lane p of j's group
// takes the central bond (j, k) in slot p and walks j's other bonds i (wave-uniform loop) and k's other bonds l. Each end keeps
// its own half (its atom, its outer partner i and outer bond (j, i)); the end the reference counts (canonical_first) also adds
// the energy and the central-bond coefficients. The outer-bond contributions of an i are summed over the group and kept by the
// lane owning slot i.
// PA: also the reference's per-atom virial split (bonded.cpp:636-647): each torsion's (x_l - x_i) (x) g_i + (x_l - x_j) (x) g_j +
// (x_l - x_k) (x) g_k (g the gradients), a quarter to each of its four atoms. Both ends evaluate every torsion, so each end gives
// the quarters of its own two atoms (j and the outer partner i, through i's slot).
template <class TR, class TS, int G, bool kCen, bool PA>
__device__ __forceinline__ void torsion_body(State st, Globals g, Tables t) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = gid / G, pk = gid % G, n = st.n;
  const bool live = j < n;
  const int tj = live ? st.type[j] : -1;
  const int nbj = live ? min(st.b_cnt[j], st.maxb) : 0;
  const int nbw = wave_max(nbj);
  double e_tor = 0, e_con = 0, cdself = 0, fjx = 0, fjy = 0, fjz = 0;
  double own_fx = 0, own_fy = 0, own_fz = 0;
  double fIx = 0, fIy = 0, fIz = 0;   // this lane's forces on the outer partners i, summed (for the virial)
  const TR thb_cut = (TR)g.thb_cut;
  const TS thb_cut_s = (TS)g.thb_cut;
  const TR p_tor2 = (TR)g.p_tor2, p_tor3 = (TR)g.p_tor3, p_tor4 = (TR)g.p_tor4, p_cot2 = (TR)g.p_cot2;
  const TR MINS = TR(1e-10);
  // this lane's central bond (j, k)
  bool cen = live && tj >= 0 && nbj >= 2 && pk < nbj;
  size_t qk = 0;
  int k = 0, sk = 0, tk = 0, nbk = 0, symk = 0;
  bool counted = false;
  TR BOjk = 0, BOA_jk = 0, rjk = 1, et2jk = 0, ec2jk = 0, f11 = 0, dfn11 = 0, BOpijk = 0;
  V3<TR> djk{0, 0, 0}, dkj{0, 0, 0};
  if (cen) {
    qk = sl(pk, j, n);
    BOjk = (TR)st.bo[qk];
    k = st.b_j[qk]; sk = st.b_s[qk]; tk = st.type[k];
    nbk = min(st.b_cnt[k], st.maxb);
    cen = (TS)st.bo[qk] > thb_cut_s && nbk >= 2;
  }
  if (cen) {
    counted = canonical_first(j, k, sk, st.tag);
    symk = st.b_sym[qk];   // slot of j in k's list
    BOA_jk = BOjk - thb_cut;
    rjk = (TR)st.b_d[qk];
    djk = V3<TR>{(TR)st.b_dx[qk], (TR)st.b_dy[qk], (TR)st.b_dz[qk]};
    dkj = V3<TR>{-djk.x, -djk.y, -djk.z};
    const TR Dj = (TR)st.Delta_boc[j], Dk = (TR)st.Delta_boc[k];
    et2jk = rexp<TR>(-p_tor2 * BOA_jk);
    ec2jk = rexp<TR>(-p_cot2 * ((BOA_jk - TR(1.5)) * (BOA_jk - TR(1.5))));
    const TR et3 = safe_exp<TR>(-p_tor3 * (Dj + Dk)), et4 = safe_exp<TR>(p_tor4 * (Dj + Dk)), it34 = TR(1) / (TR(1) + et3 + et4);
    f11 = (TR(2) + et3) * it34;
    BOpijk = (TR)st.bo_pi[qk];
    dfn11 = ratio2_deriv<TR>(et3, et4, -p_tor3, p_tor4, it34);
  }
  double cdbo_jk = 0, cdbopi_jk = 0;
  int n_tor = 0;   // census: torsions through the four-body gates that this lane's central bond counts
  double vJ[6] = {0, 0, 0, 0, 0, 0}, vIown[6] = {0, 0, 0, 0, 0, 0};   // PA: j's quarters, this lane's slot partner's quarters
  for (int pi = 0; pi < nbw; ++pi) {
    TR cdbo_ij = 0, fix = 0, fiy = 0, fiz = 0;   // this lane's terms of partner i: at most one per bond of k
    double vI[6] = {0, 0, 0, 0, 0, 0};
    if (cen && pi != pk && pi < nbj) {
      const size_t qi = sl(pi, j, n);
      const TS BOij_s = (TS)st.bo[qi];
      const TR BOij = (TR)st.bo[qi];
      if (BOij_s > thb_cut_s) {
        // the four-body gate's product, left to right as the reference (FP64 gates: the first two factors formed here)
        const TS BOijk_s = sizeof(TS) == 8 ? BOij_s * (TS)st.bo[qk] : TS(0);
        const int i = st.b_j[qi], si_ = st.b_s[qi], ti = st.type[i];
        const TR rij = (TR)st.b_d[qi], BOA_ij = BOij - thb_cut;
        const V3<TR> dji{(TR)st.b_dx[qi], (TR)st.b_dy[qi], (TR)st.b_dz[qi]};
        TR thijk, cijk_, sijk;
        V3<TR> dci, dcj, dck;   // d cos(theta_ijk) / d x_i, x_j, x_k
        angle_and_dcos<TR, false>(dji, rij, djk, rjk, thijk, cijk_, sijk, dci, dcj, dck);
        const TR cijk = cijk_, tijk = tan_inv<TR>(sijk, cijk);
        const TR et2ij = rexp<TR>(-p_tor2 * BOA_ij), ec2ij = rexp<TR>(-p_cot2 * ((BOA_ij - TR(1.5)) * (BOA_ij - TR(1.5))));
        for (int pl = 0; pl < nbk; ++pl) {
          if (pl == symk) continue;
          const size_t ql = sl(pl, k, n);
          const TS BOkl_s = (TS)st.bo[ql];
          if (!(BOkl_s > thb_cut_s)) continue;
          const TR BOkl = (TR)st.bo[ql];
          const int l = st.b_j[ql], sl_abs = shift_add(sk, st.b_s[ql]), tl = st.type[l];
          if (l == i && sl_abs == si_) continue;   // i and l are the same image
          const int ti4 = t.tors_idx[((ti * g.ntypes + tj) * g.ntypes + tk) * g.ntypes + tl];
          if (ti4 < 0) continue;
          if (sizeof(TS) == 8 ? !(BOijk_s * BOkl_s > thb_cut_s) : !(BOij * BOjk * BOkl > thb_cut)) continue;
          if (kCen) n_tor += counted ? 1 : 0;
          const DevTorsion P = t.tors[ti4];
          const TR rkl = (TR)st.b_d[ql], BOA_kl = BOkl - thb_cut;
          const V3<TR> dkl{(TR)st.b_dx[ql], (TR)st.b_dy[ql], (TR)st.b_dz[ql]};
          TR thjkl, cjkl_, sjkl;
          V3<TR> ecj, eck, ecl;   // d cos(theta_jkl) / d x_j, x_k, x_l
          angle_and_dcos<TR, false>(dkj, rjk, dkl, rkl, thjkl, cjkl_, sjkl, ecj, eck, ecl);
          const TR cjkl = cjkl_, tjkl = tan_inv<TR>(sjkl, cjkl);
          // x_i - x_l in j's frame
          const V3<TR> dli{dji.x - djk.x - dkl.x, dji.y - djk.y - dkl.y, dji.z - djk.z - dkl.z};
          const TR rli = rsqrt_<TR>(dot3(dli, dli));
          TR s1 = sijk, c1 = cijk, s2 = sjkl, c2 = cjkl;
          if (s1 >= TR(0) && s1 <= MINS) s1 = MINS; else if (s1 <= TR(0) && s1 >= -MINS) s1 = -MINS;
          if (s2 >= TR(0) && s2 <= MINS) s2 = MINS; else if (s2 <= TR(0) && s2 >= -MINS) s2 = -MINS;
          // bond vectors in the reference's orientation: dij (j->i), djk (j->k), dkl (k->l)
          const V3<TR>& bij = dji;
          const TR ucos = -dot3(bij, djk) * dot3(djk, dkl) + (rjk * rjk) * dot3(bij, dkl);
          const TR usin = -rjk * dot3(bij, cross3(djk, dkl));
          const TR htra = rij + c1 * (rkl * c2 - rjk), htrb = rjk - rij * c1 - rkl * c2;
          const TR hthd = rij * s1 * (rjk - rkl * c2), hthe = rkl * s2 * (rjk - rij * c1);
          const TR hnra = rkl * s1 * s2, hnhd = rij * rkl * c1 * s2, hnhe = rij * rkl * s1 * c2;
          const TR tel = (rij * rij) + (rjk * rjk) + (rkl * rkl) - (rli * rli) - TR(2) * (rij * rjk * c1 - rij * rkl * c1 * c2 + rjk * rkl * c2);
          const TR poem = TR(2) * rij * rkl * s1 * s2;
          TR arg = tel / poem;
          if (arg > TR(1)) arg = TR(1);
          if (arg < TR(-1)) arg = TR(-1);
          const TR ga = (htra - arg * hnra) / rij, gd = -(hthd - arg * hnhd) / s1, ge = -(hthe - arg * hnhe) / s2;
          const TR sc = TR(2) / poem;
          // d omega-like derivative vectors of i and j (the reference's dcos_omega_di / _dj)
          const V3<TR> doi{sc * (ga * bij.x - dli.x + gd * dci.x), sc * (ga * bij.y - dli.y + gd * dci.y), sc * (ga * bij.z - dli.z + gd * dci.z)};
          const V3<TR> doj{sc * (-ga * bij.x - htrb / rjk * djk.x + gd * dcj.x + ge * ecj.x), sc * (-ga * bij.y - htrb / rjk * djk.y + gd * dcj.y + ge * ecj.y),
                           sc * (-ga * bij.z - htrb / rjk * djk.z + gd * dcj.z + ge * ecj.z)};
          // cos(omega) for omega = atan2(usin, ucos), and cos(2 omega), cos(3 omega) by the multiple-angle identities (the
          // reference's cos(atan2(..)), cos(2 omega), cos(3 omega) to within a few ulp, without the four trigonometric calls;
          // atan2(0, 0) = 0 gives cos = 1)
          const TR h2 = ucos * ucos + usin * usin;
          const TR cw = h2 > TR(0) ? ucos / rsqrt_<TR>(h2) : TR(1);
          const TR c2w = TR(2) * (cw * cw) - TR(1), c3w = (TR(4) * (cw * cw) - TR(3)) * cw;
          const TR x = TR(2) - BOpijk - f11;
          const TR et1 = rexp<TR>((TR)P.p_tor1 * (x * x));
          const TR et2kl = rexp<TR>(-p_tor2 * BOA_kl), ec2kl = rexp<TR>(-p_cot2 * ((BOA_kl - TR(1.5)) * (BOA_kl - TR(1.5))));
          const TR fn10 = (TR(1) - et2ij) * (TR(1) - et2jk) * (TR(1) - et2kl);
          const TR V1 = (TR)P.V1, V2 = (TR)P.V2, V3_ = (TR)P.V3, p_tor1 = (TR)P.p_tor1, p_cot1 = (TR)P.p_cot1;
          const TR CV = TR(0.5) * (V1 * (TR(1) + cw) + V2 * et1 * (TR(1) - c2w) + V3_ * (TR(1) + c3w));
          const TR CT1 = sijk * sjkl * CV;
          const TR CT2 = -fn10 * TR(2) * p_tor1 * V2 * et1 * x * (TR(1) - (cw * cw)) * sijk * sjkl;
          const TR CT3 = CT2 * dfn11;
          const TR CT4 = CT1 * p_tor2 * et2ij * (TR(1) - et2jk) * (TR(1) - et2kl);
          const TR CT5 = CT1 * p_tor2 * (TR(1) - et2ij) * et2jk * (TR(1) - et2kl);
          const TR cmn = -fn10 * CV, CT7 = cmn * sjkl * tijk, CT8 = cmn * sijk * tjkl;
          const TR CT9 = fn10 * sijk * sjkl * (TR(0.5) * V1 - TR(2) * V2 * et1 * cw + TR(1.5) * V3_ * (c2w + TR(2) * (cw * cw)));
          const TR fn12 = ec2ij * ec2jk * ec2kl;
          const TR gg = TR(1) + ((cw * cw) - TR(1)) * sijk * sjkl;
          const TR Cc = -TR(2) * fn12 * p_cot1 * p_cot2 * gg;
          const TR CC1 = Cc * (BOA_ij - TR(1.5)), CC2 = Cc * (BOA_jk - TR(1.5));
          const TR CC4 = -p_cot1 * fn12 * ((cw * cw) - TR(1)) * sjkl * tijk;
          const TR CC5 = -p_cot1 * fn12 * ((cw * cw) - TR(1)) * sijk * tjkl;
          const TR CC6 = TR(2) * p_cot1 * fn12 * cw * sijk * sjkl;
          // this end's half: j itself, its outer partner i, its outer bond ij, its CdDelta
          cdself += (double)CT3;
          cdbo_ij += CT4 + CC1;
          const TR a7 = CT7 + CC4, a8 = CT8 + CC5, a9 = CT9 + CC6;
          const TR gix = a7 * dci.x + a9 * doi.x, giy = a7 * dci.y + a9 * doi.y, giz = a7 * dci.z + a9 * doi.z;
          const double gjx = (double)(a7 * dcj.x + a8 * ecj.x + a9 * doj.x), gjy = (double)(a7 * dcj.y + a8 * ecj.y + a9 * doj.y),
                       gjz = (double)(a7 * dcj.z + a8 * ecj.z + a9 * doj.z);
          fix += gix; fiy += giy; fiz += giz;
          fjx += gjx; fjy += gjy; fjz += gjz;
          if (PA) {   // g_k with the reference's dcos(omega)/dx_k; positions relative to l: x_l - x_i = -dli, x_l - x_j = djk + dkl
            const TR htrc = rkl + c2 * (rij * c1 - rjk), hnrc = rij * s1 * s2, gcl = (htrc - arg * hnrc) / rkl;
            const V3<TR> dok{sc * (-gcl * dkl.x + htrb / rjk * djk.x + gd * dck.x + ge * eck.x),
                             sc * (-gcl * dkl.y + htrb / rjk * djk.y + gd * dck.y + ge * eck.y),
                             sc * (-gcl * dkl.z + htrb / rjk * djk.z + gd * dck.z + ge * eck.z)};
            const double gkx = (double)(a7 * dck.x + a8 * eck.x + a9 * dok.x), gky = (double)(a7 * dck.y + a8 * eck.y + a9 * dok.y),
                         gkz = (double)(a7 * dck.z + a8 * eck.z + a9 * dok.z);
            double v[6] = {0, 0, 0, 0, 0, 0};
            vir_add(v, -1.0, (double)dli.x, (double)dli.y, (double)dli.z, (double)gix, (double)giy, (double)giz);
            vir_add(v, 1.0, (double)(djk.x + dkl.x), (double)(djk.y + dkl.y), (double)(djk.z + dkl.z), gjx, gjy, gjz);
            vir_add(v, 1.0, (double)dkl.x, (double)dkl.y, (double)dkl.z, gkx, gky, gkz);
            for (int c = 0; c < 6; ++c) { vJ[c] += v[c]; vI[c] += v[c]; }
          }
          if (counted) {   // energy and the central bond once per torsion
            e_tor += (double)(fn10 * sijk * sjkl * CV);
            e_con += (double)(p_cot1 * fn12 * gg);
            cdbopi_jk += (double)CT2;
            cdbo_jk += (double)(CT5 + CC2);
          }
        }
      }
    }
    fIx += (double)fix; fIy += (double)fiy; fIz += (double)fiz;
    const double gc = gsum<G>((double)cdbo_ij), gx = gsum<G>((double)fix), gy = gsum<G>((double)fiy), gz = gsum<G>((double)fiz);
    if (pk == pi) {   // the lane owning slot i: its BO coefficient now, its force for the virial and the end
      if (live && tj >= 0) st.Cdbo[sl(pk, j, n)] += gc;
      own_fx += gx; own_fy += gy; own_fz += gz;
    }
    if (PA)
      for (int c = 0; c < 6; ++c) {
        const double vs = gsum<G>(vI[c]);
        if (pk == pi) vIown[c] += vs;
      }
  }
  // This end's half of the torsion virial about the midpoint m of each central bond: with x_i - m = d_ji - d_jk / 2 and
  // x_j - m = -d_jk / 2, - sum (x_i - m) (x) g_i - sum (x_j - m) (x) g_j = - sum_p d_jp (x) F_p + sum_lanes (d_jk / 2) (x) (F_I + F_j),
  // F_p the group's force on the partner in slot p, F_I and F_j this lane's forces on its outer partners and on j.
  double w[6] = {0, 0, 0, 0, 0, 0};
  if (live && tj >= 0 && pk < nbj) {
    const size_t q = sl(pk, j, n);
    vir_add(w, -1.0, st.b_dx[q], st.b_dy[q], st.b_dz[q], own_fx, own_fy, own_fz);
  }
  if (cen) vir_add(w, 0.5, (double)djk.x, (double)djk.y, (double)djk.z, fIx + fjx, fIy + fjy, fIz + fjz);
  if (st.peratom && live && tj >= 0 && pk < nbj) st.eslot[sl(pk, j, n)] += 0.5 * (e_tor + e_con);   // this lane's central bond
  e_tor = gsum<G>(e_tor); e_con = gsum<G>(e_con); cdself = gsum<G>(cdself);
  fjx = gsum<G>(fjx); fjy = gsum<G>(fjy); fjz = gsum<G>(fjz);
  for (int c = 0; c < 6; ++c) w[c] = gsum<G>(w[c]);
  if (live && tj >= 0 && pk < nbj) {
    const size_t q = sl(pk, j, n);
    st.Cdbo[q] += cdbo_jk;
    st.Cdbopi[q] += cdbopi_jk;
    st.fpx[q] += own_fx; st.fpy[q] += own_fy; st.fpz[q] += own_fz;
  }
  if (live && pk == 0) {
    st.CdD_self[j] += cdself;
    st.fs_x[j] += fjx; st.fs_y[j] += fjy; st.fs_z[j] += fjz;
    for (int c = 0; c < 6; ++c) st.vb[(size_t)c * n + j] += w[c];
    st.eatom[T_TOR * n + j] = e_tor;
    st.eatom[T_CON * n + j] = e_con;
    if (st.peratom) st.eown[j] += 0.5 * (e_tor + e_con);
  }
  if (kCen) {
    n_tor = gsum_int<G>(n_tor);
    if (live && pk == 0) st.cen[kCenTorsions * n + j] = n_tor;
  }
  if (PA) {
    for (int c = 0; c < 6; ++c) vJ[c] = gsum<G>(vJ[c]);
    const size_t nsl = (size_t)st.maxb * n;
    if (live && tj >= 0 && pk < nbj)
      for (int c = 0; c < 6; ++c) st.vslot[c * nsl + sl(pk, j, n)] += 0.25 * vIown[c];
    if (live && pk == 0)
      for (int c = 0; c < 6; ++c) st.vref[(size_t)c * n + j] += 0.25 * vJ[c];
  }
}

// FP32 terms at 3 waves per SIMD (at most 168 VGPRs, no spill); FP64 at 2 (the compiler's choice, 258 VGPRs, leaves one; the
// bound spills two 8-byte values, reloaded once per outer partner i, outside the loop over l). kCen: the variant that also
// records the torsion census (launched only for a force call that asks for it, so the counter costs the ordinary step nothing)
template <class TS, int G, bool kCen, bool PA>
__global__ void __attribute__((amdgpu_waves_per_eu(3))) __launch_bounds__(256) k_torsion_f(State st, Globals g, Tables t) {
  torsion_body<float, TS, G, kCen, PA>(st, g, t);
}
template <int G, bool kCen, bool PA>
__global__ void __attribute__((amdgpu_waves_per_eu(2))) __launch_bounds__(256) k_torsion_d(State st, Globals g, Tables t) {
  torsion_body<double, double, G, kCen, PA>(st, g, t);
}

// ================================================================================================================
template <class TR>
__device__ inline void cos_and_dcos(const V3<TR>& a, TR da, const V3<TR>& b, TR db, TR& cost, V3<TR>& dca, V3<TR>& dcj, V3<TR>& dcb) {
  const TR inv = TR(1) / (da * db), ab = dot3(a, b);   // one division for the cosine and its derivatives
  cost = ab * inv;
  if (cost > TR(1)) cost = TR(1);
  if (cost < TR(-1)) cost = TR(-1);
  const TR da2 = da * da, db2 = db * db, inv3 = inv * inv * inv, cd = ab * inv3;
  dca = {b.x * inv - cd * db2 * a.x, b.y * inv - cd * db2 * a.y, b.z * inv - cd * db2 * a.z};
  dcj = {-(b.x + a.x) * inv + cd * (db2 * a.x + da2 * b.x), -(b.y + a.y) * inv + cd * (db2 * a.y + da2 * b.y),
         -(b.z + a.z) * inv + cd * (db2 * a.z + da2 * b.z)};
  dcb = {a.x * inv - cd * da2 * b.x, a.y * inv - cd * da2 * b.y, a.z * inv - cd * da2 * b.z};
}

// H-bond triple (i, j, k) [reference terms of bonded.cpp:654-715]: sin^4(theta/2) = ((1 - cos theta) / 2)^2, so no acos is
// needed; eh2 = exp(-p2 BO_ij) is the caller's (it depends on the bond only). C3 is the reference's CH3 / r_jk. Divisions by
// r_jk and r0 are products with their reciprocals (two divisions per triple instead of five)
template <class TR>
__device__ inline void hbond_term(const DevHBond& P, TR eh2, const V3<TR>& dji, TR rij, const V3<TR>& djk, TR rjk, TR& e, TR& CH1, TR& CH2,
                                  TR& C3, V3<TR>& dti, V3<TR>& dtj, V3<TR>& dtk) {
  TR ct;
  cos_and_dcos<TR>(dji, rij, djk, rjk, ct, dti, dtj, dtk);
  const TR cx = TR(1) - ct, hs = TR(0.5) * cx, s4 = hs * hs;
  const TR r0 = (TR)P.r0, p1 = (TR)P.p1, p2 = (TR)P.p2, p3 = (TR)P.p3;
  const TR ir = TR(1) / rjk, ir0 = TR(1) / r0;
  const TR eh3 = rexp<TR>(-p3 * (r0 * ir + rjk * ir0 - TR(2)));
  e = p1 * (TR(1) - eh2) * eh3 * s4;
  CH1 = p1 * p2 * eh2 * eh3 * s4;
  CH2 = -p1 / TR(2) * (TR(1) - eh2) * eh3 * cx;
  C3 = (-p3 * (-r0 * (ir * ir) + ir0) * e) * ir;
}

// Donor side: one wave per donor hydrogen j. A first pass over j's H-bond candidates (chunks of 64) computes each distance once,
// clears the acceptor-force slot of every candidate beyond hbond_cut and compacts the candidates within it (list order) into a
// per-wave LDS ring; the triple work then runs on full waves of 64 in-range candidates (typically one batch per donor instead of
// two half-empty chunks). For each bond (j, i) that can carry an H-bond (wave-uniform loop) a lane evaluates the triple (i, j, k)
// of its candidate: energy and the force on j stay in the lane, the BO derivative coefficient and the force on i are reduced over
// the wave into the bond slot, and the force on the acceptor k is written to the candidate's slot of hf for the acceptor to
// collect. Non-donors only clear their H-bond energy.
// PA (per-atom calls): also the reference's per-atom virial split (bonded.cpp:703-710): each triple's (x_j - x_i) (x) g_i +
// (x_j - x_k) (x) g_k, a third to the donor, to i (through the bond slot) and to the acceptor (through the candidate slot).
constexpr int kHbWaves = 4;
constexpr int kHbRing = 128;

template <class TR, class TS, bool PA>
__global__ void __launch_bounds__(64 * kHbWaves) k_hbond_donor(State st, Globals g, Tables t) {
  __shared__ int rk_[kHbWaves][kHbRing], rh_[kHbWaves][kHbRing], rt_[kHbWaves][kHbRing];
  __shared__ double rx_[kHbWaves][kHbRing], ry_[kHbWaves][kHbRing], rz_[kHbWaves][kHbRing], rr_[kHbWaves][kHbRing];
  const int wv = threadIdx.x >> 6;
  const int j = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (j >= st.n) return;
  const int lane = lane_id(), n = st.n, tj = st.type[j];
  const int role = (tj >= 0 && g.hbond_cut > 0) ? t.sp[tj].p_hbond : 0;
  if (role != 1) {
    if (lane == 0) {
      st.eatom[T_HB * n + j] = 0.0;
      st.h_cnt[j] = 0;
      if (st.census) st.cen[kCenHbonds * n + j] = 0;
    }
    return;
  }
  int* rk = rk_[wv]; int* rh = rh_[wv]; int* rt = rt_[wv];
  double *rx = rx_[wv], *ry = ry_[wv], *rz = rz_[wv], *rr = rr_[wv];
  const int nc = min(st.hc_cnt[j], st.maxhc);
  const size_t hbase = (size_t)j * st.maxhc;
  const double4 pj = st.pos[j];
  const double hcut = g.hbond_cut;
  const int nb = min(st.b_cnt[j], st.maxb);
  TR* hf = reinterpret_cast<TR*>(st.hf);
  double e_hb = 0, fjx = 0, fjy = 0, fjz = 0;
  double w[6] = {0, 0, 0, 0, 0, 0};   // virial of the triples about the donor
  int head = 0, fill = 0, inrange = 0;
  int n_hb = 0;   // census: triples evaluated (bond BO >= 0.01 to an acceptor type, parameters set, i and k not one atom)
  // the bonds (j, i) that can carry an H-bond (acceptor-type partner, BO >= 0.01), tested by the lanes at once (nb <= 64)
  unsigned long long hbonds;
  {
    bool ok = false;
    if (lane < nb) {
      const size_t q = sl(lane, j, n);
      ok = t.sp[st.type[st.b_j[q]]].p_hbond == 2 && (TS)st.bo[q] >= TS(1e-2);
    }
    hbonds = __ballot(ok);
  }
  // the triples of the m ring entries from head on
  auto batch = [&](int m) {
    const bool has = lane < m;
    int k = 0, tk = 0, h = 0;
    double dx = 0, dy = 0, dz = 0, r = 1;
    if (has) {
      const int e = (head + lane) & (kHbRing - 1);
      k = rk[e]; tk = rt[e]; h = rh[e]; dx = rx[e]; dy = ry[e]; dz = rz[e]; r = rr[e];
    }
    const long long tagk = has ? st.tag[k] : 0;
    const V3<TR> djk{(TR)dx, (TR)dy, (TR)dz};
    const TR rjk = (TR)r;
    double fkx = 0, fky = 0, fkz = 0, ek = 0;
    double vk[6] = {0, 0, 0, 0, 0, 0};   // PA: this candidate's triples
    for (unsigned long long left = hbonds; left; left &= left - 1) {   // wave-uniform, ascending slots
      double vt[6] = {0, 0, 0, 0, 0, 0};
      const int p = __ffsll((long long)left) - 1;
      const size_t q = sl(p, j, n);
      const int i = st.b_j[q], ti = st.type[i];
      const V3<TR> dji{(TR)st.b_dx[q], (TR)st.b_dy[q], (TR)st.b_dz[q]};
      const TR rij = (TR)st.b_d[q];
      double c1 = 0, fx = 0, fy = 0, fz = 0, et = 0;
      // the same atom as i, possibly another image, is not an acceptor of this triple (reference quirk Q-32)
      if (has && tagk != st.tag[i]) {
        const DevHBond P = t.hb[(ti * g.ntypes + tj) * g.ntypes + tk];
        if (P.r0 > 0.0) {
          ++n_hb;
          const TR e2 = rexp<TR>(-(TR)P.p2 * (TR)st.bo[q]);
          TR e, CH1, CH2, c3;
          V3<TR> dti, dtj, dtk;
          hbond_term<TR>(P, e2, dji, rij, djk, rjk, e, CH1, CH2, c3, dti, dtj, dtk);
          e_hb += (double)e;
          et = (double)e;
          c1 = (double)CH1;
          fx = (double)(CH2 * dti.x); fy = (double)(CH2 * dti.y); fz = (double)(CH2 * dti.z);
          fjx += (double)(CH2 * dtj.x - c3 * djk.x);
          fjy += (double)(CH2 * dtj.y - c3 * djk.y);
          fjz += (double)(CH2 * dtj.z - c3 * djk.z);
          const double gkx = (double)(CH2 * dtk.x + c3 * djk.x), gky = (double)(CH2 * dtk.y + c3 * djk.y),
                       gkz = (double)(CH2 * dtk.z + c3 * djk.z);
          fkx += gkx; fky += gky; fkz += gkz;
          vir_add(w, -1.0, dx, dy, dz, gkx, gky, gkz);
          vir_add(w, -1.0, (double)dji.x, (double)dji.y, (double)dji.z, fx, fy, fz);
          if (PA) {
            vir_add(vt, -1.0, dx, dy, dz, gkx, gky, gkz);
            vir_add(vt, -1.0, (double)dji.x, (double)dji.y, (double)dji.z, fx, fy, fz);
          }
        }
      }
      c1 = wave_sum(c1); fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
      if (lane == 0) { st.Cdbo[q] += c1; st.fpx[q] += fx; st.fpy[q] += fy; st.fpz[q] += fz; }
      if (PA) {
        const size_t nsl = (size_t)st.maxb * n;
        for (int c = 0; c < 6; ++c) {
          vk[c] += vt[c];
          const double vs = wave_sum(vt[c]);
          if (lane == 0) st.vslot[c * nsl + q] += vs * (1.0 / 3.0);
        }
      }
      if (st.peratom) {   // a third of each triple's energy to i (through the bond slot) and to k (through the candidate slot)
        ek += et;
        et = wave_sum(et);
        if (lane == 0) st.eslot[q] += et / 3.0;
      }
    }
    if (has) {
      const size_t o = 3 * (hbase + h);
      hf[o] = (TR)fkx; hf[o + 1] = (TR)fky; hf[o + 2] = (TR)fkz;
      if (st.peratom) st.ehc[hbase + h] = ek / 3.0;
      if (PA)
        for (int c = 0; c < 6; ++c) st.vhc[c * ((size_t)st.maxhc * n) + hbase + h] = vk[c] * (1.0 / 3.0);
    }
  };
  for (int h0 = 0; h0 < nc; h0 += 64) {
    const int h = h0 + lane;
    int k = 0, tk = -1;
    bool ok = false;
    double dx = 0, dy = 0, dz = 0, r = 0;
    if (h < nc) {
      k = st.hc_j[hbase + h];
      r = sqrt(pair_dvec_list(st, pj, k, st.hc_s[hbase + h], dx, dy, dz, tk));
      ok = r <= hcut;
      if (!ok) {   // beyond hbond_cut this step: no force on the acceptor from this slot
        const size_t o = 3 * (hbase + h);
        hf[o] = TR(0); hf[o + 1] = TR(0); hf[o + 2] = TR(0);
        if (st.peratom) st.ehc[hbase + h] = 0.0;
        if (PA)
          for (int c = 0; c < 6; ++c) st.vhc[c * ((size_t)st.maxhc * n) + hbase + h] = 0.0;
      }
    }
    const unsigned long long mk = __ballot(ok);
    if (ok) {
      const int e = (head + fill + lanes_below(mk)) & (kHbRing - 1);
      rk[e] = k; rt[e] = tk; rh[e] = h; rx[e] = dx; ry[e] = dy; rz[e] = dz; rr[e] = r;
    }
    fill += __popcll(mk);
    inrange += __popcll(mk);
    __builtin_amdgcn_wave_barrier();
    if (fill >= 64) {
      batch(64);
      head = (head + 64) & (kHbRing - 1);
      fill -= 64;
      __builtin_amdgcn_wave_barrier();
    }
  }
  if (fill > 0) batch(fill);
  e_hb = wave_sum(e_hb); fjx = wave_sum(fjx); fjy = wave_sum(fjy); fjz = wave_sum(fjz);
  for (int c = 0; c < 6; ++c) w[c] = wave_sum(w[c]);
  if (lane == 0) {
    st.fs_x[j] += fjx; st.fs_y[j] += fjy; st.fs_z[j] += fjz;
    for (int c = 0; c < 6; ++c) st.vb[(size_t)c * n + j] += w[c];
    if (PA)   // the donor's third: w is the sum of the triples' virials
      for (int c = 0; c < 6; ++c) st.vref[(size_t)c * n + j] += w[c] * (1.0 / 3.0);
    st.eatom[T_HB * n + j] = e_hb;
    st.h_cnt[j] = inrange;
  }
  if (st.census) {
    n_hb = wave_sum_int(n_hb);
    if (lane == 0) st.cen[kCenHbonds * n + j] = n_hb;
  }
  if (lane == 0) {
    if (st.peratom) st.eown[j] += e_hb / 3.0;
  }
}

// Acceptor side: one wave per acceptor k sums the forces its donors left in k's slots of their candidate lists (k's row of hc_j
// lists those slots in ascending order; fixed between neighbour rebuilds).
template <class TR>
__global__ void __launch_bounds__(256) k_hbond_acceptor(State st, Globals g, Tables t) {
  const int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (k >= st.n) return;
  const int lane = lane_id(), tk = st.type[k];
  const int role = (tk >= 0 && g.hbond_cut > 0) ? t.sp[tk].p_hbond : 0;
  if (role != 2) {
    if (st.peratom && lane == 0) st.eacc[k] = 0.0;
    return;
  }
  const int nc = min(st.hc_cnt[k], st.maxhc);
  const size_t hbase = (size_t)k * st.maxhc;
  const TR* hf = reinterpret_cast<const TR*>(st.hf);
  double fx = 0, fy = 0, fz = 0, e = 0;
  double va[6] = {0, 0, 0, 0, 0, 0};
  const size_t nhc = (size_t)st.maxhc * st.n;
  for (int h = lane; h < nc; h += 64) {
    const size_t o = (size_t)st.hc_j[hbase + h];   // flat donor slot
    fx += (double)hf[3 * o]; fy += (double)hf[3 * o + 1]; fz += (double)hf[3 * o + 2];
    if (st.peratom) {
      e += st.ehc[o];
      for (int c = 0; c < 6; ++c) va[c] += st.vhc[c * nhc + o];
    }
  }
  fx = wave_sum(fx); fy = wave_sum(fy); fz = wave_sum(fz);
  if (lane == 0) { st.fs_x[k] += fx; st.fs_y[k] += fy; st.fs_z[k] += fz; }
  if (st.peratom) {
    e = wave_sum(e);
    for (int c = 0; c < 6; ++c) va[c] = wave_sum(va[c]);
    if (lane == 0) {
      st.eacc[k] = e;
      for (int c = 0; c < 6; ++c) st.vref[(size_t)c * st.n + k] += va[c];
    }
  }
}

// per-atom energy in the reference's attribution: the atom's own shares, the valence terms of its centre, its non-bonded halves and
// polarisation, the shares its bond partners left in their mirror slots, and its H-bond thirds as an acceptor
__global__ void __launch_bounds__(128) k_eatom_ref(State st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  const int n = st.n, nb = min(st.b_cnt[i], st.maxb);
  double e = st.eown[i] + st.eatom[T_ANG * n + i] + st.eatom[T_PEN * n + i] + st.eatom[T_COA * n + i] + st.eatom[T_VDW * n + i] +
             st.eatom[T_ELE * n + i] + st.eatom[T_POL * n + i] + st.eacc[i];
  for (int p = 0; p < nb; ++p) {
    const size_t q = sl(p, i, n);
    e += st.eslot[sl(st.b_sym[q], st.b_j[q], n)];
  }
  st.eref[i] = e;
}

// ================================================================================================================
// bond-order derivative pass as gathers
// ================================================================================================================
// L lanes per atom as in k_atom_terms: a lane takes every L-th slot, the atom sums run in slot order through the group
template <int L>
__global__ void __launch_bounds__(256) k_cdelta(State st) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gid / L, p0 = gid % L, n = st.n;
  const bool live = i < n;
  const int nb = live ? min(st.b_cnt[i], st.maxb) : 0;
  const int nbw = wave_max(nb), base = lane_id() - p0;
  double c = live ? st.CdD_self[i] : 0.0;
  for (int k0 = 0; k0 < nbw; k0 += L) {
    const int p = k0 + p0;
    double v = 0;
    if (p < nb) {
      const size_t q = sl(p, i, n);
      v = st.cdp[sl(st.b_sym[q], st.b_j[q], n)];
    }
    for (int l = 0; l < L; ++l) c += __shfl(v, base + l, 64);
  }
  if (live && p0 == 0) st.CdD[i] = c;
}

template <int L>
__global__ void __launch_bounds__(256) k_dbond(State st) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = gid / L, p0 = gid % L, n = st.n;
  const bool live = i < n;
  const int nb = live ? min(st.b_cnt[i], st.maxb) : 0;
  const int nbw = wave_max(nb), base = lane_id() - p0;
  double S = 0, Dx = 0, Dy = 0, Dz = 0;
  double Ex = 0, Ey = 0, Ez = 0;   // per-atom calls: sum of the slots' coefficients of dDeltap_self times their bond vectors
  double w[6] = {0, 0, 0, 0, 0, 0};
  const double cdi = live ? st.CdD[i] : 0.0;
  for (int k0 = 0; k0 < nbw; k0 += L) {
    const int p = k0 + p0;
    double Sq = 0, sc = 0, dx = 0, dy = 0, dz = 0;
    if (p < nb) {
      const size_t q = sl(p, i, n);
      const int j = st.b_j[q];
      const size_t m = sl(st.b_sym[q], j, n);
      const double cbo = st.Cdbo[q] + st.Cdbo[m], cpi = st.Cdbopi[q] + st.Cdbopi[m], cpi2 = st.Cdbopi2[q] + st.Cdbopi2[m];
      const double cdel = cdi + st.CdD[j];
      const double C1dbo = st.C[0][q], C2pi = st.C[1][q], C2pi2 = st.C[2][q], C1pi = st.C[3][q], C1pi2 = st.C[4][q];
      const double T1 = (C1dbo * cbo + C1dbo * cdel) + C2pi * cpi + C2pi2 * cpi2;
      Sq = (st.C[5][q] * cbo + st.C[5][q] * cdel) + st.C[6][q] * cpi + st.C[7][q] * cpi2;
      sc = T1 * st.dbop_c[q] + (C1pi * cpi) * st.dlnpi_c[q] + (C1pi2 * cpi2) * st.dlnpi2_c[q];
      dx = st.b_dx[q]; dy = st.b_dy[q]; dz = st.b_dz[q];
    }
    for (int l = 0; l < L; ++l) {   // slot order; absent slots add exact zeros
      const double sq = __shfl(Sq, base + l, 64), s = __shfl(sc, base + l, 64);
      const double ax = __shfl(dx, base + l, 64), ay = __shfl(dy, base + l, 64), az = __shfl(dz, base + l, 64);
      S += sq;
      Ex += sq * ax; Ey += sq * ay; Ez += sq * az;
      Dx += s * ax; Dy += s * ay; Dz += s * az;
      // a central pair force: gradient c d on this end, virial d (x) c d, half at each end
      vir_add(w, 0.5 * s, ax, ay, az, ax, ay, az);
    }
  }
  if (!live || p0 != 0) return;
  for (int c = 0; c < 6; ++c) st.vb[(size_t)c * n + i] += w[c];
  st.S[i] = S;
  st.Dx[i] = Dx; st.Dy[i] = Dy; st.Dz[i] = Dz;
  if (st.peratom) { st.vE[i] = Ex; st.vE[n + i] = Ey; st.vE[2 * (size_t)n + i] = Ez; }
}

// Per-atom virial in the reference's split, completed (per-atom calls; before k_gather, while vatom holds the non-bonded half
// pairs). The bond-order derivative pass tallies per bond (i, j) [reference bonded.cpp:717-766]: (x_i - x_j) (x) F_i at i with
// F_i = -(t_i)/2, t_i the pair gradient of the bond plus its dDeltap_self coefficient times dDeltap_self(i); and at every
// partner k of i, ((x_k - x_i) + (x_k - x_j)) (x) F_k with F_k = -(c_k dBO'_ik)/2, the same for j's side. Summed over the bonds
// of an atom a: the own part is sum_q (sc_q / 2) d_q (x) d_q + E_a (x) dDeltap_self(a) / 2, and a partner k of a (slot r of a)
// receives (2 S_a d_r - E_a) (x) (dbop_c_r d_r) / 2, with S_a = sum_q S_q and E_a = sum_q S_q d_q.
__global__ void __launch_bounds__(128) k_vatom_ref(State st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  const int n = st.n, nb = min(st.b_cnt[i], st.maxb);
  const size_t nsl = (size_t)st.maxb * n;
  double v[6];
  for (int c = 0; c < 6; ++c) v[c] = st.vatom[(size_t)c * n + i] + st.vref[(size_t)c * n + i];
  const double cdi = st.CdD[i];
  for (int p = 0; p < nb; ++p) {
    const size_t q = sl(p, i, n);
    const int j = st.b_j[q];
    const size_t m = sl(st.b_sym[q], j, n);
    for (int c = 0; c < 6; ++c) v[c] += st.vslot[c * nsl + m];
    // this end of the bond: the pair part of t_i (the coefficient of k_dbond)
    const double cbo = st.Cdbo[q] + st.Cdbo[m], cpi = st.Cdbopi[q] + st.Cdbopi[m], cpi2 = st.Cdbopi2[q] + st.Cdbopi2[m];
    const double cdel = cdi + st.CdD[j];
    const double T1 = (st.C[0][q] * cbo + st.C[0][q] * cdel) + st.C[1][q] * cpi + st.C[2][q] * cpi2;
    const double sc = T1 * st.dbop_c[q] + (st.C[3][q] * cpi) * st.dlnpi_c[q] + (st.C[4][q] * cpi2) * st.dlnpi2_c[q];
    vir_add(v, 0.5 * sc, st.b_dx[q], st.b_dy[q], st.b_dz[q], st.b_dx[q], st.b_dy[q], st.b_dz[q]);
    // i as a partner of j (slot m of j)
    const double dmx = st.b_dx[m], dmy = st.b_dy[m], dmz = st.b_dz[m], Sj2 = 2.0 * st.S[j];
    vir_add(v, 0.5 * st.dbop_c[m], Sj2 * dmx - st.vE[j], Sj2 * dmy - st.vE[n + j], Sj2 * dmz - st.vE[2 * (size_t)n + j], dmx, dmy,
            dmz);
  }
  vir_add(v, 0.5, st.vE[i], st.vE[n + i], st.vE[2 * (size_t)n + i], st.dDps_x[i], st.dDps_y[i], st.dDps_z[i]);
  for (int c = 0; c < 6; ++c) st.vref[(size_t)c * n + i] = v[c];
}

// final gradient: bonded explicit-geometry terms (own + partner slots), BO-derivative pair forces, non-bonded
__global__ void __launch_bounds__(128) k_gather(State st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < st.n) gather_atom(st, i);
}

inline int nblk(long long items, int per) { return (int)((items + per - 1) / per); }

// lanes per atom of the slot-parallel per-atom kernels (bond-order correction, atom terms, derivative passes): 1, 2, 4 or 8,
// a run-time choice (the autotuner's); every choice gives the same sums bit for bit
template <class F>
void with_lanes(int lanes, F&& f) {
  switch (lanes) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    default: f(std::integral_constant<int, 4>{}); break;
  }
}

template <class TB>
void bond_build_t(const State& st, const Cell& c, const Globals& g, const Tables& t, hipStream_t s) {
  hipLaunchKernelGGL(k_bond_build<TB>, dim3(nblk(st.n, kBuildWaves)), dim3(64 * kBuildWaves), 0, s, st, c, g, t);
}

}  // namespace

void launch_bond_build(Precision p, const State& st, const Cell& c, const Globals& g, const Tables& t, hipStream_t s) {
  if (p == Precision::FP32) bond_build_t<float>(st, c, g, t, s);
  else bond_build_t<double>(st, c, g, t, s);
}


void launch_bo_correct(Precision p, const State& st, const Globals& g, const Tables& t, hipStream_t s) {
  with_lanes(st.lanes, [&](auto lc) {
    constexpr int L = decltype(lc)::value;
    const dim3 grid(nblk((long long)st.n * L, 256)), blk(256);
    if (p == Precision::FP32) hipLaunchKernelGGL((k_bo_correct<float, L>), grid, blk, 0, s, st, g, t);
    else hipLaunchKernelGGL((k_bo_correct<double, L>), grid, blk, 0, s, st, g, t);
  });
}

// PA: the per-atom variants (reference-split virial shares), built for G = 64 only: a group of 64 lanes gives the same sums as
// a smaller group (the extra lanes add exact zeros ahead of the same butterfly), so a per-atom call leaves the energies and the
// gradient bitwise unchanged
template <class TR, class TS, int G, bool PA>
void launch_terms_g(const State& st, const Globals& g, const Tables& t, hipStream_t s) {
  const dim3 grid(nblk((long long)st.n * G, 256)), blk(256), wgrid(nblk(st.n, 4));
  if (sizeof(TR) == 4) hipLaunchKernelGGL((k_valence<TS, G, PA>), grid, blk, 0, s, st, g, t);
  else hipLaunchKernelGGL((k_valence_d<G, PA>), grid, blk, 0, s, st, g, t);
  if (sizeof(TR) == 4) {
    if (st.census) hipLaunchKernelGGL((k_torsion_f<TS, G, true, PA>), grid, blk, 0, s, st, g, t);
    else hipLaunchKernelGGL((k_torsion_f<TS, G, false, PA>), grid, blk, 0, s, st, g, t);
  } else {
    if (st.census) hipLaunchKernelGGL((k_torsion_d<G, true, PA>), grid, blk, 0, s, st, g, t);
    else hipLaunchKernelGGL((k_torsion_d<G, false, PA>), grid, blk, 0, s, st, g, t);
  }
  hipLaunchKernelGGL((k_hbond_donor<TR, TS, PA>), dim3(nblk(st.n, kHbWaves)), dim3(64 * kHbWaves), 0, s, st, g, t);
  hipLaunchKernelGGL(k_hbond_acceptor<TR>, wgrid, blk, 0, s, st, g, t);
}

template <class TR, class TS>
void launch_terms_t(const State& st, const Globals& g, const Tables& t, int bmax, hipStream_t s) {
  if (st.peratom) launch_terms_g<TR, TS, 64, true>(st, g, t, s);
  else if (bmax <= 4) launch_terms_g<TR, TS, 4, false>(st, g, t, s);
  else if (bmax <= 8) launch_terms_g<TR, TS, 8, false>(st, g, t, s);
  else if (bmax <= 16) launch_terms_g<TR, TS, 16, false>(st, g, t, s);
  else if (bmax <= 32) launch_terms_g<TR, TS, 32, false>(st, g, t, s);
  else launch_terms_g<TR, TS, 64, false>(st, g, t, s);
}

void launch_atom_terms(Precision p, const State& st, const Globals& g, const Tables& t, hipStream_t s) {
  with_lanes(st.lanes, [&](auto lc) {
    constexpr int L = decltype(lc)::value;
    const dim3 grid(nblk((long long)st.n * L, 256)), blk(256);
    if (p == Precision::FP64) hipLaunchKernelGGL((k_atom_terms<double, double, L>), grid, blk, 0, s, st, g, t);
    else if (p == Precision::MIXED) hipLaunchKernelGGL((k_atom_terms<float, double, L>), grid, blk, 0, s, st, g, t);
    else hipLaunchKernelGGL((k_atom_terms<float, float, L>), grid, blk, 0, s, st, g, t);
  });
}

void launch_angle_terms(Precision p, const State& st, const Globals& g, const Tables& t, int bmax, hipStream_t s) {
  if (p == Precision::FP64) launch_terms_t<double, double>(st, g, t, bmax, s);
  else if (p == Precision::MIXED) launch_terms_t<float, double>(st, g, t, bmax, s);
  else launch_terms_t<float, float>(st, g, t, bmax, s);
}

void launch_bond_derivatives(const State& st, hipStream_t s) {
  with_lanes(st.lanes, [&](auto lc) {
    constexpr int L = decltype(lc)::value;
    const dim3 grid(nblk((long long)st.n * L, 256)), blk(256);
    hipLaunchKernelGGL(k_cdelta<L>, grid, blk, 0, s, st);
    hipLaunchKernelGGL(k_dbond<L>, grid, blk, 0, s, st);
  });
}

void launch_eatom_ref(const State& st, hipStream_t s) {
  hipLaunchKernelGGL(k_eatom_ref, dim3(nblk(st.n, 128)), dim3(128), 0, s, st);
}

void launch_vatom_ref(const State& st, hipStream_t s) {
  hipLaunchKernelGGL(k_vatom_ref, dim3(nblk(st.n, 128)), dim3(128), 0, s, st);
}

void launch_gather(const State& st, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3(nblk(st.n, 128)), dim3(128), 0, s, st);
}

}  // namespace rxa
