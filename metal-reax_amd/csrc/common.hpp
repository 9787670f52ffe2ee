// Shared device definitions of the MI355X ReaxFF engine: parameter blob, engine state, canonical pair geometry, wave reductions.
//
// Data model ("explicit images"): only the n owned atoms of a periodic cell exist on the device. A neighbour, bond or H-bond partner
// is (owner index j, integer lattice shift s); its position is x_j + s.L. Every per-atom quantity of an image (Delta', Delta, CdDelta,
// S, bond list) is that of its owner, so every force lands on an owned atom and no ghost fold-back is needed. This is the
// same physics as the reference's ghost-native expanded set (owner + lattice shift, reference src/core/system.hpp:43-58) for any
// ghost shell that is wide enough, which the reference requires anyway (required_shell, src/neighbor/neighbor.cpp:15).
//
// Canonical pair geometry: a pair (i, j, s) is always measured from its canonical end - the smaller tag, the smaller index for
// equal tags, and for a self image (i == j) the shift that is positive in (s2, s1, s0) order, which is the coordinate tie-break of
// the reference for LAMMPS-form cells (reference src/neighbor/neighbor.cpp:295-310, src/cpu/bonded.cpp:226-232). Both ends of a pair
// therefore see bitwise-identical distances (dvec exactly negated), so every cutoff decision (bond_cut, bo_cut, hbond_cut, nonb_cut)
// is taken identically from both sides, and the counted end is the end the reference counts.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rxa {

constexpr int kTerms = 13;   // e_bond, e_ov, e_un, e_lp, e_ang, e_pen, e_coa, e_hb, e_tor, e_con, e_vdW, e_ele, e_pol
enum TermIndex { T_BOND, T_OV, T_UN, T_LP, T_ANG, T_PEN, T_COA, T_HB, T_TOR, T_CON, T_VDW, T_ELE, T_POL };

struct DevSpecies {
  double r_s, r_pi, r_pi_pi, valency, valency_e, valency_boc, valency_val, nlp_opt;
  double p_lp2, p_ovun2, p_ovun5, p_val3, p_val5, chi, eta, gamma;
  int p_hbond, c2, heavy, pad;
};

struct DevPair {
  double p_bo1, p_bo2, p_bo3, p_bo4, p_bo5, p_bo6, r_s, r_p, r_pp, p_boc3, p_boc4, p_boc5;
  double p_be1, p_be2, De_s, De_p, De_pp, p_ovun1;
  double D, alpha, r_vdW, gamma_w, rcore, ecore, acore, lgcij, lgre, gamma, v13cor, ovc;
  double ln_rs, ln_rp, ln_rpp;   // log of r_s, r_p, r_pp (0 where the bond kind is absent)
  double pgw, irvdw, ircore, re6;   // (1 / gamma_w)^p_vdW1, 1 / r_vdW, 1 / rcore, lgre^6
  int tbs, s_ok, pi_ok, pp_ok;
};

// van der Waals / Coulomb parameters of a type pair for the single-precision non-bonded kernel (pgw = (1 / gamma_w)^p_vdW1)
// single-precision non-bonded pair constants, reciprocals and quotients precomputed (in FP64, then rounded) so the pair
// arithmetic needs no division by a constant
struct NbPairF { float D, alpha, irvdw, a_r, pgw, ecore, acore, ircore, ac_r, gamma, lgcij, re6; };   // a_r = alpha / r_vdW, ac_r = acore / rcore, re6 = lgre^6

struct DevAngle { double theta_00, p_val1, p_val2, p_val4, p_val7, p_pen1, p_coa1, pad; };
struct DevTorsion { double V1, V2, V3, p_tor1, p_cot1, pad[3]; };
struct DevHBond { double r0, p1, p2, p3; };

// global scalars (general parameters by role, cutoffs, taper coefficients), passed by value to every kernel
struct Globals {
  double p_boc1, p_boc2, p_lp1, p_lp3, gp3, gp4, gp7, gp10;
  double p_ovun3, p_ovun4, p_ovun6, p_ovun7, p_ovun8;
  double p_val6, p_val8, p_val9, p_val10, p_pen2, p_pen3, p_pen4, p_coa2, p_coa3, p_coa4;
  double p_tor2, p_tor3, p_tor4, p_cot2, p_vdW1;
  double bo_cut, nonb_low, nonb_cut, thb_cut, thb_cutsq, bond_cut, hbond_cut;
  double tap[8];      // non-bonded taper (swa = gp[11], swb = gp[12])
  double qtap[8];     // charge-equilibration taper (the fix's swa, swb)
  double q_swa, q_swb;
  int ntypes, vdw_type, lg, enobonds, c2_active, pad[3];
};

// pointers to the packed type tables
struct Tables {
  const DevSpecies* sp;     // ntypes
  const DevPair* pair;      // ntypes^2
  const int2* angle_idx;    // ntypes^3: (offset, count) into angle
  const DevAngle* angle;
  const int* tors_idx;      // ntypes^4: index into tors or -1
  const DevTorsion* tors;
  const DevHBond* hb;       // ntypes^3, r0 <= 0: none
  const double* qeq_chi;    // ntypes (fix parameters)
  const double* qeq_eta;
  const double* qeq_shld;   // ntypes^2: (gamma_i gamma_j)^-1.5
  const NbPairF* nbf;       // ntypes^2
};

struct Cell {   // LAMMPS-form periodic cell: a = (lx,0,0), b = (xy,ly,0), c = (xz,yz,lz), origin
  double o[3];
  double v[3][3];
  double cr[3][3];   // cr[d] = v[d+1] x v[d+2]; fractional coordinate f_d = ((r - o) . cr[d]) / det
  double det;        // v0 . (v1 x v2) > 0
  double h[3];       // distance between the two faces normal to direction d
  int per[3];
};

#pragma clang fp contract(off)
__host__ __device__ inline double frac_coord(const Cell& c, int d, double x, double y, double z) {
  const double dx = x - c.o[0], dy = y - c.o[1], dz = z - c.o[2];
  return (dx * c.cr[d][0] + dy * c.cr[d][1] + dz * c.cr[d][2]) / c.det;
}
#pragma clang fp contract(on)

// device state of one engine (owned atoms only); arrays of bond / H-bond slots are slot-major: index = p * n + i
struct State {
  int n, maxnb, maxb;
  double *x, *y, *z;          // owned positions (wrapped into the cell at every rebuild)
  double4* pos;               // the same positions packed (x, y, z, w) for one 32-byte gather per neighbour; w = type + 4096 * system
  const double4* shv;         // lattice shift vectors of every code in [-smx, smx]^3 (see shift_index), one table per system
  int smx0, smx1, smx2;
  int shv_stride;             // entries per system table (0 with one system)
  int* type;
  int* sysid;                 // system of each atom (several independent systems packed in one engine; all 0 otherwise)
  long long* tag;
  double* q;
  // skin neighbour list (atom-major: i * maxnb + k), full (both directions); nb_s = shift code | kCanon if i is the canonical end
  int *nb_j, *nb_s, *nb_cnt;
  // sub-lists of the skin list written at every rebuild (atom-major, same entry format): bond candidates (r <= bond_cut + skin)
  // and the acceptor candidates of H-bond donors (r <= hbond_cut + skin). An acceptor's row of hc_j instead holds, in ascending
  // order, the flat donor slots (j * maxhc + h) that list it; hc_m is the unsorted scratch of that transpose.
  int maxbc, maxhc;
  int *bc_j, *bc_s, *bc_cnt, *hc_j, *hc_s, *hc_cnt;
  int* hc_m;
  double* hf;                 // donors: force on the acceptor of each candidate slot (3 per slot, TR-typed)
  int* flags;                 // [0] neighbour overflow, [1] bond overflow, [2] H-bond overflow, [3] missing mirror bond,
                              // [4] bond-candidate overflow, [5] H-bond-candidate overflow
  // bond slots
  int *b_cnt, *b_j, *b_s, *b_sym;
  double *b_d, *b_dx, *b_dy, *b_dz;
  double *bop_s, *bop_pi, *bop_pi2, *dbop_c, *dlnpi_c, *dlnpi2_c;
  double *bo, *bo_s, *bo_pi, *bo_pi2;
  double* C[11];              // C1dbo C2dbo C3dbo C1dbopi C2dbopi C3dbopi C4dbopi C1dbopi2 C2dbopi2 C3dbopi2 C4dbopi2 (this end's view)
  double *Cdbo, *Cdbopi, *Cdbopi2, *fpx, *fpy, *fpz, *cdp;
  int* h_cnt;                 // H-bond partners within hbond_cut of this donor or acceptor at the last evaluation
  // per atom
  double *Dp, *Dpb, *tbo, *Delta, *Delta_e, *Delta_boc, *Delta_val, *vlpex, *nlp, *Delta_lp, *dDelta_lp, *Delta_lp_temp;
  double *dDps_x, *dDps_y, *dDps_z;
  double *CdD_self, *CdD, *S, *Dx, *Dy, *Dz;
  double *fs_x, *fs_y, *fs_z;     // explicit-geometry gradient of bonded terms computed at this atom
  double *gx, *gy, *gz;           // final gradient dE/dx (force = -g)
  double* eatom;                  // kTerms * n per-atom energies
  double* vatom;                  // 6 * n per-atom virial (xx yy zz xy xz yz): non-bonded half pairs + vb, complete after k_gather
  double* vb;                     // 6 * n bonded part of the per-atom virial, accumulated by the bonded kernels
  // per-atom energies in the reference's attribution (filled only when peratom != 0; bonded.cpp:212-704 ev_tally convention):
  int peratom;
  double* eslot;                  // maxb * n: energy share of the slot's partner (bond, C2, torsion halves; H-bond third of i)
  double* eown;                   // n: the atom's own share of the bond, lone-pair/C2, over/under, torsion and H-bond terms
  double* ehc;                    // maxhc * n: H-bond energy third of the acceptor in each donor candidate slot
  double* eacc;                   // n: an acceptor's gathered H-bond thirds
  double* eref;                   // n: the per-atom energy (k_eatom_ref)
  // per-atom virial in the reference's split (filled only when peratom != 0; bonded.cpp:462-466, 636-647, 703-710, 737-766):
  // valence a third of each angle's virial to each of its three atoms, torsion a quarter to each of four, H-bond a third to each
  // of three, the bond-order derivative pass per bond as the reference tallies it, non-bonded half pairs. Component-major.
  double* vref;                   // 6 * n: the atom's own shares, complete after k_vatom_ref
  double* vslot;                  // 6 * maxb * n: shares of the slot's partner (valence, torsion, H-bond i)
  double* vhc;                    // 6 * maxhc * n: H-bond share of the acceptor in each donor candidate slot
  double* vE;                     // 3 * n: sum over the atom's slots of the bond-order derivative coefficient times the bond vector
  double *gnb_x, *gnb_y, *gnb_z;  // non-bonded gradient
  // decision census (filled only when census != 0): kCenCount * n ints, [k * n + i] for the categories below
  int census;
  int* cen;
  int lanes;   // lanes per atom of the slot-parallel per-atom kernels (1, 2, 4 or 8; the same results for each)
};

// census categories recorded by the kernels that take the decisions (the bond count and the lone-pair truncation are read from
// the bond list and Delta_e): valence centre bond pairs through the three-body gates, torsions through the four-body gates at
// the central bond's counted end, H-bond triples evaluated at the donor, non-bonded pairs within nonb_cut at the counted end,
// the SBO branch of the valence centre (0: SBO <= 0, 1: (0, 1], 2: (1, 2), 3: >= 2)
enum CensusIndex { kCenAngles, kCenTorsions, kCenHbonds, kCenNonbonded, kCenSbo, kCenCount };

// ---- shift codes ------------------------------------------------------------------------------------------------
__host__ __device__ inline int shift_code(int s0, int s1, int s2) { return (s0 & 0xFF) | ((s1 & 0xFF) << 8) | ((s2 & 0xFF) << 16); }
__host__ __device__ inline void shift_decode(int c, int& s0, int& s1, int& s2) {
  s0 = (int)(signed char)(c & 0xFF);
  s1 = (int)(signed char)((c >> 8) & 0xFF);
  s2 = (int)(signed char)((c >> 16) & 0xFF);
}
__host__ __device__ inline int shift_neg(int c) { int a, b, d; shift_decode(c, a, b, d); return shift_code(-a, -b, -d); }
__host__ __device__ inline int shift_add(int c1, int c2) {
  int a, b, d, e, f, g;
  shift_decode(c1, a, b, d);
  shift_decode(c2, e, f, g);
  return shift_code(a + e, b + f, d + g);
}
__host__ __device__ inline bool shift_positive(int c) {
  int s0, s1, s2;
  shift_decode(c, s0, s1, s2);
  return s2 > 0 || (s2 == 0 && (s1 > 0 || (s1 == 0 && s0 > 0)));
}

// the end of pair (i, j, s) whose view is canonical
__device__ inline bool canonical_first(int i, int j, int s, const long long* tag) {
  const long long ti = tag[i], tj = tag[j];
  if (ti != tj) return ti < tj;
  if (i != j) return i < j;
  return shift_positive(s);
}

#pragma clang fp contract(off)
// lattice shift vector, accumulated as the image expander does (sum_k s_k v_k, k = 0, 1, 2)
__device__ inline void shift_vec(const Cell& c, int code, double& sx, double& sy, double& sz) {
  int s[3];
  shift_decode(code, s[0], s[1], s[2]);
  sx = 0.0; sy = 0.0; sz = 0.0;
  for (int k = 0; k < 3; ++k) {
    sx += (double)s[k] * c.v[k][0];
    sy += (double)s[k] * c.v[k][1];
    sz += (double)s[k] * c.v[k][2];
  }
}
// dvec = x_(j,s) - x_i measured from the canonical end; returns d^2 (no FMA contraction: identical to the host arithmetic)
__device__ inline double pair_dvec(const State& st, const Cell& c, int i, int j, int s, double& dx, double& dy, double& dz) {
  double sx, sy, sz;
  if (canonical_first(i, j, s, st.tag)) {
    shift_vec(c, s, sx, sy, sz);
    dx = (st.x[j] + sx) - st.x[i];
    dy = (st.y[j] + sy) - st.y[i];
    dz = (st.z[j] + sz) - st.z[i];
  } else {
    shift_vec(c, shift_neg(s), sx, sy, sz);
    dx = -((st.x[i] + sx) - st.x[j]);
    dy = -((st.y[i] + sy) - st.y[j]);
    dz = -((st.z[i] + sz) - st.z[j]);
  }
  return dx * dx + dy * dy + dz * dz;
}
#pragma clang fp contract(on)

// x^(-1/3) in double for normal x > 0: a single-precision seed from the hardware log2/exp2 (relative error ~1e-7), then two
// Newton steps for f(y) = y^-3 - x, y <- y (4 - x y^3) / 3, whose error goes e -> -2 e^2 (1e-7 -> 1e-13 -> rounding level).
// Replaces 1 / pow(x, 1/3) (a software pow plus a division) where the result only has to agree to a few ulp.
__device__ inline double rcbrt_f64(double x) {
  double y = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)x) * (-1.0f / 3.0f));
  y = y * (4.0 - x * (y * y * y)) * (1.0 / 3.0);
  y = y * (4.0 - x * (y * y * y)) * (1.0 / 3.0);
  return y;
}

// ---- virial: w[6] (xx yy zz xy xz yz) += s * (d (x) g) for a relative position d and a gradient g --------------------------
__device__ inline void vir_add(double* w, double s, double dx, double dy, double dz, double gx, double gy, double gz) {
  w[0] += s * dx * gx; w[1] += s * dy * gy; w[2] += s * dz * gz;
  w[3] += s * dx * gy; w[4] += s * dx * gz; w[5] += s * dy * gz;
}

// ---- list entries ---------------------------------------------------------------------------------------------------
constexpr int kCanon = 1 << 24;   // flag of a list entry: the listing atom is the canonical end of the pair

// pos.w carries the atom type (-1 for an untyped atom) and the atom's system: w = type + 4096 * system (exact in a double)
constexpr int kSysShift = 12;
__host__ __device__ inline double pos_w(int type, int sys) { return (double)(type + (sys << kSysShift)); }
__device__ inline int pos_sys(double w) { return ((int)w + 1) >> kSysShift; }
__device__ inline int pos_type(double w) { const int iw = (int)w; return iw - (((iw + 1) >> kSysShift) << kSysShift); }
__device__ inline double pos_w(const State& st, int i) { return pos_w(st.type[i], st.sysid[i]); }
__host__ __device__ inline int entry_code(int e) { return e & 0xFFFFFF; }

// index of a shift code in the shift-vector table; |s_d| <= smx_d is guaranteed by the list builders
__device__ inline int shift_index(const State& st, int code) {
  int s0, s1, s2;
  shift_decode(code, s0, s1, s2);
  return ((s2 + st.smx2) * (2 * st.smx1 + 1) + (s1 + st.smx1)) * (2 * st.smx0 + 1) + (s0 + st.smx0);
}

// table index of the shift that list entry e applies to its canonical end: the entry's code when the listing atom is that end,
// else the code's negation; the table enumerates the box [-smx, smx]^3 in shift_index order, so negating a code maps index k
// to N - 1 - k (N the table size)
__device__ inline int entry_shift_index(const State& st, int e) {
  const int k = shift_index(st, entry_code(e));
  return (e & kCanon) ? k : (2 * st.smx0 + 1) * (2 * st.smx1 + 1) * (2 * st.smx2 + 1) - 1 - k;
}

#pragma clang fp contract(off)
// pair_dvec for a list entry e = code | flag of neighbour j (position pj) of the atom at pi, without tags or branches: the same
// operands in the same order as pair_dvec (the table holds shift_vec of every code), so the result is bitwise identical
// (pi and pj by value: selecting between a reference into global memory and one to a local would put the local on the stack)
__device__ inline double pair_dvec_shift(const double4 pi, const double4 pj, bool canon, const double4 S, double& dx, double& dy,
                                         double& dz) {
  const double ax = canon ? pj.x : pi.x, ay = canon ? pj.y : pi.y, az = canon ? pj.z : pi.z;
  const double bx = canon ? pi.x : pj.x, by = canon ? pi.y : pj.y, bz = canon ? pi.z : pj.z;
  const double ex = (ax + S.x) - bx, ey = (ay + S.y) - by, ez = (az + S.z) - bz;
  dx = canon ? ex : -ex;
  dy = canon ? ey : -ey;
  dz = canon ? ez : -ez;
  return dx * dx + dy * dy + dz * dz;
}
__device__ inline double pair_dvec_pos(const State& st, const double4 pi, const double4 pj, int e, double& dx, double& dy, double& dz) {
  return pair_dvec_shift(pi, pj, (e & kCanon) != 0, st.shv[pos_sys(pi.w) * st.shv_stride + entry_shift_index(st, e)], dx, dy, dz);
}
__device__ inline double pair_dvec_list(const State& st, const double4& pi, int j, int e, double& dx, double& dy, double& dz) {
  return pair_dvec_pos(st, pi, st.pos[j], e, dx, dy, dz);
}
// the same, also returning the neighbour's type (carried in pos.w)
__device__ inline double pair_dvec_list(const State& st, const double4& pi, int j, int e, double& dx, double& dy, double& dz, int& tj) {
  tj = pos_type(st.pos[j].w);
  return pair_dvec_list(st, pi, j, e, dx, dy, dz);
}
#pragma clang fp contract(on)

// ---- wave (64-lane) reductions in a fixed butterfly order: deterministic run to run --------------------------------
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline int wave_sum_int(int v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline int lane_id() { return threadIdx.x & 63; }
// number of set bits of `mask` below this lane
__device__ inline int lanes_below(unsigned long long mask) { return __popcll(mask & ((1ull << lane_id()) - 1ull)); }

// slot p of atom i in the slot-major per-bond arrays
__device__ inline size_t sl(int p, int i, int n) { return (size_t)p * (size_t)n + (size_t)i; }

// The last step of a force call for atom i (the bonded chain and the non-bonded pass complete): the gradient from the atom's own
// sums, the forces its bond partners left in the mirror slots and the BO-derivative terms of its bonds, plus the non-bonded
// gradient; the per-atom virial gets the bonded part. Writes st.gx/gy/gz[i] and returns them.
__device__ inline void gather_atom(const State& st, int i, double& gx_out, double& gy_out, double& gz_out) {
  const int n = st.n, nb = min(st.b_cnt[i], st.maxb);
  const double Si = st.S[i];
  double w[6] = {0, 0, 0, 0, 0, 0};
  double gx = st.fs_x[i] + st.Dx[i] + Si * st.dDps_x[i];
  double gy = st.fs_y[i] + st.Dy[i] + Si * st.dDps_y[i];
  double gz = st.fs_z[i] + st.Dz[i] + Si * st.dDps_z[i];
  for (int p = 0; p < nb; ++p) {
    const size_t q = sl(p, i, n);
    const int j = st.b_j[q];
    const size_t m = sl(st.b_sym[q], j, n);
    gx += st.fpx[m]; gy += st.fpy[m]; gz += st.fpz[m];
    const double c = st.S[j] * st.dbop_c[q];
    gx += c * st.b_dx[q]; gy += c * st.b_dy[q]; gz += c * st.b_dz[q];
    vir_add(w, 0.5 * (c + Si * st.dbop_c[q]), st.b_dx[q], st.b_dy[q], st.b_dz[q], st.b_dx[q], st.b_dy[q], st.b_dz[q]);
  }
  gx_out = gx + st.gnb_x[i];
  gy_out = gy + st.gnb_y[i];
  gz_out = gz + st.gnb_z[i];
  st.gx[i] = gx_out;
  st.gy[i] = gy_out;
  st.gz[i] = gz_out;
  for (int c = 0; c < 6; ++c) st.vatom[(size_t)c * n + i] += st.vb[(size_t)c * n + i] + w[c];
}
__device__ inline void gather_atom(const State& st, int i) {
  double gx, gy, gz;
  gather_atom(st, i, gx, gy, gz);
}

}  // namespace rxa
