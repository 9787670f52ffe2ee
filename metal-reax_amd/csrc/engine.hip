// Host orchestration of the MI355X ReaxFF engine and its C ABI (include/reaxamd.h).
//
// One Engine per context: force-field tables packed once at pair_coeff time, device buffers sized for the atom set and regrown
// (grow-and-retry, the reference's policy at src/neighbor/neighbor.cpp:362-374 and src/neighbor/bonded_device.cpp:123-139) when a
// kernel reports that a list overflowed, and one HIP stream on which every kernel of a step is queued in order. The host reads back
// only the overflow flags, energies and (when asked) forces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>
#include <unordered_map>

#include "../../include/reaxamd.h"
#include "common.hpp"
#include "ffield.hpp"
#include "launch.hpp"

namespace rxa {

struct ApiError : std::runtime_error {
  int code;
  ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                      \
  do {                                                                                                 \
    hipError_t e_ = (x);                                                                               \
    if (e_ != hipSuccess) throw ApiError(RXA_EDEVICE, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

// device allocation owned by the engine
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    HIPCHK(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
};

// Optional per-phase device timing: a pair of HIP events around each timed launch sequence on the engine's stream, resolved at the
// next host synchronisation. Off by default (no events are recorded).
enum Phase { PH_NBR, PH_BONDS, PH_BOCORR, PH_TERMS, PH_NONB, PH_DBOND, PH_QEQ_H, PH_QEQ_SPMV, PH_QEQ_CG, PH_MD, PH_COUNT };
struct PhaseTimer {
  bool on = false;
  unsigned mask = 0;   // phases being timed (bit k = phase k)
  int every = 1;       // time every `every`-th launch sequence of a phase (the others carry no events)
  long long seen[PH_COUNT] = {0};
  hipStream_t s = nullptr;
  std::vector<hipEvent_t> pool;
  struct Pending { int ph; hipEvent_t a, b; };
  std::vector<Pending> pending;
  double ms[PH_COUNT] = {0};
  long long cnt[PH_COUNT] = {0};
  hipEvent_t take() {
    if (pool.empty()) {
      hipEvent_t e;
      HIPCHK(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t begin(int ph) {
    if (!on || !(mask >> ph & 1u)) return nullptr;
    if (seen[ph]++ % every != 0) return nullptr;
    hipEvent_t a = take();
    HIPCHK(hipEventRecord(a, s));
    return a;
  }
  void end(int ph, hipEvent_t a) {
    if (!on || !a) return;
    hipEvent_t b = take();
    HIPCHK(hipEventRecord(b, s));
    pending.push_back({ph, a, b});
  }
  void resolve() {   // call after a stream synchronisation
    for (auto& p : pending) {
      float t = 0;
      HIPCHK(hipEventElapsedTime(&t, p.a, p.b));
      ms[p.ph] += t;
      cnt[p.ph] += 1;
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pending.clear();
  }
  void reset() { for (int k = 0; k < PH_COUNT; ++k) { ms[k] = 0; cnt[k] = 0; seen[k] = 0; } }
  ~PhaseTimer() {
    for (auto e : pool) (void)hipEventDestroy(e);
    for (auto& p : pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  }
};

// LAMMPS' ghost-native atom set (nlocal owned atoms followed by nghost ghosts) checked against the explicit-image model: on one
// rank every ghost must be an image of an owned atom with the same id and type, displaced by a whole lattice vector along
// periodic directions only (the reference's host_atom_set, plugin/adapter/pair_reaxff_metal.cpp:244-291, and AtomSet::validate
// at 1e-8). More than one rank is refused: the owners of a sub-domain's ghosts live on other ranks, and this engine evaluates
// whole periodic cells. Throws ApiError with the first violation.
inline void check_ghosts(const Cell& c, long long nlocal, long long nghost, const int* types, const int64_t* tags, const double* x,
                         int nranks) {
  if (nranks != 1)
    throw ApiError(RXA_EINVAL, "multi-rank domain decomposition is not supported: this engine evaluates whole periodic cells on "
                               "one rank (got " + std::to_string(nranks) + " ranks; run one MPI rank per system)");
  if (nlocal <= 0 || nghost < 0) throw ApiError(RXA_EINVAL, "atom count out of range");
  std::unordered_map<long long, long long> owner;
  owner.reserve(2 * (size_t)nlocal);
  for (long long i = 0; i < nlocal; ++i)
    if (!owner.emplace(tags[i], i).second) throw ApiError(RXA_EINVAL, "duplicate atom id " + std::to_string(tags[i]) + " among owned atoms");
  double scale = 1.0;
  for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) scale = std::max(scale, std::fabs(c.v[a][b]));
  for (long long g = nlocal; g < nlocal + nghost; ++g) {
    const auto it = owner.find(tags[g]);
    if (it == owner.end())
      throw ApiError(RXA_EINVAL, "ghost " + std::to_string(g) + " (id " + std::to_string(tags[g]) + ") has no owned atom on this rank");
    const long long o = it->second;
    if (types[g] != types[o]) throw ApiError(RXA_EINVAL, "ghost " + std::to_string(g) + " has another type than its owned atom");
    double shift[3];
    for (int d = 0; d < 3; ++d) {
      const double fg = frac_coord(c, d, x[3 * g], x[3 * g + 1], x[3 * g + 2]), fo = frac_coord(c, d, x[3 * o], x[3 * o + 1], x[3 * o + 2]);
      shift[d] = std::nearbyint(fg - fo);
      if (!c.per[d] && shift[d] != 0.0)
        throw ApiError(RXA_EINVAL, "ghost " + std::to_string(g) + " is shifted along a non-periodic direction");
    }
    for (int k = 0; k < 3; ++k) {
      const double img = x[3 * o + k] + shift[0] * c.v[0][k] + shift[1] * c.v[1][k] + shift[2] * c.v[2][k];
      if (!(std::fabs(x[3 * g + k] - img) <= 1e-8 * scale))
        throw ApiError(RXA_EINVAL, "ghost " + std::to_string(g) + " is not a lattice image of its owned atom (id " + std::to_string(tags[g]) + ")");
    }
  }
}

struct Engine {
  int device = 0;
  PhaseTimer timer;
  hipStream_t stream = nullptr;
  std::string err;

  // force field
  bool have_ff = false, have_style = false;
  PairSettings ps;
  ForceField ff;
  std::vector<int> lmp2ff;   // LAMMPS type (1-based) -> ffield type, -1 NULL; index 0 unused
  Precision prec = Precision::FP64;
  Globals g{};
  DBuf<DevSpecies> d_sp;
  DBuf<DevPair> d_pair;
  DBuf<int2> d_angle_idx;
  DBuf<DevAngle> d_angle;
  DBuf<int> d_tors_idx;
  DBuf<DevTorsion> d_tors;
  DBuf<DevHBond> d_hb;
  DBuf<NbPairF> d_nbf;
  DBuf<double> d_qchi, d_qeta, d_qshld;
  Tables tab{};

  // atoms
  Cell cell{};
  bool have_box = false;
  int n = 0;
  State st{};
  int maxnb = 0, maxb = 12, maxbc = 0, maxhc = 0, bmax = 0;
  double skin = 2.0;
  DBuf<double> a_x, a_y, a_z, a_q;
  DBuf<double4> a_pos, a_shv;
  // engine order vs the caller's order: engine atom i is the caller's atom perm_h[i] (device copy d_perm). In cell mode the atoms
  // are put in the order of their cells at every neighbour rebuild, so neighbour gathers touch nearby memory.
  bool sort_atoms = true;
  std::vector<int> perm_h;
  DBuf<int> d_perm, d_src, d_psi;
  DBuf<double> d_psd;
  DBuf<long long> d_psl;
  int smx[3] = {-1, -1, -1};
  DBuf<int> a_type;
  DBuf<long long> a_tag;
  DBuf<double> a_atom;   // all per-atom double arrays
  DBuf<int> a_iatom;     // per-atom int arrays (nb_cnt, b_cnt, h_cnt, bc_cnt, hc_cnt) + flags
  DBuf<int> a_nbj, a_nbs;
  DBuf<int> a_bint;      // b_j, b_s, b_sym
  DBuf<double> a_bdbl;   // bond slot doubles
  DBuf<int> a_bc, a_hc;  // bond / H-bond candidate sub-lists (j, entry; H-bond: + mirror position)
  DBuf<double> a_hf;     // per H-bond candidate slot: force on the acceptor
  DBuf<double> a_ehc;    // per H-bond candidate slot: the acceptor's energy third (per-atom calls)
  DBuf<double> a_vref;   // per-atom calls: the reference-split virial (vref, vE, vslot, vhc), sized on first use
  BinGrid grid;
  BinBuffers bins;
  DBuf<unsigned> b_keys, b_skeys, b_akey;
  DBuf<int> b_idx, b_sidx, b_cstart, b_cend;
  DBuf<char> b_temp;
  bool cell_mode = false;
  int sm[3] = {0, 0, 0};
  bool nbr_valid = false;
  long long rebuilds = 0;
  DBuf<double> red;      // energy reduction output + scratch
  std::vector<double> host_types_mass;

  // charge equilibration
  QeqBuffers qb{};
  DBuf<int> q_int;
  DBuf<double> q_dbl;
  int q_dbl_n = -1;   // atom count the charge buffers were sized for
  DBuf<float> q_f32;        // FP32 values (and their FP32 remainders under refinement): mixed and FP32 modes
  DBuf<unsigned short> q_u16;
  DBuf<unsigned short> q_key;
  DBuf<double> q_scal;       // CG scalars and flags per iteration, sized by the iteration cap alone (never reallocated with the
                             // matrix or the per-atom vectors, so growing it cannot invalidate a built matrix)
  bool qeq_ready = false, qeq_h_valid = false;
  int qeq_imax_cap = 0;      // iteration cap the scalars, the packed flags and the pinned readback buffers are sized for
  int q_fp32_alloc = -1;     // value format the matrix arrays were allocated for (qb.fp32 at the time)
  // FP64 iterative refinement of the charge solve: -1 the precision's default (on in mixed mode), 0 off, 1 on (the fix's
  // 'resident' keyword; no effect in FP64 mode, whose matrix values are FP64)
  int qeq_refine_mode = -1;
  int hist_count = 0;
  double q_swa = 0, q_swb = 0;
  std::vector<double> q_chi_t, q_eta_t, q_gam_t;   // per ffield type

  // MD
  MdBuffers md{};
  DBuf<double> md_buf;
  bool md_ready = false, md_nve = false;
  bool v_scale_pending = false;   // the last thermostat velocity factor (md.nh[19]) is still to be applied
  double md_dt = 0.25;
  int md_qeq_every = 1, md_qeq_imax = 200;
  double md_qeq_tol = 1e-6;
  long long md_steps = 0;
  double qeq_iter_sum = 0;
  long long qeq_calls = 0;
  long long cg_batches = 0;   // readbacks of the single-system solve (host round trips)
  int last_iters[2] = {0, 0};
  double last_rel[2] = {0, 0};

  // pinned host words the device writes asynchronously (bond-list flags, displacement), and the events that mark them ready
  int* flags_h = nullptr;
  double* disp_h = nullptr;
  hipEvent_t ev_flags = nullptr, ev_disp = nullptr;
  bool nb_rows = true;   // non-bonded pass over the charge-matrix rows when they are current (RXA_NB_ROWS=0: skin list always)
  // The bonded chain (bond list ... BO-derivative pass) does not depend on the charges, and the charge solve and the non-bonded
  // pass do not depend on it: it runs on a second stream (default priority) between a fork and a join event, overlapping the
  // latency-bound CG kernels. RXA_OVERLAP=0, or timing a bonded phase, keeps everything on one stream.
  bool overlap = true;
  int g_mult = 1;   // lanes per atom of the angle kernels: the group size for the largest bond count times this (1 or 2)
  hipStream_t sb = nullptr;
  hipEvent_t ev_fork = nullptr, ev_bonded = nullptr, ev_cg = nullptr;
  int* cgf_h = nullptr;      // pinned: packed systems' any-active flags read back per batch
  double* cgs_h = nullptr;   // pinned: CG scalars and activity flags (the device layout of QeqBuffers::scal) read back per batch

  Engine(int dev) : device(dev) {
    const char* sort_env = std::getenv("RXA_ATOM_SORT");
    if (sort_env && std::atoi(sort_env) == 0) sort_atoms = false;
    const char* rows_env = std::getenv("RXA_NB_ROWS");
    if (rows_env && std::atoi(rows_env) == 0) nb_rows = false;
    const char* ovl_env = std::getenv("RXA_OVERLAP");
    if (ovl_env && std::atoi(ovl_env) == 0) overlap = false;
    // initial bond-slot capacity per atom (default 12); the list regrows on overflow whatever the start
    const char* bcap_env = std::getenv("RXA_BOND_CAP");
    if (bcap_env && std::atoi(bcap_env) > 0) maxb = std::atoi(bcap_env);
    HIPCHK(hipSetDevice(dev));
    st.shv_stride = 0;
    st.lanes = 4;
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&sb, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_bonded, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_cg, hipEventDisableTiming));
    timer.s = stream;
    HIPCHK(hipHostMalloc((void**)&flags_h, 8 * sizeof(int), hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void**)&disp_h, md_disp_partials() * sizeof(double), hipHostMallocDefault));
    HIPCHK(hipEventCreateWithFlags(&ev_flags, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev_disp, hipEventDisableTiming));
  }
  ~Engine() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (sb) (void)hipStreamSynchronize(sb);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_bonded) (void)hipEventDestroy(ev_bonded);
    if (ev_cg) (void)hipEventDestroy(ev_cg);
    if (cgf_h) (void)hipHostFree(cgf_h);
    if (pki_h) (void)hipHostFree(pki_h);
    if (pk_h) (void)hipHostFree(pk_h);
    if (rebany_h) (void)hipHostFree(rebany_h);
    if (cgs_h) (void)hipHostFree(cgs_h);
    if (sb) (void)hipStreamDestroy(sb);
    if (ev_flags) (void)hipEventDestroy(ev_flags);
    if (ev_disp) (void)hipEventDestroy(ev_disp);
    if (flags_h) (void)hipHostFree(flags_h);
    if (disp_h) (void)hipHostFree(disp_h);
    if (stream) (void)hipStreamDestroy(stream);
  }
  void sync() {
    HIPCHK(hipStreamSynchronize(sb));
    HIPCHK(hipStreamSynchronize(stream));
    timer.resolve();
  }
  // Host <-> device transfers and fills ordered on the engine's stream. The engine's streams are non-blocking, so they do not
  // wait for the null stream: a plain hipMemset, or a hipMemcpy from pageable memory, may still be in flight when a kernel on
  // them runs. h2d returns with the copy done (the host buffer may go out of scope); zero is ordered before later work on the
  // stream (and on the bonded stream, which forks from it).
  void h2d(void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void d2h(void* dst, const void* src, size_t bytes) {
    sync();
    if (!bytes) return;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
  }
  void zero(void* dst, size_t bytes) {
    if (bytes) HIPCHK(hipMemsetAsync(dst, 0, bytes, stream));
  }
  // the stream of the bonded chain
  hipStream_t bstream() const {
    const unsigned bonded_phases = 1u << PH_BONDS | 1u << PH_BOCORR | 1u << PH_TERMS | 1u << PH_DBOND;
    return overlap && !(timer.on && (timer.mask & bonded_phases)) ? sb : stream;
  }

  // ---------------------------------------------------------------------------------------------------------------
  void pack_tables() {
    const int nt = ff.ntypes();
    std::vector<DevSpecies> sp(nt);
    for (int i = 0; i < nt; ++i) {
      const Species& s = ff.sp[i];
      DevSpecies d{};
      d.r_s = s.r_s; d.r_pi = s.r_pi; d.r_pi_pi = s.r_pi_pi; d.valency = s.valency; d.valency_e = s.valency_e;
      d.valency_boc = s.valency_boc; d.valency_val = s.valency_val; d.nlp_opt = s.nlp_opt; d.p_lp2 = s.p_lp2;
      d.p_ovun2 = s.p_ovun2; d.p_ovun5 = s.p_ovun5; d.p_val3 = s.p_val3; d.p_val5 = s.p_val5; d.chi = s.chi; d.eta = s.eta; d.gamma = s.gamma;
      d.p_hbond = s.p_hbond; d.c2 = s.c2_species; d.heavy = s.heavy;
      sp[i] = d;
    }
    std::vector<DevPair> pr((size_t)nt * nt);
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nt; ++j) {
        const PairParams& p = ff.P(i, j);
        DevPair d{};
        d.p_bo1 = p.p_bo1; d.p_bo2 = p.p_bo2; d.p_bo3 = p.p_bo3; d.p_bo4 = p.p_bo4; d.p_bo5 = p.p_bo5; d.p_bo6 = p.p_bo6;
        d.r_s = p.r_s; d.r_p = p.r_p; d.r_pp = p.r_pp; d.p_boc3 = p.p_boc3; d.p_boc4 = p.p_boc4; d.p_boc5 = p.p_boc5;
        d.p_be1 = p.p_be1; d.p_be2 = p.p_be2; d.De_s = p.De_s; d.De_p = p.De_p; d.De_pp = p.De_pp; d.p_ovun1 = p.p_ovun1;
        d.D = p.D; d.alpha = p.alpha; d.r_vdW = p.r_vdW; d.gamma_w = p.gamma_w; d.rcore = p.rcore; d.ecore = p.ecore; d.acore = p.acore;
        d.lgcij = p.lgcij; d.lgre = p.lgre; d.gamma = p.gamma; d.v13cor = p.v13cor; d.ovc = p.ovc;
        d.tbs = p.triple_stab;
        d.s_ok = ff.sp[i].r_s > 0.0 && ff.sp[j].r_s > 0.0;
        d.pi_ok = ff.sp[i].r_pi > 0.0 && ff.sp[j].r_pi > 0.0;
        d.pp_ok = ff.sp[i].r_pi_pi > 0.0 && ff.sp[j].r_pi_pi > 0.0;
        d.ln_rs = d.s_ok && p.r_s > 0.0 ? std::log(p.r_s) : 0.0;
        d.ln_rp = d.pi_ok && p.r_p > 0.0 ? std::log(p.r_p) : 0.0;
        d.ln_rpp = d.pp_ok && p.r_pp > 0.0 ? std::log(p.r_pp) : 0.0;
        d.pgw = std::pow(1.0 / p.gamma_w, ff.gp[28]);
        d.irvdw = 1.0 / p.r_vdW;
        d.ircore = 1.0 / p.rcore;
        d.re6 = (p.lgre * p.lgre * p.lgre) * (p.lgre * p.lgre * p.lgre);
        pr[(size_t)i * nt + j] = d;
      }
    std::vector<int2> aidx((size_t)nt * nt * nt, int2{0, 0});
    std::vector<DevAngle> ang;
    for (const auto& kv : ff.angles) {
      const size_t key = ((size_t)kv.first[0] * nt + kv.first[1]) * nt + kv.first[2];
      aidx[key] = int2{(int)ang.size(), (int)kv.second.size()};
      for (const AngleSet& a : kv.second) ang.push_back(DevAngle{a.theta_00, a.p_val1, a.p_val2, a.p_val4, a.p_val7, a.p_pen1, a.p_coa1, 0});
    }
    std::vector<int> tidx((size_t)nt * nt * nt * nt, -1);
    std::vector<DevTorsion> tors;
    for (const auto& kv : ff.torsions) {
      const size_t key = (((size_t)kv.first[0] * nt + kv.first[1]) * nt + kv.first[2]) * nt + kv.first[3];
      tidx[key] = (int)tors.size();
      tors.push_back(DevTorsion{kv.second.V1, kv.second.V2, kv.second.V3, kv.second.p_tor1, kv.second.p_cot1, {0, 0, 0}});
    }
    std::vector<DevHBond> hb((size_t)nt * nt * nt, DevHBond{-1, 0, 0, 0});
    for (const auto& kv : ff.hbonds) hb[((size_t)kv.first[0] * nt + kv.first[1]) * nt + kv.first[2]] = DevHBond{kv.second.r0, kv.second.p1, kv.second.p2, kv.second.p3};
    auto up = [&](auto& buf, const auto& vec) {
      buf.alloc(vec.size());
      if (!vec.empty()) h2d(buf.p, vec.data(), vec.size() * sizeof(vec[0]));
    };
    std::vector<NbPairF> nbf((size_t)nt * nt);
    for (size_t k = 0; k < nbf.size(); ++k) {
      const DevPair& d = pr[k];
      NbPairF f{};
      f.D = (float)d.D; f.alpha = (float)d.alpha; f.irvdw = (float)(1.0 / d.r_vdW); f.a_r = (float)(d.alpha / d.r_vdW);
      f.pgw = (float)std::pow(1.0 / d.gamma_w, ff.gp[28]);
      f.ecore = (float)d.ecore; f.acore = (float)d.acore; f.ircore = (float)(1.0 / d.rcore); f.ac_r = (float)(d.acore / d.rcore);
      f.gamma = (float)d.gamma; f.lgcij = (float)d.lgcij; f.re6 = (float)std::pow(d.lgre, 6.0);
      nbf[k] = f;
    }
    up(d_sp, sp); up(d_pair, pr); up(d_angle_idx, aidx); up(d_angle, ang); up(d_tors_idx, tidx); up(d_tors, tors); up(d_hb, hb);
    up(d_nbf, nbf);
    d_qchi.alloc(nt); d_qeta.alloc(nt); d_qshld.alloc((size_t)nt * nt);
    tab.sp = d_sp.p; tab.pair = d_pair.p; tab.angle_idx = d_angle_idx.p; tab.angle = d_angle.p; tab.tors_idx = d_tors_idx.p; tab.tors = d_tors.p;
    tab.hb = d_hb.p; tab.qeq_chi = d_qchi.p; tab.qeq_eta = d_qeta.p; tab.qeq_shld = d_qshld.p; tab.nbf = d_nbf.p;
  }

  static void taper(double swa, double swb, double* c) {
    const double d1 = swb - swa, d7 = std::pow(d1, 7.0);
    const double a2 = swa * swa, a3 = swa * swa * swa, b2 = swb * swb, b3 = swb * swb * swb;
    c[7] = 20.0 / d7;
    c[6] = -70.0 * (swa + swb) / d7;
    c[5] = 84.0 * (a2 + 3.0 * swa * swb + b2) / d7;
    c[4] = -35.0 * (a3 + 9.0 * a2 * swb + 9.0 * swa * b2 + b3) / d7;
    c[3] = 140.0 * (a3 * swb + 3.0 * a2 * b2 + swa * b3) / d7;
    c[2] = -210.0 * (a3 * b2 + a2 * b3) / d7;
    c[1] = 140.0 * a3 * b3 / d7;
    c[0] = (-35.0 * a3 * b2 * b2 + 21.0 * a2 * b3 * b2 - 7.0 * swa * b3 * b3 + b3 * b3 * swb) / d7;
  }

  void make_globals() {
    const auto& gp = ff.gp;
    const Control& c = ps.control;
    g = Globals{};
    g.p_boc1 = gp[0]; g.p_boc2 = gp[1]; g.p_lp1 = gp[15]; g.p_lp3 = gp[5]; g.gp3 = gp[3]; g.gp4 = gp[4]; g.gp7 = gp[7]; g.gp10 = gp[10];
    g.p_ovun3 = gp[32]; g.p_ovun4 = gp[31]; g.p_ovun6 = gp[6]; g.p_ovun7 = gp[8]; g.p_ovun8 = gp[9];
    g.p_val6 = gp[14]; g.p_val8 = gp[33]; g.p_val9 = gp[16]; g.p_val10 = gp[17];
    if (gp.size() <= 38) throw ApiError(RXA_EFFIELD, "the force field needs general parameter 39 (p_coa3)");
    g.p_pen2 = gp[19]; g.p_pen3 = gp[20]; g.p_pen4 = gp[21]; g.p_coa2 = gp[2]; g.p_coa3 = gp[38]; g.p_coa4 = gp[30];
    g.p_tor2 = gp[23]; g.p_tor3 = gp[24]; g.p_tor4 = gp[25]; g.p_cot2 = gp[27]; g.p_vdW1 = gp[28];
    g.bo_cut = ff.bo_cut; g.nonb_low = ff.nonb_low; g.nonb_cut = ff.nonb_cut;
    g.thb_cut = c.thb_cut; g.thb_cutsq = c.thb_cutsq; g.bond_cut = c.bond_cut; g.hbond_cut = c.hbond_cut;
    taper(ff.nonb_low, ff.nonb_cut, g.tap);
    g.ntypes = ff.ntypes(); g.vdw_type = ff.vdw_type; g.lg = ps.lgvdw; g.enobonds = ps.enobonds; g.c2_active = ff.c2_active;
    if (qeq_ready) { taper(q_swa, q_swb, g.qtap); g.q_swa = q_swa; g.q_swb = q_swb; }
    const std::string p = ps.effective_precision();
    prec = p == "fp64" ? Precision::FP64 : (p == "fp32" ? Precision::FP32 : Precision::MIXED);
    qb.fp32 = prec == Precision::FP64 ? 0 : 1;   // the charge solve multiplies with FP32 matrix values outside FP64 mode
    apply_refine();
    qeq_h_valid = false;
  }
  void apply_refine() {
    const int r = (prec == Precision::MIXED && qeq_refine_mode != 0) || (prec == Precision::FP32 && qeq_refine_mode == 1) ? 1 : 0;
    if (r != qb.refine) qeq_h_valid = false;   // the matrix must be rebuilt with (or may drop) its FP32 remainders
    qb.refine = r;
  }

  double list_cut() const { return std::max({g.nonb_cut, g.hbond_cut, g.bond_cut, qeq_ready ? q_swb : 0.0}); }

  // ---------------------------------------------------------------------------------------------------------------
  void set_box(const double lo[3], const double hi[3], const double tilt[3], const int per[3]) {
    cell = make_cell(lo, hi, tilt, per);
    have_box = true;
    nbr_valid = false;
  }
  static Cell make_cell(const double lo[3], const double hi[3], const double tilt[3], const int per[3]) {
    Cell c{};
    for (int d = 0; d < 3; ++d) c.o[d] = lo[d];
    const double v[3][3] = {{hi[0] - lo[0], 0, 0}, {tilt[0], hi[1] - lo[1], 0}, {tilt[1], tilt[2], hi[2] - lo[2]}};
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) c.v[a][b] = v[a][b];
    auto cross = [](const double* a, const double* b, double* r) { r[0] = a[1] * b[2] - a[2] * b[1]; r[1] = a[2] * b[0] - a[0] * b[2]; r[2] = a[0] * b[1] - a[1] * b[0]; };
    cross(c.v[1], c.v[2], c.cr[0]);
    cross(c.v[2], c.v[0], c.cr[1]);
    cross(c.v[0], c.v[1], c.cr[2]);
    c.det = c.v[0][0] * c.cr[0][0] + c.v[0][1] * c.cr[0][1] + c.v[0][2] * c.cr[0][2];
    if (!(c.det > 0) || !std::isfinite(c.det)) throw ApiError(RXA_EINVAL, "box is degenerate or left-handed");
    for (int d = 0; d < 3; ++d) {
      const double nr = std::sqrt(c.cr[d][0] * c.cr[d][0] + c.cr[d][1] * c.cr[d][1] + c.cr[d][2] * c.cr[d][2]);
      c.h[d] = std::fabs(c.det) / nr;
      c.per[d] = per[d] != 0;
    }
    return c;
  }

  double* atom_arr(int k) { return a_atom.p + (size_t)k * n; }

  void alloc_atoms() {
    a_x.alloc(n); a_y.alloc(n); a_z.alloc(n); a_q.alloc(n); a_type.alloc(n); a_tag.alloc(n); a_pos.alloc(n);
    d_perm.alloc(n); d_src.alloc(n); d_psd.alloc(24 * (size_t)n); d_psi.alloc(4 * (size_t)n); d_psl.alloc(2 * (size_t)n);
    zero(a_q.p, n * sizeof(double));
    const int NA = 38 + kTerms + 9;   // number of per-atom double arrays below
    a_atom.alloc((size_t)NA * n);
    zero(a_atom.p, (size_t)NA * n * sizeof(double));
    const size_t ni = (6 + kCenCount) * (size_t)n + 8;   // 6 count arrays, 8 flags, the census
    a_iatom.alloc(ni);
    zero(a_iatom.p, ni * sizeof(int));
    State& s = st;
    s.n = n;
    s.x = a_x.p; s.y = a_y.p; s.z = a_z.p; s.q = a_q.p; s.type = a_type.p; s.tag = a_tag.p; s.pos = a_pos.p;
    int k = 0;
    s.Dp = atom_arr(k++); s.Dpb = atom_arr(k++); s.tbo = atom_arr(k++); s.Delta = atom_arr(k++); s.Delta_e = atom_arr(k++);
    s.Delta_boc = atom_arr(k++); s.Delta_val = atom_arr(k++); s.vlpex = atom_arr(k++); s.nlp = atom_arr(k++); s.Delta_lp = atom_arr(k++);
    s.dDelta_lp = atom_arr(k++); s.Delta_lp_temp = atom_arr(k++); s.dDps_x = atom_arr(k++); s.dDps_y = atom_arr(k++); s.dDps_z = atom_arr(k++);
    s.CdD_self = atom_arr(k++); s.CdD = atom_arr(k++); s.S = atom_arr(k++); s.Dx = atom_arr(k++); s.Dy = atom_arr(k++); s.Dz = atom_arr(k++);
    s.fs_x = atom_arr(k++); s.fs_y = atom_arr(k++); s.fs_z = atom_arr(k++); s.gx = atom_arr(k++); s.gy = atom_arr(k++); s.gz = atom_arr(k++);
    s.gnb_x = atom_arr(k++); s.gnb_y = atom_arr(k++); s.gnb_z = atom_arr(k++);
    s.vb = atom_arr(k);   // 6 arrays: 30..35
    k = 38;
    s.eatom = atom_arr(k);
    s.vatom = atom_arr(k + kTerms);
    s.eown = atom_arr(k + kTerms + 6); s.eacc = atom_arr(k + kTerms + 7); s.eref = atom_arr(k + kTerms + 8);
    s.nb_cnt = a_iatom.p; s.b_cnt = a_iatom.p + n; s.h_cnt = a_iatom.p + 2 * (size_t)n; s.bc_cnt = a_iatom.p + 3 * (size_t)n;
    s.hc_cnt = a_iatom.p + 4 * (size_t)n; s.sysid = a_iatom.p + 5 * (size_t)n; s.flags = a_iatom.p + 6 * (size_t)n;
    s.cen = a_iatom.p + 6 * (size_t)n + 8;
    s.census = 0;
    census_valid = false;
    red.alloc(19 + 240 * 19 + 8);
    // binning scratch
    b_keys.alloc(n); b_skeys.alloc(n); b_akey.alloc(n); b_idx.alloc(n); b_sidx.alloc(n);
    bins.keys = b_keys.p; bins.skeys = b_skeys.p; bins.atom_key = b_akey.p; bins.idx = b_idx.p; bins.sidx = b_sidx.p;
    bins.temp_bytes = bin_temp_bytes(n);
    b_temp.alloc(bins.temp_bytes + 16);
    bins.temp = b_temp.p;
  }

  void alloc_neighbors(int cap) {
    maxnb = cap;
    a_nbj.alloc((size_t)n * maxnb);
    a_nbs.alloc((size_t)n * maxnb);
    st.maxnb = maxnb; st.nb_j = a_nbj.p; st.nb_s = a_nbs.p;
    qeq_h_valid = false;
    if (qeq_ready) alloc_qeq_matrix();
  }

  void alloc_sublists(int cap_bc, int cap_hc) {
    maxbc = cap_bc;
    maxhc = cap_hc;
    a_bc.alloc(2 * (size_t)n * maxbc);
    a_hc.alloc(3 * (size_t)n * maxhc);
    a_hf.alloc(3 * (size_t)n * maxhc);
    a_ehc.alloc((size_t)n * maxhc);
    st.maxbc = maxbc; st.maxhc = maxhc;
    st.bc_j = a_bc.p; st.bc_s = a_bc.p + (size_t)n * maxbc;
    st.hc_j = a_hc.p; st.hc_s = a_hc.p + (size_t)n * maxhc; st.hc_m = a_hc.p + 2 * (size_t)n * maxhc;
    st.hf = a_hf.p;
    st.ehc = a_ehc.p;
    if (st.peratom) alloc_peratom();
  }

  // the per-atom virial buffers of the reference's split for the current atom count and list capacities (per-atom calls only;
  // called with no work queued that uses them)
  void alloc_peratom() {
    const size_t need = 9 * (size_t)n + 6 * (size_t)n * maxb + 6 * (size_t)n * maxhc;
    if (a_vref.n != need) a_vref.alloc(need);
    st.vref = a_vref.p;
    st.vE = a_vref.p + 6 * (size_t)n;
    st.vslot = a_vref.p + 9 * (size_t)n;
    st.vhc = st.vslot + 6 * (size_t)n * maxb;
  }

  void alloc_bonds(int cap_b) {
    maxb = cap_b;
    const size_t nb = (size_t)n * maxb;
    a_bint.alloc(3 * nb);
    a_bdbl.alloc(30 * nb);
    st.maxb = maxb;
    st.b_j = a_bint.p; st.b_s = a_bint.p + nb; st.b_sym = a_bint.p + 2 * nb;
    double* b = a_bdbl.p;
    int k = 0;
    auto nxt = [&]() { return b + (size_t)(k++) * nb; };
    st.b_d = nxt(); st.b_dx = nxt(); st.b_dy = nxt(); st.b_dz = nxt();
    st.bop_s = nxt(); st.bop_pi = nxt(); st.bop_pi2 = nxt(); st.dbop_c = nxt(); st.dlnpi_c = nxt(); st.dlnpi2_c = nxt();
    st.bo = nxt(); st.bo_s = nxt(); st.bo_pi = nxt(); st.bo_pi2 = nxt();
    for (int c = 0; c < 8; ++c) st.C[c] = nxt();
    for (int c = 8; c < 11; ++c) st.C[c] = nullptr;
    st.Cdbo = nxt(); st.Cdbopi = nxt(); st.Cdbopi2 = nxt(); st.fpx = nxt(); st.fpy = nxt(); st.fpz = nxt();
    st.cdp = nxt();
    st.eslot = nxt();
    if (k != 30) throw ApiError(RXA_EINVAL, "internal: bond slot layout");
    if (st.peratom) alloc_peratom();
  }

  // expected entries within r of an atom at the mean density (at least that of a condensed phase for open cells), with margin
  int estimate_within(double r) const {
    double dens = packed ? pk_dens : n / cell.det;
    if (!packed && !(cell.per[0] && cell.per[1] && cell.per[2])) dens = std::max(dens, 0.12);
    const double est = dens * 4.0 / 3.0 * M_PI * r * r * r;
    const int cap = (int)(est * 1.25) + 32;
    return std::min(cap, std::max(64, 27 * n + 32));
  }
  int estimate_neighbors() const { return estimate_within(list_cut() + skin); }

  // moves every persistent per-atom array to a new order: new[i] = old[src[i]] (src on the device)
  void permute(const int* src) {
    PermuteSet ps;
    ps.n = n;
    auto D = [&](double* p) { ps.d[ps.nd++] = p; };
    D(st.x); D(st.y); D(st.z); D(st.q);
    if (md_ready) { D(md.vx); D(md.vy); D(md.vz); D(md.mass); }
    if (qeq_ready && qb.s_hist)
      for (int k = 0; k < 4; ++k) { D(qb.s_hist + (size_t)k * n); D(qb.t_hist + (size_t)k * n); }
    ps.i[ps.ni++] = st.type;
    ps.i[ps.ni++] = d_perm.p;
    ps.l[ps.nl++] = st.tag;
    launch_permute(ps, src, d_psd.p, d_psi.p, d_psl.p, stream);
  }

  // ---- several independent systems packed in one engine --------------------------------------------------------------
  // Atoms of system s occupy [sys_start[s], sys_start[s+1]) in the caller's order; each system has its own cell, shift table,
  // thermostat, charge-solve scalars and energy sums, and its own neighbour rebuilds. Lists never cross systems.
  bool packed = false;
  int S = 1;
  double pk_dens = 0;
  std::vector<int> sys_start_h;
  std::vector<Cell> cells_h;
  std::vector<int> ext_h;
  bool shv_valid = false;
  DBuf<int> d_sys_start, d_reb, d_rebany, d_pki, d_cg_any;
  DBuf<Cell> d_cells;
  DBuf<int4> d_ext;
  DBuf<double> d_pk, d_rowv, d_nh_pk, d_disp_pk, d_sys_out;
  PackedQeq pq;
  int* rebany_h = nullptr;   // pinned
  int* pki_h = nullptr;      // pinned, 4 S
  double* pk_h = nullptr;    // pinned, 10 S
  size_t pk_cap_h = 0;

  void leave_packed() {
    packed = false;
    S = 1;
    st.shv_stride = 0;
    smx[0] = smx[1] = smx[2] = -1;
  }

  void set_systems(int ns, const int64_t* counts, const double* boxes, const int* periodic, const int* types, const int64_t* tags,
                   const double* x) {
    if (!have_ff) throw ApiError(RXA_ESTATE, "pair_coeff (the force field) must be set before the atoms");
    if (ns < 1 || ns > (1 << 18)) throw ApiError(RXA_EINVAL, "system count out of range");
    sync();
    std::vector<int> start(ns + 1, 0);
    std::vector<Cell> cells(ns);
    double dens = 0;
    for (int k = 0; k < ns; ++k) {
      if (counts[k] <= 0) throw ApiError(RXA_EINVAL, "system " + std::to_string(k) + " has no atoms");
      if (start[k] + counts[k] > (1LL << 30)) throw ApiError(RXA_EINVAL, "atom count out of range");
      start[k + 1] = start[k] + (int)counts[k];
      const double* b = boxes + 9 * (size_t)k;
      cells[k] = make_cell(b, b + 3, b + 6, periodic + 3 * (size_t)k);
      double d = counts[k] / cells[k].det;
      if (!(cells[k].per[0] && cells[k].per[1] && cells[k].per[2])) d = std::max(d, 0.12);
      dens = std::max(dens, d);
    }
    const int nn = start[ns];
    std::vector<int> ft(nn), sid(nn);
    std::vector<long long> tg(nn);
    std::vector<int> used;
    for (int k = 0; k < ns; ++k) {
      std::vector<long long> t2;
      for (int i = start[k]; i < start[k + 1]; ++i) {
        const int t = types[i];
        if (t < 1 || t >= (int)lmp2ff.size()) throw ApiError(RXA_EINVAL, "atom type " + std::to_string(t) + " outside the pair_coeff map");
        ft[i] = lmp2ff[t];
        tg[i] = tags[i];
        sid[i] = k;
        t2.push_back(tags[i]);
      }
      std::sort(t2.begin(), t2.end());
      for (size_t q = 1; q < t2.size(); ++q)
        if (t2[q] == t2[q - 1]) throw ApiError(RXA_EINVAL, "duplicate atom id " + std::to_string(t2[q]) + " in system " + std::to_string(k));
    }
    {
      std::vector<int> seen(ff.ntypes(), 0);
      for (int v : ft) if (v >= 0 && !seen[v]) { seen[v] = 1; used.push_back(v); }
      ff.require_bond_blocks(used);
    }
    packed = true;
    S = ns;
    pk_dens = dens;
    sys_start_h = start;
    cells_h = cells;
    cell = cells[0];
    have_box = true;
    n = nn;
    hist_count = 0;
    alloc_atoms();
    alloc_neighbors(estimate_neighbors());
    alloc_sublists(estimate_within(g.bond_cut + skin), g.hbond_cut > 0 ? estimate_within(g.hbond_cut + skin) : 1);
    alloc_bonds(maxb);
    qeq_h_valid = false;
    if (qeq_ready) alloc_qeq();
    md_ready = false;
    d_sys_start.alloc(S + 1); d_cells.alloc(S); d_ext.alloc(S); d_reb.alloc(S); d_rebany.alloc(1);
    d_pk.alloc(10 * (size_t)S); d_pki.alloc(4 * (size_t)S); d_rowv.alloc(4 * (size_t)n); d_nh_pk.alloc(32 * (size_t)S);
    d_disp_pk.alloc(S); d_sys_out.alloc((size_t)S * (kTerms + 6));
    zero(d_pki.p, 4 * (size_t)S * sizeof(int));
    h2d(d_sys_start.p, start.data(), (S + 1) * sizeof(int));
    h2d(d_cells.p, cells.data(), S * sizeof(Cell));
    h2d(st.sysid, sid.data(), n * sizeof(int));
    if ((size_t)S > pk_cap_h) {
      if (pki_h) HIPCHK(hipHostFree(pki_h));
      if (pk_h) HIPCHK(hipHostFree(pk_h));
      HIPCHK(hipHostMalloc((void**)&pki_h, 4 * (size_t)S * sizeof(int), hipHostMallocDefault));
      HIPCHK(hipHostMalloc((void**)&pk_h, 10 * (size_t)S * sizeof(double), hipHostMallocDefault));
      pk_cap_h = S;
    }
    if (!rebany_h) HIPCHK(hipHostMalloc((void**)&rebany_h, sizeof(int), hipHostMallocDefault));
    pq.S = S; pq.sys_start = d_sys_start.p; pq.sysid = st.sysid; pq.pk = d_pk.p; pq.pki = d_pki.p; pq.rowv = d_rowv.p;
    ext_h.clear();
    smx[0] = smx[1] = smx[2] = -1;
    shv_valid = false;
    perm_h.resize(n);
    for (int i = 0; i < n; ++i) perm_h[i] = i;
    h2d(d_perm.p, perm_h.data(), n * sizeof(int));
    h2d(a_type.p, ft.data(), n * sizeof(int));
    h2d(a_tag.p, tg.data(), n * sizeof(long long));
    host_types_mass.assign(n, 0.0);
    set_positions(x);
    sync();
  }

  // Neighbour lists of the systems flagged in d_reb (all of them when `all`): wrap, direct search, sub-lists.
  void build_neighbors_packed(bool all) {
    const double rs = list_cut() + skin;
    int want[3] = {1, 1, 1};
    std::vector<int> ext(4 * (size_t)S, 0);
    for (int k = 0; k < S; ++k)
      for (int d = 0; d < 3; ++d) {
        const Cell& c = cells_h[k];
        const int e = c.per[d] ? (int)std::floor(rs / c.h[d] + 0.5) : 0;
        if (e > 60) throw ApiError(RXA_EINVAL, "periodic cell of system " + std::to_string(k) + " too thin for the neighbour cut-off");
        ext[4 * (size_t)k + d] = e;
        want[d] = std::max(want[d], 1 + e);
      }
    if (ext != ext_h) {
      HIPCHK(hipMemcpyAsync(d_ext.p, ext.data(), ext.size() * sizeof(int), hipMemcpyHostToDevice, stream));
      sync();
      ext_h = ext;
    }
    if (want[0] != smx[0] || want[1] != smx[1] || want[2] != smx[2]) {
      const long long T = (2LL * want[0] + 1) * (2LL * want[1] + 1) * (2LL * want[2] + 1);
      if (T * S > (1LL << 26) || T > 32768) throw ApiError(RXA_EINVAL, "periodic cells too thin for the neighbour cut-off");
      a_shv.alloc((size_t)(T * S));
      for (int d = 0; d < 3; ++d) smx[d] = want[d];
      st.shv = a_shv.p; st.smx0 = smx[0]; st.smx1 = smx[1]; st.smx2 = smx[2]; st.shv_stride = (int)T;
      shv_valid = false;
    }
    hipEvent_t t0 = timer.begin(PH_NBR);
    if (!shv_valid) { launch_pk_shift_table(d_cells.p, S, a_shv.p, smx, stream); shv_valid = true; }
    if (all) launch_fill_int(d_reb.p, S, 1, stream);
    launch_pk_wrap(st, d_cells.p, d_reb.p, stream);
    timer.end(PH_NBR, t0);
    const double rb = g.bond_cut + skin, rh = g.hbond_cut > 0 ? g.hbond_cut + skin : 0.0;
    bool need_full = true;
    for (int attempt = 0; attempt < 8; ++attempt) {
      t0 = timer.begin(PH_NBR);
      launch_fill_int(st.flags, 8, 0, stream);
      if (need_full) launch_pk_nbr(st, d_cells.p, d_sys_start.p, d_ext.p, d_reb.p, rs, stream);
      launch_sublists(st, g, tab, rb, rh, stream);
      timer.end(PH_NBR, t0);
      int flags[8];
      HIPCHK(hipMemcpyAsync(flags, st.flags, sizeof(flags), hipMemcpyDeviceToHost, stream));
      sync();
      if (flags[0] > maxnb) {   // the skin list overflowed: grow it and list every system again (without a new wrap)
        alloc_neighbors((flags[0] + 31) / 32 * 32 + 32);
        launch_fill_int(d_reb.p, S, 1, stream);
        need_full = true;
        continue;
      }
      need_full = false;
      if (flags[4] <= maxbc && flags[5] <= maxhc) {
        nbr_valid = true;
        ++rebuilds;
        qeq_h_valid = false;
        return;
      }
      alloc_sublists(std::max(maxbc, (flags[4] + 31) / 32 * 32 + 32), std::max(maxhc, (flags[5] + 31) / 32 * 32 + 32));
    }
    throw ApiError(RXA_EINVAL, "neighbour list capacity did not converge");
  }

  // per-system energy terms, virial and thermostat state after a force call: out[s*(kTerms+6)], ke, temperature, iterations
  void system_results(double* energies, double* virial, double* thermo) {
    if (!packed) throw ApiError(RXA_ESTATE, "no packed systems (rxa_set_systems)");
    launch_pk_energy(st, d_sys_start.p, S, d_sys_out.p, stream);
    std::vector<double> out((size_t)S * (kTerms + 6)), nh;
    HIPCHK(hipMemcpyAsync(out.data(), d_sys_out.p, out.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (md_ready) {
      nh.resize(32 * (size_t)S);
      HIPCHK(hipMemcpyAsync(nh.data(), d_nh_pk.p, nh.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    HIPCHK(hipMemcpyAsync(pki_h, d_pki.p, 4 * (size_t)S * sizeof(int), hipMemcpyDeviceToHost, stream));
    sync();
    for (int k = 0; k < S; ++k) {
      if (energies) for (int t = 0; t < kTerms; ++t) energies[(size_t)k * kTerms + t] = out[(size_t)k * (kTerms + 6) + t];
      if (virial) for (int t = 0; t < 6; ++t) virial[(size_t)k * 6 + t] = out[(size_t)k * (kTerms + 6) + kTerms + t];
      if (thermo) {
        thermo[4 * (size_t)k] = md_ready ? nh[32 * (size_t)k + 21] : 0.0;
        thermo[4 * (size_t)k + 1] = md_ready ? nh[32 * (size_t)k + 20] : 0.0;
        thermo[4 * (size_t)k + 2] = pki_h[4 * k + 1];
        thermo[4 * (size_t)k + 3] = pki_h[4 * k + 3];
      }
    }
  }

  void set_atoms(long long nn, const int* types, const int64_t* tags, const double* x) {
    if (!have_ff) throw ApiError(RXA_ESTATE, "pair_coeff (the force field) must be set before the atoms");
    if (!have_box) throw ApiError(RXA_ESTATE, "the box must be set before the atoms");
    if (nn <= 0 || nn > (1LL << 30)) throw ApiError(RXA_EINVAL, "atom count out of range");
    sync();   // blocking copies below touch arrays that queued work may still use
    const bool realloc = (int)nn != n || !a_x.p || packed;
    if (packed) leave_packed();
    n = (int)nn;
    if (realloc) {
      hist_count = 0;   // a new atom set: no charge history to carry
      alloc_atoms();
      alloc_neighbors(estimate_neighbors());
      alloc_sublists(estimate_within(g.bond_cut + skin), g.hbond_cut > 0 ? estimate_within(g.hbond_cut + skin) : 1);
      alloc_bonds(maxb);
      qeq_h_valid = false;
      hist_count = 0;
      if (qeq_ready) alloc_qeq();
      md_ready = false;
    }
    std::vector<int> ft(n);
    std::vector<long long> tg(n);
    std::vector<int> used;
    for (int i = 0; i < n; ++i) {
      const int t = types[i];
      if (t < 1 || t >= (int)lmp2ff.size()) throw ApiError(RXA_EINVAL, "atom type " + std::to_string(t) + " outside the pair_coeff map");
      ft[i] = lmp2ff[t];
      tg[i] = tags[i];
    }
    {
      std::vector<int> seen(ff.ntypes(), 0);
      for (int v : ft) if (v >= 0 && !seen[v]) { seen[v] = 1; used.push_back(v); }
      ff.require_bond_blocks(used);
    }
    {
      std::vector<long long> s = tg;
      std::sort(s.begin(), s.end());
      for (size_t k = 1; k < s.size(); ++k)
        if (s[k] == s[k - 1]) throw ApiError(RXA_EINVAL, "duplicate atom id " + std::to_string(s[k]));
    }
    // the same atoms in a new order (e.g. LAMMPS re-sorted or re-neighboured): carry charges, charge history and velocities by tag
    if (!realloc && !perm_h.empty()) {
      std::vector<long long> old_tag(n);
      d2h(old_tag.data(), a_tag.p, n * sizeof(long long));
      std::unordered_map<long long, int> where;
      where.reserve(2 * (size_t)n);
      for (int i = 0; i < n; ++i) where[old_tag[i]] = i;
      std::vector<int> src(n);
      bool same = true;
      for (int i = 0; i < n && same; ++i) {
        auto it = where.find(tg[i]);
        if (it == where.end()) same = false; else src[i] = it->second;
      }
      if (same) {
        h2d(d_src.p, src.data(), n * sizeof(int));
        permute(d_src.p);
      } else {
        reset_history();
        md_ready = false;
      }
    }
    perm_h.resize(n);
    for (int i = 0; i < n; ++i) perm_h[i] = i;
    HIPCHK(hipMemcpyAsync(d_perm.p, perm_h.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(a_type.p, ft.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(a_tag.p, tg.data(), n * sizeof(long long), hipMemcpyHostToDevice, stream));
    host_types_mass.assign(n, 0.0);
    set_positions(x);
    sync();
  }

  void set_positions(const double* x) {
    std::vector<double> xs(n), ys(n), zs(n);
    for (int i = 0; i < n; ++i) {
      const int u = perm_h[i];
      xs[i] = x[3 * u]; ys[i] = x[3 * u + 1]; zs[i] = x[3 * u + 2];
      if (!std::isfinite(xs[i]) || !std::isfinite(ys[i]) || !std::isfinite(zs[i])) throw ApiError(RXA_EINVAL, "non-finite coordinate of atom " + std::to_string(u));
    }
    HIPCHK(hipMemcpyAsync(a_x.p, xs.data(), n * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(a_y.p, ys.data(), n * sizeof(double), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(a_z.p, zs.data(), n * sizeof(double), hipMemcpyHostToDevice, stream));
    launch_pos4(st, stream);
    sync();
    nbr_valid = false;
    qeq_h_valid = false;
  }

  // ---------------------------------------------------------------------------------------------------------------
  void build_neighbors() {
    if (packed) { build_neighbors_packed(true); return; }
    const double rs = list_cut() + skin;
    cell_mode = true;
    int nc[3];
    for (int d = 0; d < 3; ++d) {
      nc[d] = (int)std::floor(cell.h[d] / rs);
      if (!cell.per[d] || nc[d] < 3) cell_mode = false;
    }
    if (cell_mode && (long long)nc[0] * nc[1] * nc[2] > (1LL << 24)) cell_mode = false;
    if (!cell_mode && n > 20000) throw ApiError(RXA_EINVAL, "a cell narrower than 3 list cut-offs along some direction is limited to 20000 atoms");
    if (cell_mode) {
      for (int d = 0; d < 3; ++d) grid.nc[d] = nc[d];
      if ((size_t)grid.ncells() > b_cstart.n) { b_cstart.alloc(grid.ncells()); b_cend.alloc(grid.ncells()); }
      bins.cstart = b_cstart.p;
      bins.cend = b_cend.p;
    } else {
      for (int d = 0; d < 3; ++d) sm[d] = cell.per[d] ? (int)std::ceil(rs / cell.h[d]) + 1 : 0;
      for (int d = 0; d < 3; ++d)
        if (sm[d] > 100) throw ApiError(RXA_EINVAL, "periodic cell too thin for the neighbour cut-off");
    }
    // shift-vector table of every code the lists can hold (cell stencil: one image in each direction)
    int want[3];
    for (int d = 0; d < 3; ++d) want[d] = cell_mode ? 1 : sm[d];
    if (want[0] != smx[0] || want[1] != smx[1] || want[2] != smx[2]) {
      const long long m = (2LL * want[0] + 1) * (2LL * want[1] + 1) * (2LL * want[2] + 1);
      if (m > (1LL << 22)) throw ApiError(RXA_EINVAL, "periodic cell too thin for the neighbour cut-off");
      a_shv.alloc((size_t)m);
      for (int d = 0; d < 3; ++d) smx[d] = want[d];
      st.shv = a_shv.p; st.smx0 = smx[0]; st.smx1 = smx[1]; st.smx2 = smx[2];
    }
    hipEvent_t t0 = timer.begin(PH_NBR);
    launch_shift_table(cell, a_shv.p, smx, stream);
    launch_wrap(st, cell, stream);
    timer.end(PH_NBR, t0);
    const double rb = g.bond_cut + skin, rh = g.hbond_cut > 0 ? g.hbond_cut + skin : 0.0;
    bool need_full = true;
    for (int attempt = 0; attempt < 8; ++attempt) {
      t0 = timer.begin(PH_NBR);
      launch_fill_int(st.flags, 8, 0, stream);
      bool sorted = false;
      if (need_full) {
        if (cell_mode) {
          launch_bin(st, cell, grid, bins, stream);
          if (sort_atoms) {   // atoms into the order of their cells
            permute(bins.sidx);
            launch_bin_identity(n, bins, stream);
            launch_pos4(st, stream);
            sorted = true;
          }
          launch_nbr_cells(st, cell, grid, bins, rs, stream);
        } else {
          launch_nbr_brute(st, cell, sm, rs, stream);
        }
      }
      launch_sublists(st, g, tab, rb, rh, stream);
      timer.end(PH_NBR, t0);
      int flags[8];
      if (sorted) HIPCHK(hipMemcpyAsync(perm_h.data(), d_perm.p, n * sizeof(int), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipMemcpyAsync(flags, st.flags, sizeof(flags), hipMemcpyDeviceToHost, stream));
      sync();
      if (flags[0] > maxnb) {   // the skin list overflowed: grow it and redo everything
        alloc_neighbors((flags[0] + 31) / 32 * 32 + 32);
        need_full = true;
        continue;
      }
      need_full = false;
      if (flags[4] <= maxbc && flags[5] <= maxhc) {
        nbr_valid = true;
        ++rebuilds;
        qeq_h_valid = false;
        return;
      }
      alloc_sublists(std::max(maxbc, (flags[4] + 31) / 32 * 32 + 32), std::max(maxhc, (flags[5] + 31) / 32 * 32 + 32));
    }
    throw ApiError(RXA_EINVAL, "neighbour list capacity did not converge");
  }

  // Forces and energies of the current positions and charges. The bonded chain is queued on the bonded stream after a fork
  // from the main stream: bond list and uncorrected BO, the capacity flags (read back through pinned memory), BO correction and
  // the per-atom terms (valid whenever the lists fit); the main stream meanwhile runs the charge solve and the non-bonded pass.
  // Only the group-size dependent angle kernels wait for the flags. An overflow drains both streams, grows the lists and redoes
  // the chain. The gather on the main stream joins the two.
  // an event on the main stream after which the bonded chain may start (set by the MD loop for a step that keeps its lists:
  // the positions are final there, before the charge matrix of the step is queued); otherwise the fork is recorded here
  hipEvent_t fork_after = nullptr;
  void bonded_begin() {
    const hipStream_t b = bstream();
    if (b != stream) {
      if (!fork_after) HIPCHK(hipEventRecord(ev_fork, stream));
      HIPCHK(hipStreamWaitEvent(b, fork_after ? fork_after : ev_fork, 0));
    }
    fork_after = nullptr;
    hipEvent_t t0 = timer.begin(PH_BONDS);
    launch_fill_int(st.flags, 8, 0, b);
    launch_bond_build(prec, st, cell, g, tab, b);
    timer.end(PH_BONDS, t0);
    t0 = timer.begin(PH_BOCORR);
    launch_bo_correct(prec, st, g, tab, b);
    timer.end(PH_BOCORR, t0);
    HIPCHK(hipMemcpyAsync(flags_h, st.flags, 8 * sizeof(int), hipMemcpyDeviceToHost, b));
    HIPCHK(hipEventRecord(ev_flags, b));
    bonded_pending = true;
    t0 = timer.begin(PH_TERMS);
    launch_atom_terms(prec, st, g, tab, b);
    timer.end(PH_TERMS, t0);
  }
  // queues the second half of the bonded chain as soon as the capacity flags are back, without waiting for them: called between
  // the launches of a CG batch, so the angle kernels need not wait for the host to finish queueing the batch
  bool bonded_pending = false;
  void bonded_poll() {
    if (bonded_pending && hipEventQuery(ev_flags) == hipSuccess) bonded_end();
  }
  void bonded_end() {
    for (int attempt = 0;; ++attempt) {
      HIPCHK(hipEventSynchronize(ev_flags));
      if (flags_h[1] <= maxb) {
        if (flags_h[3]) throw ApiError(RXA_EDEVICE, "internal: a bond has no mirror slot");
        bmax = flags_h[7];
        if (bmax > 64) throw ApiError(RXA_EINVAL, "an atom has more than 64 bonds");
        break;
      }
      if (attempt >= 5) throw ApiError(RXA_EINVAL, "bond list capacity did not converge");
      sync();
      alloc_bonds(std::max(maxb, (flags_h[1] + 3) / 4 * 4 + 4));
      bonded_begin();
    }
    const hipStream_t b = bstream();
    hipEvent_t t0 = timer.begin(PH_TERMS);
    launch_angle_terms(prec, st, g, tab, std::min(64, bmax * g_mult), b);
    timer.end(PH_TERMS, t0);
    t0 = timer.begin(PH_DBOND);
    launch_bond_derivatives(st, b);
    timer.end(PH_DBOND, t0);
    if (b != stream) HIPCHK(hipEventRecord(ev_bonded, b));
    bonded_pending = false;
  }
  void nonbonded_launch() {
    hipEvent_t t0 = timer.begin(PH_NONB);
    // the charge-matrix rows of these positions already list the non-bonded pairs in skin-list order
    const bool rows = qeq_h_valid && qb.keyed && qb.key && g.q_swb >= g.nonb_cut && nb_rows;
    if (rows) launch_nonbonded_rows(prec, st, g, tab, qb, stream);
    else launch_nonbonded(prec, st, cell, g, tab, stream);
    timer.end(PH_NONB, t0);
  }
  void join_gather(bool energies) {
    if (bstream() != stream) HIPCHK(hipStreamWaitEvent(stream, ev_bonded, 0));
    hipEvent_t t0 = timer.begin(PH_DBOND);
    if (st.peratom) launch_vatom_ref(st, stream);   // reads the non-bonded half pairs in vatom, before the gather adds to them
    launch_gather(st, stream);
    if (st.peratom) launch_eatom_ref(st, stream);
    if (energies) launch_energy_reduce(st, red.p, stream);   // the per-atom arrays hold the energies either way
    timer.end(PH_DBOND, t0);
  }
  void forces(bool rebuild, bool energies = true) {
    if (!nbr_valid || rebuild) build_neighbors();
    bonded_begin();
    nonbonded_launch();
    bonded_end();
    join_gather(energies);
  }

  bool peratom_valid = false;   // the last force call of rxa_compute left per-atom tallies
  bool census_valid = false;    // the last force call of rxa_compute recorded the decision census
  void compute(int flags, double* energies, double* grad, double* virial) {
    if (!n) throw ApiError(RXA_ESTATE, "no atoms");
    st.peratom = (flags & RXA_COMPUTE_PERATOM) ? 1 : 0;
    st.census = (flags & RXA_COMPUTE_CENSUS) ? 1 : 0;
    peratom_valid = false;
    census_valid = false;
    try {
      if (st.peratom) { sync(); alloc_peratom(); }
      forces((flags & RXA_COMPUTE_REBUILD) != 0);
      if (st.census) launch_census_nb(st, g, stream);
    } catch (...) {
      st.peratom = 0;
      st.census = 0;
      throw;
    }
    peratom_valid = st.peratom != 0;
    census_valid = st.census != 0;
    st.peratom = 0;
    st.census = 0;
    double out[kTerms + 6];
    HIPCHK(hipMemcpyAsync(out, red.p, sizeof(out), hipMemcpyDeviceToHost, stream));
    std::vector<double> gx, gy, gz;
    if (grad) {
      gx.resize(n); gy.resize(n); gz.resize(n);
      HIPCHK(hipMemcpyAsync(gx.data(), st.gx, n * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipMemcpyAsync(gy.data(), st.gy, n * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipMemcpyAsync(gz.data(), st.gz, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    }
    sync();
    if (energies) for (int k = 0; k < kTerms; ++k) energies[k] = out[k];
    if (virial) for (int k = 0; k < 6; ++k) virial[k] = out[kTerms + k];
    if (grad)
      for (int i = 0; i < n; ++i) {
        const int u = perm_h[i];
        grad[3 * u] = gx[i]; grad[3 * u + 1] = gy[i]; grad[3 * u + 2] = gz[i];
      }
  }

  // per-atom energy and virial of the last rxa_compute call made with RXA_COMPUTE_PERATOM, in the caller's order
  void peratom_out(double* eatom, double* vatom) {
    if (!peratom_valid) throw ApiError(RXA_ESTATE, "no per-atom tallies: call rxa_compute with RXA_COMPUTE_PERATOM first");
    std::vector<double> e(n), v(6 * (size_t)n);
    HIPCHK(hipMemcpyAsync(e.data(), st.eref, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipMemcpyAsync(v.data(), st.vref, 6 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    for (int i = 0; i < n; ++i) {
      const int u = perm_h[i];
      if (eatom) eatom[u] = e[i];
      if (vatom)
        for (int c = 0; c < 6; ++c) vatom[6 * (size_t)u + c] = v[(size_t)c * n + i];
    }
  }

  // ---------------------------------------------------------------------------------------------------------------
  void alloc_qeq_matrix() {
    qb.narrow = n <= 65535 ? 1 : 0;
    q_int.alloc(qb.narrow ? (size_t)n : (size_t)n * maxnb + n);
    qb.cnt = q_int.p;
    qb.col = qb.narrow ? nullptr : q_int.p + n;
    if (qb.narrow) { q_u16.alloc((size_t)n * maxnb); qb.col16 = q_u16.p; } else { q_u16.release(); qb.col16 = nullptr; }
    // FP32 values and their FP32 remainders (mixed / FP32), or FP64 values in the double pool (FP64)
    if (qb.fp32) {
      q_f32.alloc(2 * (size_t)n * maxnb);
      qb.val32 = q_f32.p;
      qb.lo32 = q_f32.p + (size_t)n * maxnb;
    } else {
      q_f32.release();
      qb.val32 = nullptr;
      qb.lo32 = nullptr;
    }
    q_fp32_alloc = qb.fp32;
    q_key.alloc((size_t)n * maxnb);
    qb.key = q_key.p;
    // val shares the double pool
    alloc_qeq_doubles();
  }
  void alloc_qeq_doubles() {
    // sub-arrays start on 256-byte boundaries (the CG reads (s, t) pairs as 16-byte loads)
    auto rnd = [](size_t k) { return (k + 31) & ~(size_t)31; };
    const size_t nmv = 4 * (size_t)qeq_mv_partials(n), nqb = 4 * (size_t)qeq_build_blocks(n);
    const size_t nval = qb.fp32 ? 0 : (size_t)n * maxnb;   // FP64 values (FP64 mode only)
    const size_t sz = (nval ? rnd(nval) : 0) + 3 * rnd(n) + 6 * rnd(2 * (size_t)n) + rnd(4 * kCgBlocks) + 2 * rnd(nmv) + rnd(nqb) +
                      2 * rnd(4 * (size_t)n) + rnd(n) + rnd(2 * (size_t)n) + rnd(3 * kCgBlocks);
    std::vector<double> keep;
    const bool had_hist = q_dbl.p && hist_count > 0 && q_dbl_n == n;   // history of the same atoms only
    if (had_hist) {
      keep.resize(8 * (size_t)n);
      d2h(keep.data(), qb.s_hist, 4 * (size_t)n * sizeof(double));
      d2h(keep.data() + 4 * (size_t)n, qb.t_hist, 4 * (size_t)n * sizeof(double));
    }
    q_dbl.alloc(sz);
    q_dbl_n = n;
    zero(q_dbl.p, sz * sizeof(double));
    double* p = q_dbl.p;
    auto take = [&](size_t k) { double* r = p; p += rnd(k); return r; };
    qb.val = nval ? take(nval) : nullptr;
    qb.eta = take(n);
    qb.hinv = take(n);
    qb.chi = take(n);
    qb.b = take(2 * (size_t)n);
    qb.x = take(2 * (size_t)n);
    qb.r = take(2 * (size_t)n);
    qb.d = take(2 * (size_t)n);
    qb.p = take(2 * (size_t)n);
    qb.q = take(2 * (size_t)n);
    qb.partial = take(4 * kCgBlocks);
    qb.partial_mv = take(nmv);
    qb.partial_rs = take(nmv);
    qb.partial_qb = take(nqb);
    qb.s_hist = take(4 * (size_t)n);
    qb.t_hist = take(4 * (size_t)n);
    qb.lorow = take(n);
    qb.x0 = take(2 * (size_t)n);
    qb.partial_gate = take(3 * kCgBlocks);
    static_assert(2 * sizeof(int) <= sizeof(double), "");
    if (had_hist) {
      h2d(qb.s_hist, keep.data(), 4 * (size_t)n * sizeof(double));
      h2d(qb.t_hist, keep.data() + 4 * (size_t)n, 4 * (size_t)n * sizeof(double));
    } else {
      hist_count = 0;
    }
  }
  void alloc_qeq() { alloc_qeq_matrix(); }

  // A new atom set starts the stock extrapolation from a zero history: the first solve starts from zero and the later ones
  // extrapolate over the solutions so far and zeros, exactly as for atoms the stock fix has just created.
  void reset_history() {
    hist_count = 0;
    qeq_started = false;
    if (qb.s_hist && q_dbl_n == n) {
      HIPCHK(hipMemsetAsync(qb.s_hist, 0, 4 * (size_t)n * sizeof(double), stream));
      HIPCHK(hipMemsetAsync(qb.t_hist, 0, 4 * (size_t)n * sizeof(double), stream));
    }
  }

  void qeq_setup(double swa, double swb, const double* chi, const double* eta, const double* gamma) {
    if (!have_ff) throw ApiError(RXA_ESTATE, "pair_coeff must precede the charge fix setup");
    if (!(swb > swa) || !(swb > 0)) throw ApiError(RXA_EINVAL, "charge equilibration taper: need swb > swa");
    if (swb > ff.nonb_cut + 1e-12 && ps.shellcheck)
      throw ApiError(RXA_EINVAL, "fix qeq/reaxff/metal: taper radius exceeds nonb_cut (the stock fix would silently truncate the matrix)");
    const int nt = ff.ntypes();
    q_chi_t.assign(nt, 0.0); q_eta_t.assign(nt, 1.0); q_gam_t.assign(nt, 1.0);
    std::vector<int> have(nt, 0);
    for (size_t t = 1; t < lmp2ff.size(); ++t) {
      const int f = lmp2ff[t];
      if (f < 0) continue;
      if (have[f] && (q_chi_t[f] != chi[t] || q_eta_t[f] != eta[t] || q_gam_t[f] != gamma[t]))
        throw ApiError(RXA_EINVAL, "two LAMMPS types of one element have different charge parameters");
      q_chi_t[f] = chi[t]; q_eta_t[f] = eta[t]; q_gam_t[f] = gamma[t];
      have[f] = 1;
    }
    std::vector<double> shld((size_t)nt * nt);
    for (int a = 0; a < nt; ++a) for (int b = 0; b < nt; ++b) shld[(size_t)a * nt + b] = std::pow(q_gam_t[a] * q_gam_t[b], -1.5);
    h2d(d_qchi.p, q_chi_t.data(), nt * sizeof(double));
    h2d(d_qeta.p, q_eta_t.data(), nt * sizeof(double));
    h2d(d_qshld.p, shld.data(), shld.size() * sizeof(double));
    const bool first = !qeq_ready;
    q_swa = swa;
    q_swb = swb;
    qeq_ready = true;
    make_globals();
    if (n && (first || !qb.val)) { alloc_qeq(); }
    nbr_valid = false;
    qeq_h_valid = false;
  }

  // The charge matrix of the current positions. start (one system): the build also forms the solve's start from the stock
  // guess (the history extrapolation, queued before it), so the next qeq_begin skips its own start pass (qeq_started).
  bool qeq_started = false;
  void qeq_matrix(bool start = false) {
    if (!qeq_ready) throw ApiError(RXA_ESTATE, "charge equilibration is not set up");
    if (!nbr_valid) build_neighbors();
    if (q_fp32_alloc != qb.fp32) {   // the precision changed since the arrays were sized
      sync();
      alloc_qeq_matrix();
      qeq_h_valid = false;
    }
    if (!qeq_h_valid) {
      hipEvent_t t0 = timer.begin(PH_QEQ_H);
      qb.keyed = (2 * smx[0] + 1) * (2 * smx[1] + 1) * (2 * smx[2] + 1) <= 32768 ? 1 : 0;
      const bool fuse = start && !packed;
      if (fuse) launch_qeq_guess(st, qb, hist_count > 0 ? 1 : 0, stream);
      launch_qeq_build(st, cell, g, tab, qb, fuse ? 1 : 0, stream);
      timer.end(PH_QEQ_H, t0);
      qeq_h_valid = true;
      qeq_started = fuse;
    }
  }

  // runs CG iterations until both systems stopped; batch launches, one readback per batch
  // The CG's per-iteration scalars (device), the packed systems' per-iteration flags (device) and their pinned host copies, all
  // sized for iterations 0 .. cap + 1 (cap >= imax). Only these grow here: the matrix, the diagonal and the per-atom vectors
  // are untouched, so a built matrix stays valid. Queued work may still read the old buffers or copy them into the old pinned
  // memory, so the streams are drained before anything is freed.
  void ensure_cg_cap(int imax) {
    if (imax + 2 > qeq_imax_cap || !q_scal.p) {
      sync();
      const int cap = std::max(imax + 2, 2048);
      q_scal.alloc(qeq_scal_doubles(cap));
      zero(q_scal.p, q_scal.n * sizeof(double));
      d_cg_any.alloc((size_t)cap + 2);
      if (cgf_h) HIPCHK(hipHostFree(cgf_h));
      if (cgs_h) HIPCHK(hipHostFree(cgs_h));
      cgf_h = nullptr;
      cgs_h = nullptr;
      HIPCHK(hipHostMalloc((void**)&cgf_h, ((size_t)cap + 2) * sizeof(int), hipHostMallocDefault));
      HIPCHK(hipHostMalloc((void**)&cgs_h, qeq_scal_doubles(cap) * sizeof(double), hipHostMallocDefault));
      qeq_imax_cap = cap;
    }
    qb.scal = q_scal.p;
    pq.any = d_cg_any.p;
  }
  int cg_guess = 8;
  int cg_it = 0, cg_batch = 0;
  // refinement: the last batch ended with update(cg_it) and a check of slot cg_it, so the next launch is the product of
  // iteration cg_it + 1 in its checking form
  bool cg_checked = false;
  // activity flags, restart records and sig of iteration slot k in the read-back scalars (the layout of QeqBuffers::scal)
  const int* cg_act(int k) const { return reinterpret_cast<const int*>(cgs_h + 8 + kScalSlot * (size_t)k + 2); }
  const int* cg_rst(int k) const { return reinterpret_cast<const int*>(cgs_h + 8 + kScalSlot * (size_t)k + 3); }
  double cg_sig(int k, int sys) const { return cgs_h[8 + kScalSlot * (size_t)k + sys]; }
  // launches the next batch of iterations, then the asynchronous readback of the flags and scalars (pinned) and its event
  void cg_launch_batch(int imax) {
    if (packed) {   // iteration it sets any[it]: some system active for the next one
      for (int b = 0; b < cg_batch && cg_it < imax; ++b) {
        hipEvent_t t0 = timer.begin(PH_QEQ_SPMV);
        launch_pk_qeq_iteration(st, qb, pq, cg_it, imax, cg_tol, stream);
        timer.end(PH_QEQ_SPMV, t0);
        ++cg_it;
        bonded_poll();
      }
      if (qb.refine) {   // stopped solves: true residual, restart or verify (sets any[cg_it - 1] on a restart)
        hipEvent_t t0 = timer.begin(PH_QEQ_CG);
        launch_pk_qeq_restart(st, qb, pq, cg_it - 1, imax, cg_tol, stream);
        timer.end(PH_QEQ_CG, t0);
      }
      HIPCHK(hipMemcpyAsync(cgf_h, d_cg_any.p, (size_t)cg_it * sizeof(int), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipEventRecord(ev_cg, stream));
      return;
    }
    for (int b = 0; b < cg_batch && cg_it < imax; ++b) {
      hipEvent_t t0 = timer.begin(PH_QEQ_CG);
      if (!cg_checked) launch_qeq_update(st, qb, cg_it, stream);
      timer.end(PH_QEQ_CG, t0);
      ++cg_it;
      t0 = timer.begin(PH_QEQ_SPMV);
      if (cg_checked) launch_qeq_mvdir_chk(st, qb, cg_it, imax, cg_tol, stream);
      else launch_qeq_mvdir(st, qb, cg_it, imax, cg_tol, stream);
      timer.end(PH_QEQ_SPMV, t0);
      cg_checked = false;
      bonded_poll();
    }
    int slots = cg_it;
    if (qb.refine) {   // close the batch with the update of iteration cg_it and a refinement check in slot cg_it
      hipEvent_t t0 = timer.begin(PH_QEQ_CG);
      if (!cg_checked) launch_qeq_update(st, qb, cg_it, stream);
      launch_qeq_restart(st, qb, cg_it, imax, cg_tol, stream);
      timer.end(PH_QEQ_CG, t0);
      cg_checked = true;
      slots = cg_it + 1;
    }
    HIPCHK(hipMemcpyAsync(cgs_h, qb.scal, (8 + kScalSlot * (size_t)slots) * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipEventRecord(ev_cg, stream));
    ++cg_batches;
  }
  double cg_tol = 1e-6;
  // Charge solve in two halves so that other work can be queued in between: qeq_begin queues the matrix (if not current), the
  // initial residuals and the first batch of iterations (launch order mvdir(1) [update(1) mvdir(2)] ...; mvdir(it) decides
  // whether iteration it runs); qeq_end reads each batch's flags back and queues more until both systems stopped, then the
  // charges. The first batch is as long as the previous solve (warm-started solves of consecutive steps need about the same
  // number of iterations), later ones two iterations; iterations past convergence exit at once on the device flags.
  void qeq_begin(int imax, double tol) {
    if (!n) throw ApiError(RXA_ESTATE, "no atoms");
    if (imax < 1) throw ApiError(RXA_EINVAL, "charge equilibration: maxiter must be >= 1");
    ensure_cg_cap(imax);
    qeq_matrix(true);
    if (packed) {
      launch_fill_int(d_cg_any.p, (int)d_cg_any.n, 0, stream);
      launch_pk_qeq_init(st, qb, pq, imax, tol, hist_count > 0 ? 1 : 0, stream);
      cg_tol = tol;
      cg_it = 1;
      cg_batch = std::max(1, cg_guess);
      cg_launch_batch(imax);
      return;
    }
    if (qeq_started) launch_qeq_init_fused(st, qb, imax, tol, stream);   // the build formed the start
    else launch_qeq_init(st, qb, imax, tol, hist_count > 0 ? 1 : 0, stream);
    qeq_started = false;
    cg_start(imax, tol);
  }
  void cg_start(int imax, double tol) {
    cg_tol = tol;
    cg_it = 1;
    cg_checked = false;
    cg_batch = std::max(1, cg_guess);
    hipEvent_t t0 = timer.begin(PH_QEQ_SPMV);
    launch_qeq_mvdir(st, qb, 1, imax, tol, stream);
    timer.end(PH_QEQ_SPMV, t0);
    cg_launch_batch(imax);
  }
  void cg_wait_end(int imax) {
    while (true) {
      HIPCHK(hipEventSynchronize(ev_cg));
      if (qb.refine) {   // done when the batch's check found both true residuals within tol (or the budget is spent)
        const int* r = cg_rst(cg_it);
        if ((r[0] == 0 && r[1] == 0) || cg_it >= imax) break;
      } else {
        const int* a = cg_act(cg_it - 1);
        if ((!a[0] && !a[1]) || cg_it >= imax) break;
      }
      cg_batch = 2;
      cg_launch_batch(imax);
    }
    for (int s = 0; s < 2; ++s) {
      int done = 0;   // iterations the system ran (over every restart)
      for (int k = 0; k < cg_it; ++k) done += cg_act(k)[s] == 1;
      last_iters[s] = done + 1;   // the stock loop index at exit
      // the relative residual the solve stopped on: the true one of the last check under refinement, else the recursion's
      last_rel[s] = std::sqrt(qb.refine ? cg_sig(cg_it, s) : cg_sig(last_iters[s] - 1, s)) / cgs_h[s];
    }
    cg_guess = std::max(last_iters[0], last_iters[1]) - 1;
  }
  // CG iterations on the current initial residuals until both systems stopped (the raw solve of rxa_qeq_solve)
  void qeq_iterate(int imax, double tol) {
    cg_start(imax, tol);
    cg_wait_end(imax);
  }
  void qeq_end(int imax, int strict) {
    if (packed) {
      while (true) {
        HIPCHK(hipEventSynchronize(ev_cg));
        if (!cgf_h[cg_it - 1] || cg_it >= imax) break;
        cg_batch = 2;
        cg_launch_batch(imax);
      }
      int done = 0;   // iterations the slowest system ran
      for (int k = 0; k < cg_it; ++k) if (cgf_h[k]) done = k + 1;
      cg_guess = std::max(1, done);
      launch_pk_qeq_finish(st, qb, pq, stream);
      HIPCHK(hipMemcpyAsync(pki_h, d_pki.p, 4 * (size_t)S * sizeof(int), hipMemcpyDeviceToHost, stream));
      HIPCHK(hipMemcpyAsync(pk_h, d_pk.p, 10 * (size_t)S * sizeof(double), hipMemcpyDeviceToHost, stream));
      hist_count = std::min(hist_count + 1, 4);
      ++qeq_calls;
      pk_collect_pending = true;
      if (strict) {
        sync();
        pk_collect();
        for (int q = 0; q < 2; ++q)
          if (!(last_rel[q] <= cg_tol))
            throw ApiError(RXA_ENOCONV, "fix qeq/reaxff/metal strict: CG did not converge in some system after " + std::to_string(last_iters[q]) + " iterations");
      }
      return;
    }
    cg_wait_end(imax);
    launch_qeq_finish(st, qb, stream);
    hist_count = std::min(hist_count + 1, 4);
    qeq_iter_sum += 0.5 * (last_iters[0] + last_iters[1]);
    ++qeq_calls;
    if (strict)
      for (int s = 0; s < 2; ++s)
        if (!(last_rel[s] <= cg_tol))
          throw ApiError(RXA_ENOCONV, "fix qeq/reaxff/metal strict: CG did not converge after " + std::to_string(last_iters[s]) + " iterations");
  }
  void qeq_charges(int imax, double tol, int strict) {
    qeq_begin(imax, tol);
    qeq_end(imax, strict);
  }
  // packed systems: the largest iteration count and relative residual over the systems of the last solve (after a sync)
  bool pk_collect_pending = false;
  void pk_collect() {
    if (!pk_collect_pending) return;
    pk_collect_pending = false;
    for (int q = 0; q < 2; ++q) {
      int it = 0;
      double rel = 0;
      for (int k = 0; k < S; ++k) {
        it = std::max(it, pki_h[4 * k + 2 * q + 1]);
        rel = std::max(rel, std::sqrt(pk_h[10 * k + 4 * q + 1]) / pk_h[10 * k + 4 * q]);
      }
      last_iters[q] = it;
      last_rel[q] = rel;
    }
    qeq_iter_sum += 0.5 * (last_iters[0] + last_iters[1]);
  }

  // the device charges in the caller's order
  void charges_out(double* q) {
    std::vector<double> qi(n);
    sync();
    d2h(qi.data(), st.q, n * sizeof(double));
    for (int i = 0; i < n; ++i) q[perm_h[i]] = qi[i];
  }

  // ---------------------------------------------------------------------------------------------------------------
  void md_setup(const double* v, const double* mass_by_type, double dt, double tt, double tdamp, double sk, int qeq_every, int qimax, double qtol) {
    md_setup_buffers(v, mass_by_type, dt, tt, tdamp, sk, qeq_every, qimax, qtol);
    // initial forces (and charges)
    build_neighbors();
    launch_md_ref_positions(st, md, stream);
    if (md_qeq_every > 0) qeq_charges(md_qeq_imax, md_qeq_tol, 0);
    forces(false);
    sync();
  }
  // velocities, masses, thermostat chains and the run's parameters (the common part of rxa_md_setup and rxa_md_resume)
  void md_setup_buffers(const double* v, const double* mass_by_type, double dt, double tt, double tdamp, double sk, int qeq_every,
                        int qimax, double qtol) {
    if (!n) throw ApiError(RXA_ESTATE, "no atoms");
    sync();
    md_buf.alloc(8 * (size_t)n + 64 + 256);
    double* p = md_buf.p;
    md.vx = p; p += n; md.vy = p; p += n; md.vz = p; p += n; md.mass = p; p += n;
    md.x0 = p; p += n; md.y0 = p; p += n; md.z0 = p; p += n;
    p += n;
    md.nh = p; p += 64;
    md.partial = p;
    std::vector<double> vx(n), vy(n), vz(n), m(n);
    std::vector<int> ty(n);
    d2h(ty.data(), st.type, n * sizeof(int));
    // masses by LAMMPS type: map through the host copy of the LAMMPS types kept at set_atoms
    for (int i = 0; i < n; ++i) {
      const int u = perm_h[i];
      vx[i] = v[3 * u]; vy[i] = v[3 * u + 1]; vz[i] = v[3 * u + 2];
      m[i] = mass_by_type[lmp_types[u]];
      if (!(m[i] > 0)) throw ApiError(RXA_EINVAL, "non-positive mass");
    }
    h2d(md.vx, vx.data(), n * sizeof(double));
    h2d(md.vy, vy.data(), n * sizeof(double));
    h2d(md.vz, vz.data(), n * sizeof(double));
    h2d(md.mass, m.data(), n * sizeof(double));
    std::vector<double> nh(64, 0.0);
    nh[13] = tt; nh[14] = 1.0 / tdamp; nh[15] = dt; nh[16] = 3.0 * n - 3.0; nh[17] = 0.0019872067; nh[18] = 48.88821291 * 48.88821291;
    nh[19] = 1.0;
    md_nve = !(tdamp > 0);   // t_damp <= 0: no thermostat (NVE)
    nh[22] = md_nve ? 1.0 : 0.0;
    h2d(md.nh, nh.data(), nh.size() * sizeof(double));
    if (packed) {   // one chain per system, the same layout
      std::vector<double> nhp(32 * (size_t)S, 0.0);
      for (int k = 0; k < S; ++k) {
        double* h = nhp.data() + 32 * (size_t)k;
        for (int q = 13; q < 23; ++q) h[q] = nh[q];
        h[16] = 3.0 * (sys_start_h[k + 1] - sys_start_h[k]) - 3.0;
      }
      h2d(d_nh_pk.p, nhp.data(), nhp.size() * sizeof(double));
    }
    md_dt = dt;
    skin = sk;
    md_qeq_every = qeq_every; md_qeq_imax = qimax; md_qeq_tol = qtol;
    md_ready = true;
    md_steps = 0;
    v_scale_pending = false;
  }
  std::vector<int> lmp_types;

  // Autotuner: launch configurations timed on the running MD (each trial advances the run by 2 + 2 x `steps` steps; the
  // configurations give the same results, so only the speed changes). Coordinate descent over: lanes per atom of the
  // slot-parallel per-atom kernels (1, 2, 4, 8), the angle kernels' group size (the bond count's power of two, or twice it),
  // the two-stream overlap of the bonded chain with the charge solve, and the non-bonded pass over the charge-matrix rows
  // versus the skin list. Keeps the fastest setting of each in turn; returns the choices and every trial's ms per step.
  void autotune(int steps, int64_t* choice, double* trial_ms, int trial_cap) {
    if (!md_ready) throw ApiError(RXA_ESTATE, "rxa_md_setup first");
    if (steps < 1) throw ApiError(RXA_EINVAL, "autotune: steps must be >= 1");
    int ntrial = 0;
    auto timed = [&]() {
      md_run(2, nullptr);   // first-call effects (a rebuild, capacities) outside the timing
      double best = 1e300;
      for (int r = 0; r < 2; ++r) {
        const auto t0 = std::chrono::steady_clock::now();
        md_run(steps, nullptr);
        best = std::min(best, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps);
      }
      if (trial_ms && ntrial < trial_cap) trial_ms[ntrial] = best;
      ++ntrial;
      return best;
    };
    // each value is timed twice, the values in order and then in reverse, and scored by the sum of its two trials: a clock or
    // trajectory that drifts linearly during the search then adds the same to every value's score (the better of the two
    // trials would favour whichever value drew one lucky trial)
    auto pick = [&](auto set, std::initializer_list<int> values) {
      const std::vector<int> vals(values);
      const int nv = (int)vals.size();
      std::vector<double> t(nv, 0.0);
      for (int pass = 0; pass < 2; ++pass)
        for (int k = 0; k < nv; ++k) {
          const int i = pass ? nv - 1 - k : k;
          set(vals[i]);
          t[i] += timed();
        }
      int best = 0;
      for (int i = 1; i < nv; ++i)
        if (t[i] < t[best]) best = i;
      set(vals[best]);
    };
    pick([&](int v) { st.lanes = v; }, {4, 1, 2, 8});
    pick([&](int v) { g_mult = v; }, {1, 2});
    pick([&](int v) { overlap = v != 0; }, {1, 0});
    pick([&](int v) { nb_rows = v != 0; }, {1, 0});
    choice[0] = st.lanes; choice[1] = g_mult; choice[2] = overlap ? 1 : 0; choice[3] = nb_rows ? 1 : 0; choice[4] = ntrial;
  }

  // The MD state a run continues from, beyond positions, velocities and charges: the charge-solve history of s and t (4 deep,
  // newest first, [k n + atom] in the caller's order) and each system's thermostat chain (32 doubles per system).
  void md_state_out(double* s_hist, double* t_hist, double* nh) {
    if (!md_ready) throw ApiError(RXA_ESTATE, "rxa_md_setup first");
    sync();
    std::vector<double> h(4 * (size_t)n);
    for (int w = 0; w < 2; ++w) {
      double* out = w ? t_hist : s_hist;
      if (!out) continue;
      if (!qb.s_hist || q_dbl_n != n) throw ApiError(RXA_ESTATE, "no charge solve set up");
      d2h(h.data(), w ? qb.t_hist : qb.s_hist, h.size() * sizeof(double));
      for (int k = 0; k < 4; ++k)
        for (int i = 0; i < n; ++i) out[(size_t)k * n + perm_h[i]] = h[(size_t)k * n + i];
    }
    if (nh) d2h(nh, packed ? d_nh_pk.p : md.nh, 32 * (size_t)(packed ? S : 1) * sizeof(double));
  }

  // rxa_md_setup for a run that continues some systems (keep[k] != 0) from a state read with rxa_get_state and
  // rxa_md_state: the other systems start as rxa_md_setup starts them (their first charge solve from no history), the kept ones
  // take the given charges, charge history and thermostat chain instead, so their next step is the one they would have taken
  // (the lists are rebuilt, which changes only the order of sums). v: velocities of all atoms, the kept systems' as read.
  void md_resume(const double* v, const double* mass_by_type, double dt, double tt, double tdamp, double sk, int qeq_every, int qimax,
                 double qtol, const int* keep, const double* q, const double* s_hist, const double* t_hist, const double* nh) {
    md_setup_buffers(v, mass_by_type, dt, tt, tdamp, sk, qeq_every, qimax, qtol);
    build_neighbors();
    launch_md_ref_positions(st, md, stream);
    if (md_qeq_every > 0) qeq_charges(md_qeq_imax, md_qeq_tol, 0);
    sync();
    if (packed) pk_collect();
    const int ns = packed ? S : 1;
    std::vector<char> kept_atom(n, 0);   // engine order
    for (int k = 0; k < ns; ++k) {
      if (!keep[k]) continue;
      const int a = packed ? sys_start_h[k] : 0, b = packed ? sys_start_h[k + 1] : n;
      for (int i = a; i < b; ++i) kept_atom[i] = 1;
    }
    auto merge = [&](double* dev, const double* src, int depth) {   // kept atoms' values from src (caller order)
      std::vector<double> h((size_t)depth * n);
      d2h(h.data(), dev, h.size() * sizeof(double));
      for (int k = 0; k < depth; ++k)
        for (int i = 0; i < n; ++i)
          if (kept_atom[i]) h[(size_t)k * n + i] = src[(size_t)k * n + perm_h[i]];
      h2d(dev, h.data(), h.size() * sizeof(double));
    };
    merge(st.q, q, 1);
    if (md_qeq_every > 0) {
      if (!qb.s_hist || q_dbl_n != n) throw ApiError(RXA_ESTATE, "no charge solve set up");
      merge(qb.s_hist, s_hist, 4);
      merge(qb.t_hist, t_hist, 4);
      hist_count = 4;   // systems without history (zero arrays) extrapolate to zero, the start of a solve without history
    }
    {
      double* dnh = packed ? d_nh_pk.p : md.nh;
      std::vector<double> h(32 * (size_t)ns);
      d2h(h.data(), dnh, h.size() * sizeof(double));
      for (int k = 0; k < ns; ++k)
        if (keep[k]) std::memcpy(h.data() + 32 * (size_t)k, nh + 32 * (size_t)k, 32 * sizeof(double));
      h2d(dnh, h.data(), h.size() * sizeof(double));
    }
    qeq_h_valid = false;
    forces(false);
    sync();
  }

  void md_run(int steps, double* thermo) {
    if (!md_ready) throw ApiError(RXA_ESTATE, "rxa_md_setup first");
    const double ftm2v = 1.0 / 48.88821291 / 48.88821291;
    const double dtf = 0.5 * md_dt * ftm2v;
    long long reb0 = rebuilds;
    double disp = 0;
    for (int s = 0; s < steps; ++s) {
      hipEvent_t t0 = timer.begin(PH_MD);
      if (packed) {
        launch_pk_thermostat(st, md, d_sys_start.p, S, d_nh_pk.p, stream);
        launch_md_half_kick(st, md, dtf, stream);
        launch_md_drift(st, md, md_dt, stream);
      } else {   // thermostat, half kick, drift and the displacement partials in fused kernels
        launch_md_step_start(st, md, dtf, md_dt, !md_nve, v_scale_pending, stream);
        v_scale_pending = false;
      }
      if (packed) {   // each system's own displacement decides its own rebuild
        launch_fill_int(d_rebany.p, 1, 0, stream);
        launch_pk_disp(st, md, d_sys_start.p, S, 0.5 * skin, d_reb.p, d_rebany.p, d_disp_pk.p, stream);
        HIPCHK(hipMemcpyAsync(rebany_h, d_rebany.p, sizeof(int), hipMemcpyDeviceToHost, stream));
      } else {   // the kick-drift kernel's per-block largest squared displacements, reduced on the host (exact: a maximum)
        HIPCHK(hipMemcpyAsync(disp_h, md.partial, md_disp_partials() * sizeof(double), hipMemcpyDeviceToHost, stream));
      }
      timer.end(PH_MD, t0);
      HIPCHK(hipEventRecord(ev_disp, stream));
      qeq_h_valid = false;
      ++md_steps;
      const bool do_qeq = md_qeq_every > 0 && md_steps % md_qeq_every == 0;
      if (do_qeq && nbr_valid) qeq_matrix(true);   // speculative: the common step keeps its lists; redone after a rebuild
      HIPCHK(hipEventSynchronize(ev_disp));
      fork_after = ev_disp;   // the bonded chain need not wait for the charge matrix queued above
      if (packed) {
        if (*rebany_h) {
          build_neighbors_packed(false);   // the flagged systems only; invalidates the matrix
          launch_pk_ref(st, md, d_reb.p, stream);
          fork_after = nullptr;
        }
      } else {
        double m = 0;
        for (int k = 0; k < md_disp_partials(); ++k) m = std::max(m, disp_h[k]);
        disp = std::sqrt(m);
        if (disp > 0.5 * skin) {
          build_neighbors();   // invalidates the matrix
          launch_md_ref_positions(st, md, stream);
          fork_after = nullptr;
        }
      }
      // bonded chain on the bonded stream while the main stream solves for the charges, then the non-bonded pass and the join;
      // energy totals are read once, after the last step
      bonded_begin();
      if (do_qeq) qeq_begin(md_qeq_imax, md_qeq_tol);   // polls for the bonded chain's flags between its launches
      if (bonded_pending) bonded_end();
      if (do_qeq) qeq_end(md_qeq_imax, 0);
      nonbonded_launch();
      if (packed) {
        join_gather(s == steps - 1);
        t0 = timer.begin(PH_MD);
        launch_md_half_kick(st, md, dtf, stream);
        launch_pk_thermostat(st, md, d_sys_start.p, S, d_nh_pk.p, stream);
        timer.end(PH_MD, t0);
      } else {   // the gather fused with the half kick; NVE: thermostat pass (kinetic energy, factor 1) after the last step only;
                 // NVT: the velocity factor is applied by the next kernel
        if (bstream() != stream) HIPCHK(hipStreamWaitEvent(stream, ev_bonded, 0));
        t0 = timer.begin(PH_DBOND);
        launch_md_gather_step_end(st, md, dtf, !md_nve || s == steps - 1, stream);
        if (s == steps - 1) launch_energy_reduce(st, red.p, stream);
        timer.end(PH_DBOND, t0);
        v_scale_pending = !md_nve;
      }
    }
    if (v_scale_pending) {   // the last thermostat factor, before anything reads the velocities
      launch_md_scale(st, md, stream);
      v_scale_pending = false;
    }
    sync();
    if (packed) pk_collect();
    if (thermo) {
      double e[kTerms];
      d2h(e, red.p, sizeof(e));
      double nh[22];
      d2h(nh, md.nh, sizeof(nh));
      if (packed) {   // totals over the systems: kinetic energy, temperature from the summed degrees of freedom
        std::vector<double> nhp(32 * (size_t)S);
        d2h(nhp.data(), d_nh_pk.p, nhp.size() * sizeof(double));
        double ke = 0, dof = 0;
        for (int k = 0; k < S; ++k) { ke += nhp[32 * (size_t)k + 21]; dof += nhp[32 * (size_t)k + 16]; }
        nh[21] = ke;
        nh[20] = dof > 0 ? 2.0 * ke / (dof * nhp[17]) : 0.0;
      }
      double pe = 0;
      for (double v : e) pe += v;
      thermo[0] = pe;
      thermo[1] = nh[21];
      thermo[2] = nh[20];
      thermo[3] = (double)(rebuilds - reb0);
      thermo[4] = packed ? 0.5 * (last_iters[0] + last_iters[1]) : (qeq_calls ? qeq_iter_sum / qeq_calls : 0.0);
      thermo[5] = last_iters[0];
      thermo[6] = last_iters[1];
      thermo[7] = disp;
    }
  }
};

}  // namespace rxa

// =====================================================================================================================
// C ABI
// =====================================================================================================================
using rxa::ApiError;
using rxa::Engine;

struct rxa_context {
  std::unique_ptr<Engine> e;
  std::string err;
};

// ---- allocation-free charge-matrix product (rxa_qeq_matvec_noalloc) ----------------------------------------------------------
// Its staging vectors (the caller's vector and the product, both in engine order) live in a file-scope device array instead of
// engine allocations: kNoAllocMaxAtoms atoms at most, 2 * kNoAllocMaxAtoms doubles (4 MiB) on each device the library runs on,
// held until the process ends. Host code never uses the array's name as an address: rxa_create looks its address up once per
// device (hipGetSymbolAddress, with that device current) and the entry point passes that address to the kernels.
namespace rxa {
constexpr int kNoAllocMaxAtoms = 262144;
static __device__ __align__(256) double scratch[2 * (size_t)kNoAllocMaxAtoms];
constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

namespace {
__global__ void __launch_bounds__(256) k_perm_gather(int n, const int* __restrict__ perm, const double* __restrict__ src, double* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}
__global__ void __launch_bounds__(256) k_perm_scatter(int n, const int* __restrict__ perm, const double* __restrict__ src, double* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = src[i];
}
}  // namespace

// the device address of `scratch` on the current device (device index dev), looked up once
int scratch_init(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return RXA_EDEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[dev]) return RXA_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RXA_EDEVICE;
  g_scratch_dev[dev] = static_cast<double*>(p);
  return RXA_OK;
}
}  // namespace rxa

namespace {
template <class F>
int guard(rxa_context* c, F&& f) {
  if (!c) return RXA_EINVAL;
  try {
    f(*c->e);
    c->err.clear();
    return RXA_OK;
  } catch (const ApiError& e) {
    c->err = e.what();
    return e.code;
  } catch (const rxa::FfieldError& e) {
    c->err = e.what();
    return RXA_EFFIELD;
  } catch (const std::exception& e) {
    c->err = e.what();
    return RXA_EINVAL;
  }
}
}  // namespace

extern "C" {

int rxa_abi_version(void) { return 100; }

int rxa_create(rxa_context** out, int device) {
  if (!out) return RXA_EINVAL;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return RXA_EDEVICE;
  if (device < 0 || device >= count) return RXA_EINVAL;
  try {
    auto c = std::make_unique<rxa_context>();
    c->e = std::make_unique<Engine>(device);   // makes `device` current
    if (rxa::scratch_init(device) != RXA_OK) return RXA_EDEVICE;
    *out = c.release();
    return RXA_OK;
  } catch (...) {
    return RXA_EDEVICE;
  }
}

void rxa_destroy(rxa_context* c) { delete c; }

const char* rxa_last_error(const rxa_context* c) { return c ? c->err.c_str() : "null context"; }

int rxa_pair_style(rxa_context* c, int nargs, const char* const* args, const char* control_text) {
  return guard(c, [&](Engine& e) {
    std::vector<std::string> a;
    for (int i = 0; i < nargs; ++i) a.emplace_back(args[i]);
    e.ps = rxa::parse_pair_style(a, control_text ? control_text : "");
    e.have_style = true;
    if (e.have_ff) e.make_globals();
  });
}

int rxa_pair_coeff(rxa_context* c, const char* text, int ntypes, const char* const* elements, int* map_out) {
  return guard(c, [&](Engine& e) {
    if (!e.have_style) throw ApiError(RXA_ESTATE, "pair_style must precede pair_coeff");
    if (!text) throw ApiError(RXA_EINVAL, "no force field text");
    e.ff = rxa::read_ffield(text, e.ps.lgvdw);
    if (e.ps.lgvdw && e.ff.vdw_type == 1)
      throw ApiError(RXA_EFFIELD, "lgvdw with a shielding-only force field: the lg parameters would be read but never used");
    e.lmp2ff.assign(ntypes + 1, -1);
    for (int t = 0; t < ntypes; ++t) {
      const std::string el = elements[t];
      if (el == "NULL") continue;
      const int m = e.ff.match_element(el);
      if (m == -1) throw ApiError(RXA_EFFIELD, "element " + el + " not found in the force field");
      if (m == -2) throw ApiError(RXA_EFFIELD, "element " + el + " is not unique in the force field");
      e.lmp2ff[t + 1] = m;
    }
    if (map_out) for (int t = 0; t < ntypes; ++t) map_out[t] = e.lmp2ff[t + 1];
    e.pack_tables();
    e.have_ff = true;
    e.make_globals();
  });
}

int rxa_extract(rxa_context* c, const char* name, double* out) {
  return guard(c, [&](Engine& e) {
    if (!e.have_ff) throw ApiError(RXA_ESTATE, "no force field");
    const std::string nm = name ? name : "";
    out[0] = 0;
    for (size_t t = 1; t < e.lmp2ff.size(); ++t) {
      const int f = e.lmp2ff[t];
      double v = 0;
      if (f >= 0) {
        if (nm == "chi") v = e.ff.sp[f].chi;
        else if (nm == "eta") v = e.ff.sp[f].eta;
        else if (nm == "gamma") v = e.ff.sp[f].gamma;
        else if (nm == "bcut_acks2") v = e.ff.sp[f].bcut_acks2;
        else throw ApiError(RXA_EINVAL, "extract: unknown name " + nm);
      }
      out[t] = v;
    }
  });
}

int rxa_cutoffs(rxa_context* c, double* out) {
  return guard(c, [&](Engine& e) {
    if (!e.have_ff) throw ApiError(RXA_ESTATE, "no force field");
    out[0] = e.ff.nonb_cut;
    out[1] = e.ps.control.bond_cut;
    out[2] = e.ps.control.hbond_cut;
    out[3] = std::max({out[0], out[2], 2.0 * out[1]});
  });
}

long rxa_table_dump(rxa_context* c, int portable, char* buf, long cap) {
  if (!c || !c->e->have_ff) return -1;
  const std::string d = rxa::table_dump(c->e->ff, portable != 0);
  if (buf && cap > 0) {
    const long m = std::min<long>(cap - 1, (long)d.size());
    std::memcpy(buf, d.data(), (size_t)m);
    buf[m] = 0;
  }
  return (long)d.size();
}

int rxa_set_box(rxa_context* c, const double lo[3], const double hi[3], const double tilt[3], const int periodic[3]) {
  return guard(c, [&](Engine& e) { e.set_box(lo, hi, tilt, periodic); });
}

int rxa_set_atoms(rxa_context* c, int64_t n, const int* types, const int64_t* tags, const double* x) {
  return guard(c, [&](Engine& e) {
    e.set_atoms(n, types, tags, x);
    e.lmp_types.assign(types, types + n);
  });
}

int rxa_set_atoms_ghosted(rxa_context* c, int64_t nlocal, int64_t nghost, const int* types, const int64_t* tags, const double* x,
                          int nranks) {
  return guard(c, [&](Engine& e) {
    if (!e.have_box) throw ApiError(RXA_ESTATE, "the box must be set before the atoms");
    if (!types || !tags || !x) throw ApiError(RXA_EINVAL, "null atom arrays");
    rxa::check_ghosts(e.cell, nlocal, nghost, types, tags, x, nranks);
    e.set_atoms(nlocal, types, tags, x);   // the owned atoms; the engine generates every image itself
  });
}

int rxa_check_ghosts(const double lo[3], const double hi[3], const double tilt[3], const int periodic[3], int64_t nlocal, int64_t nghost,
                     const int* types, const int64_t* tags, const double* x, int nranks, char* err, long errcap) {
  try {
    const rxa::Cell cell = Engine::make_cell(lo, hi, tilt, periodic);
    rxa::check_ghosts(cell, nlocal, nghost, types, tags, x, nranks);
    if (err && errcap > 0) err[0] = 0;
    return RXA_OK;
  } catch (const ApiError& ex) {
    if (err && errcap > 0) { std::strncpy(err, ex.what(), (size_t)errcap - 1); err[errcap - 1] = 0; }
    return ex.code;
  }
}

int rxa_set_systems(rxa_context* c, int nsys, const int64_t* counts, const double* boxes, const int* periodic, const int* types,
                    const int64_t* tags, const double* x) {
  return guard(c, [&](Engine& e) {
    if (nsys < 1 || !counts || !boxes || !periodic || !types || !tags || !x) throw ApiError(RXA_EINVAL, "rxa_set_systems: null argument");
    long long nn = 0;
    for (int k = 0; k < nsys; ++k) nn += counts[k];
    e.set_systems(nsys, counts, boxes, periodic, types, tags, x);
    e.lmp_types.assign(types, types + nn);
  });
}

int rxa_launch_config(rxa_context* c, const int* set, int* get) {
  return guard(c, [&](Engine& e) {
    if (set) {
      if (set[0] != 1 && set[0] != 2 && set[0] != 4 && set[0] != 8) throw ApiError(RXA_EINVAL, "lanes per atom must be 1, 2, 4 or 8");
      if (set[1] != 1 && set[1] != 2) throw ApiError(RXA_EINVAL, "the group multiplier must be 1 or 2");
      e.sync();
      e.st.lanes = set[0]; e.g_mult = set[1]; e.overlap = set[2] != 0; e.nb_rows = set[3] != 0;
    }
    if (get) { get[0] = e.st.lanes; get[1] = e.g_mult; get[2] = e.overlap ? 1 : 0; get[3] = e.nb_rows ? 1 : 0; }
  });
}

int rxa_autotune(rxa_context* c, int steps, int64_t* choice, double* trial_ms, int trial_cap) {
  return guard(c, [&](Engine& e) {
    if (!choice) throw ApiError(RXA_EINVAL, "rxa_autotune: null argument");
    e.autotune(steps, choice, trial_ms, trial_cap);
  });
}

int rxa_md_state(rxa_context* c, double* s_hist, double* t_hist, double* nh) {
  return guard(c, [&](Engine& e) { e.md_state_out(s_hist, t_hist, nh); });
}

int rxa_md_resume(rxa_context* c, const double* v, const double* mass_by_type, double dt, double t_target, double t_damp, double skin,
                  int qeq_every, int qeq_imax, double qeq_tol, const int* keep, const double* q, const double* s_hist,
                  const double* t_hist, const double* nh) {
  return guard(c, [&](Engine& e) {
    if (!v || !mass_by_type || !keep || !q || !nh) throw ApiError(RXA_EINVAL, "rxa_md_resume: null argument");
    if (qeq_every > 0 && (!s_hist || !t_hist)) throw ApiError(RXA_EINVAL, "rxa_md_resume: the charge history is required");
    if (!(dt > 0) || !(skin >= 0)) throw ApiError(RXA_EINVAL, "invalid MD parameters");
    e.md_resume(v, mass_by_type, dt, t_target, t_damp, skin, qeq_every, qeq_imax, qeq_tol, keep, q, s_hist, t_hist, nh);
  });
}

int rxa_system_results(rxa_context* c, double* energies, double* virial, double* thermo) {
  return guard(c, [&](Engine& e) { e.system_results(energies, virial, thermo); });
}

int rxa_md_system_targets(rxa_context* c, const double* t_target, const double* t_damp) {
  return guard(c, [&](Engine& e) {
    if (!e.packed) throw ApiError(RXA_ESTATE, "no packed systems (rxa_set_systems)");
    if (!e.md_ready) throw ApiError(RXA_ESTATE, "rxa_md_setup first");
    std::vector<double> nh(32 * (size_t)e.S);
    e.d2h(nh.data(), e.d_nh_pk.p, nh.size() * sizeof(double));
    for (int k = 0; k < e.S; ++k) {
      double* h = nh.data() + 32 * (size_t)k;
      if (t_target) h[13] = t_target[k];
      if (t_damp) { h[14] = t_damp[k] > 0 ? 1.0 / t_damp[k] : 0.0; h[22] = t_damp[k] > 0 ? 0.0 : 1.0; }
    }
    e.h2d(e.d_nh_pk.p, nh.data(), nh.size() * sizeof(double));
  });
}

int rxa_set_positions(rxa_context* c, const double* x) {
  return guard(c, [&](Engine& e) {
    if (!e.n) throw ApiError(RXA_ESTATE, "no atoms");
    e.set_positions(x);
  });
}

int rxa_set_positions_device(rxa_context* c, const double* dx) {
  return guard(c, [&](Engine& e) {
    if (!e.n) throw ApiError(RXA_ESTATE, "no atoms");
    rxa::launch_positions_in(e.st, dx, e.d_perm.p, e.stream);
    e.nbr_valid = false;
    e.qeq_h_valid = false;
  });
}

int rxa_stream(rxa_context* c, void** out) {
  return guard(c, [&](Engine& e) { *out = (void*)e.stream; });
}

int rxa_stream_copy(rxa_context* c, int64_t bytes, int reps) {
  return guard(c, [&](Engine& e) {
    if (bytes <= 0 || bytes % 16 || reps < 1) throw ApiError(RXA_EINVAL, "stream copy: bytes must be a positive multiple of 16");
    rxa::DBuf<char> a, b;
    a.alloc((size_t)bytes);
    b.alloc((size_t)bytes);
    HIPCHK(hipMemsetAsync(a.p, 0, (size_t)bytes, e.stream));
    for (int r = 0; r < reps; ++r) rxa::launch_stream_copy(a.p, b.p, (size_t)bytes, e.stream);
    e.sync();
  });
}

int rxa_set_charges(rxa_context* c, const double* q) {
  return guard(c, [&](Engine& e) {
    if (!e.n) throw ApiError(RXA_ESTATE, "no atoms");
    std::vector<double> qi(e.n);
    for (int i = 0; i < e.n; ++i) qi[i] = q[e.perm_h[i]];
    e.sync();
    e.h2d(e.st.q, qi.data(), e.n * sizeof(double));
  });
}

int rxa_get_charges(rxa_context* c, double* q) {
  return guard(c, [&](Engine& e) { e.charges_out(q); });
}

int rxa_compute(rxa_context* c, int flags, double* energies, double* grad, double* virial) {
  return guard(c, [&](Engine& e) { e.compute(flags, energies, grad, virial); });
}

int rxa_peratom(rxa_context* c, double* eatom, double* vatom) {
  return guard(c, [&](Engine& e) { e.peratom_out(eatom, vatom); });
}

void rxa_pvector(const double* e, double* pv) {
  // eb ea elp emol ev epen ecoa ehb et eco ew ep efi eqeq
  pv[0] = e[0]; pv[1] = e[1] + e[2]; pv[2] = e[3]; pv[3] = 0; pv[4] = e[4]; pv[5] = e[5]; pv[6] = e[6];
  pv[7] = e[7]; pv[8] = e[8]; pv[9] = e[9]; pv[10] = e[10]; pv[11] = e[11]; pv[12] = 0; pv[13] = e[12];
}

int rxa_census(rxa_context* c, int* out) {
  return guard(c, [&](Engine& e) {
    if (!e.n) throw ApiError(RXA_ESTATE, "no atoms");
    // the census of the last rxa_compute made with RXA_COMPUTE_CENSUS, else a force call at the current positions and charges
    // that records it (the same energies and forces)
    if (!e.census_valid) e.compute(RXA_COMPUTE_CENSUS, nullptr, nullptr, nullptr);
    const int n = e.n;
    std::vector<int> b(n), cen((size_t)rxa::kCenCount * n);
    std::vector<double> de(n);
    e.sync();
    e.d2h(b.data(), e.st.b_cnt, n * sizeof(int));
    e.d2h(de.data(), e.st.Delta_e, n * sizeof(double));
    e.d2h(cen.data(), e.st.cen, cen.size() * sizeof(int));
    // reference order (census.hpp): bonds, angle_sets, torsions, hbonds, nonbonded, sbo_region, lp_trunc
    const int from[7] = {-1, rxa::kCenAngles, rxa::kCenTorsions, rxa::kCenHbonds, rxa::kCenNonbonded, rxa::kCenSbo, -1};
    for (int i = 0; i < n; ++i) {
      const int u = e.perm_h[i];
      out[u] = std::min(b[i], e.maxb);
      for (int k = 1; k < 6; ++k) out[(size_t)k * n + u] = cen[(size_t)from[k] * n + i];
      out[6 * (size_t)n + u] = (int)(de[i] / 2.0);
    }
  });
}

long rxa_bond_table(rxa_context* c, int32_t* start, int32_t* nbr, double* bo, double* total_bo, double* nlp) {
  if (!c) return -1;
  long total = -1;
  const int rc = guard(c, [&](Engine& e) {
    const int n = e.n, mb = e.maxb;
    e.sync();
    std::vector<int> cnt(n), bj((size_t)n * mb);
    std::vector<double> bv((size_t)n * mb), tb(n), np(n);
    e.d2h(cnt.data(), e.st.b_cnt, n * sizeof(int));
    total = 0;
    for (int i = 0; i < n; ++i) total += std::min(cnt[i], mb);
    if (!nbr) return;
    e.d2h(bj.data(), e.st.b_j, bj.size() * sizeof(int));
    e.d2h(bv.data(), e.st.bo, bv.size() * sizeof(double));
    e.d2h(tb.data(), e.st.tbo, n * sizeof(double));
    e.d2h(np.data(), e.st.nlp, n * sizeof(double));
    std::vector<int> inv(n);
    for (int i = 0; i < n; ++i) inv[e.perm_h[i]] = i;
    long k = 0;
    for (int u = 0; u < n; ++u) {   // rows and partners in the caller's order
      const int i = inv[u];
      if (start) start[u] = (int32_t)k;
      for (int p = 0; p < std::min(cnt[i], mb); ++p, ++k) {
        nbr[k] = e.perm_h[bj[(size_t)p * n + i]];
        if (bo) bo[k] = bv[(size_t)p * n + i];
      }
      if (total_bo) total_bo[u] = tb[i];
      if (nlp) nlp[u] = np[i];
    }
    if (start) start[n] = (int32_t)k;
  });
  return rc == RXA_OK ? total : rc;
}

int rxa_qeq_setup(rxa_context* c, double swa, double swb, const double* chi, const double* eta, const double* gamma) {
  return guard(c, [&](Engine& e) { e.qeq_setup(swa, swb, chi, eta, gamma); });
}

int rxa_qeq_matvec(rxa_context* c, const double* x, double* y) {
  return guard(c, [&](Engine& e) {
    e.qeq_matrix();
    e.qeq_started = false;   // the product stages its vectors in the solve's r and p
    const int n = e.n;
    std::vector<double> xi(n), yo(n);
    for (int i = 0; i < n; ++i) xi[i] = x[e.perm_h[i]];
    HIPCHK(hipMemcpyAsync(e.qb.p, xi.data(), n * sizeof(double), hipMemcpyHostToDevice, e.stream));
    rxa::launch_qeq_matvec1(e.st, e.qb, e.qb.p, e.qb.r, e.stream);
    HIPCHK(hipMemcpyAsync(yo.data(), e.qb.r, n * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    e.sync();
    for (int i = 0; i < n; ++i) y[e.perm_h[i]] = yo[i];
  });
}

int rxa_qeq_matvec_noalloc(rxa_context* c, const double* x_dev, double* y_dev) {
  return guard(c, [&](Engine& e) {
    const int n = e.n;
    if (!n) throw ApiError(RXA_ESTATE, "no atoms");
    if (n > rxa::kNoAllocMaxAtoms) throw ApiError(RXA_EINVAL, "rxa_qeq_matvec_noalloc: more than 262144 atoms (the size of its scratch)");
    if (!x_dev || !y_dev) throw ApiError(RXA_EINVAL, "rxa_qeq_matvec_noalloc: null vector");
    if (e.packed || !e.qeq_ready || !e.qeq_h_valid)
      throw ApiError(RXA_ESTATE, "rxa_qeq_matvec_noalloc: the charge matrix of the current positions is not built (call rxa_qeq_matvec, "
                                 "rxa_qeq_solve, rxa_qeq_charges or an MD step first)");
    double* const sx = (e.device >= 0 && e.device < rxa::kMaxDevices) ? rxa::g_scratch_dev[e.device] : nullptr;
    if (!sx) throw ApiError(RXA_EDEVICE, "rxa_qeq_matvec_noalloc: scratch address of this device unknown");
    double* const sy = sx + rxa::kNoAllocMaxAtoms;
    const dim3 blk(256), grid((n + 255) / 256);
    hipLaunchKernelGGL(rxa::k_perm_gather, grid, blk, 0, e.stream, n, (const int*)e.d_perm.p, x_dev, sx);
    rxa::launch_qeq_matvec1(e.st, e.qb, sx, sy, e.stream);
    hipLaunchKernelGGL(rxa::k_perm_scatter, grid, blk, 0, e.stream, n, (const int*)e.d_perm.p, (const double*)sy, y_dev);
    HIPCHK(hipGetLastError());
  });
}

int rxa_qeq_solve(rxa_context* c, const double* b, const double* x0, int imax, double tol, double* x, int* iters, double* rel, int* status) {
  return guard(c, [&](Engine& e) {
    if (imax < 1) throw ApiError(RXA_EINVAL, "maxiter must be >= 1");
    e.ensure_cg_cap(imax);
    e.qeq_matrix();
    e.qeq_started = false;   // the caller's b and x0 replace the stock start
    const int n = e.n;
    std::vector<double> bi(2 * (size_t)n), xi(2 * (size_t)n);
    for (int i = 0; i < n; ++i) {
      const int u = e.perm_h[i];
      bi[2 * i] = b[u]; bi[2 * i + 1] = b[n + u]; xi[2 * i] = x0[u]; xi[2 * i + 1] = x0[n + u];
    }
    HIPCHK(hipMemcpyAsync(e.qb.b, bi.data(), bi.size() * sizeof(double), hipMemcpyHostToDevice, e.stream));
    HIPCHK(hipMemcpyAsync(e.qb.x, xi.data(), xi.size() * sizeof(double), hipMemcpyHostToDevice, e.stream));
    rxa::launch_qeq_init(e.st, e.qb, imax, tol, 2, e.stream);
    e.qeq_iterate(imax, tol);
    e.sync();
    std::vector<double> xo(2 * (size_t)n);
    e.d2h(xo.data(), e.qb.x, xo.size() * sizeof(double));
    for (int i = 0; i < n; ++i) { x[e.perm_h[i]] = xo[2 * i]; x[n + e.perm_h[i]] = xo[2 * i + 1]; }
    for (int s = 0; s < 2; ++s) {
      iters[s] = e.last_iters[s];
      rel[s] = e.last_rel[s];
      status[s] = e.last_rel[s] <= tol ? 1 : 3;
    }
  });
}

int rxa_qeq_charges(rxa_context* c, int imax, double tol, int strict, double* q_out, int* iters, double* rel) {
  return guard(c, [&](Engine& e) {
    e.qeq_charges(imax, tol, strict);
    if (e.packed) { e.sync(); e.pk_collect(); }
    if (q_out) e.charges_out(q_out);
    if (iters) { iters[0] = e.last_iters[0]; iters[1] = e.last_iters[1]; }
    if (rel) { rel[0] = e.last_rel[0]; rel[1] = e.last_rel[1]; }
  });
}

int rxa_qeq_refine(rxa_context* c, int mode) {
  return guard(c, [&](Engine& e) {
    if (mode < -1 || mode > 1) throw ApiError(RXA_EINVAL, "qeq refine mode must be -1, 0 or 1");
    e.sync();
    e.qeq_refine_mode = mode;
    e.apply_refine();
  });
}

int rxa_qeq_reset_history(rxa_context* c) {
  return guard(c, [&](Engine& e) { e.reset_history(); });
}

int rxa_md_setup(rxa_context* c, const double* v, const double* mass_by_type, double dt, double t_target, double t_damp, double skin,
                 int qeq_every, int qeq_imax, double qeq_tol) {
  return guard(c, [&](Engine& e) { e.md_setup(v, mass_by_type, dt, t_target, t_damp, skin, qeq_every, qeq_imax, qeq_tol); });
}

int rxa_md_run(rxa_context* c, int steps, double* thermo) {
  return guard(c, [&](Engine& e) { e.md_run(steps, thermo); });
}

int rxa_get_state(rxa_context* c, double* x, double* v, double* q) {
  return guard(c, [&](Engine& e) {
    e.sync();
    const int n = e.n;
    std::vector<double> a(n), b(n), d(n);
    const std::vector<int>& P = e.perm_h;
    if (x) {
      e.d2h(a.data(), e.st.x, n * sizeof(double));
      e.d2h(b.data(), e.st.y, n * sizeof(double));
      e.d2h(d.data(), e.st.z, n * sizeof(double));
      for (int i = 0; i < n; ++i) { x[3 * P[i]] = a[i]; x[3 * P[i] + 1] = b[i]; x[3 * P[i] + 2] = d[i]; }
    }
    if (v) {
      if (!e.md_ready) throw ApiError(RXA_ESTATE, "no velocities (rxa_md_setup)");
      e.d2h(a.data(), e.md.vx, n * sizeof(double));
      e.d2h(b.data(), e.md.vy, n * sizeof(double));
      e.d2h(d.data(), e.md.vz, n * sizeof(double));
      for (int i = 0; i < n; ++i) { v[3 * P[i]] = a[i]; v[3 * P[i] + 1] = b[i]; v[3 * P[i] + 2] = d[i]; }
    }
    if (q) e.charges_out(q);
  });
}

static void copy_out(const std::string& s, char* buf, long cap) {
  if (!buf || cap <= 0) return;
  const long m = std::min<long>(cap - 1, (long)s.size());
  std::memcpy(buf, s.data(), (size_t)m);
  buf[m] = 0;
}

int rxa_ffield_table_dump(const char* text, int lgvdw, int portable, char* buf, long cap, long* len, char* err, long errcap) {
  try {
    if (!text) throw std::invalid_argument("no force field text");
    const rxa::ForceField ff = rxa::read_ffield(text, lgvdw != 0);
    const std::string d = rxa::table_dump(ff, portable != 0);
    if (len) *len = (long)d.size();
    copy_out(d, buf, cap);
    return RXA_OK;
  } catch (const rxa::FfieldError& e) {
    copy_out(e.what(), err, errcap);
    return RXA_EFFIELD;
  } catch (const std::exception& e) {
    copy_out(e.what(), err, errcap);
    return RXA_EINVAL;
  }
}

int rxa_check_pair_style(int nargs, const char* const* args, const char* control_text, char* summary, long cap) {
  try {
    std::vector<std::string> a;
    for (int i = 0; i < nargs; ++i) a.emplace_back(args[i]);
    const rxa::PairSettings p = rxa::parse_pair_style(a, control_text ? control_text : "");
    std::string o;
    char b[128];
    o += "precision=" + p.effective_precision() + "\n";
    o += "backend=" + p.backend + "\nbonded=" + p.bonded + "\n";
    o += std::string("checkqeq=") + (p.checkqeq ? "yes" : "no") + "\nenobonds=" + (p.enobonds ? "yes" : "no") + "\nlgvdw=" + (p.lgvdw ? "yes" : "no") + "\n";
    std::snprintf(b, sizeof b, "bond_cut=%.17g\nhbond_cut=%.17g\nthb_cut=%.17g\nthb_cutsq=%.17g\n", p.control.bond_cut, p.control.hbond_cut,
                  p.control.thb_cut, p.control.thb_cutsq);
    o += b;
    for (const auto& nt : p.notices) o += "notice=" + nt + "\n";
    copy_out(o, summary, cap);
    return RXA_OK;
  } catch (const std::exception& e) {
    copy_out(e.what(), summary, cap);
    return RXA_EFFIELD;
  }
}

int rxa_profile(rxa_context* c, int enable) {
  return guard(c, [&](Engine& e) {
    e.sync();
    e.timer.on = enable != 0;
    e.timer.mask = enable == 1 ? ((1u << rxa::PH_COUNT) - 1u) : (unsigned)enable;
    e.timer.reset();
  });
}

int rxa_profile_read(rxa_context* c, double* ms, int64_t* counts) {
  return guard(c, [&](Engine& e) {
    e.sync();
    for (int k = 0; k < rxa::PH_COUNT; ++k) {   // sampled phases: the timed mean times every launch sequence
      const long long all = e.timer.seen[k], timed = e.timer.cnt[k];
      ms[k] = timed ? e.timer.ms[k] * (double)all / (double)timed : 0.0;
      counts[k] = timed ? all : 0;
    }
  });
}

int rxa_profile_every(rxa_context* c, int every) {
  return guard(c, [&](Engine& e) {
    if (every < 1) throw ApiError(RXA_EINVAL, "profile sampling period must be >= 1");
    e.sync();
    e.timer.every = every;
    e.timer.reset();
  });
}

int rxa_qeq_stats(rxa_context* c, int64_t* out) {
  return guard(c, [&](Engine& e) {
    e.sync();
    out[0] = e.qeq_calls;
    out[1] = e.cgs_h ? (int64_t)e.cgs_h[4] : 0;
    out[2] = e.cgs_h ? (int64_t)e.cgs_h[5] : 0;
    out[3] = e.cg_batches;
  });
}

int rxa_stats(rxa_context* c, int64_t* out) {
  return guard(c, [&](Engine& e) {
    for (int k = 0; k < 12; ++k) out[k] = 0;
    out[0] = e.n; out[1] = e.maxnb; out[2] = e.maxb; out[3] = e.maxhc; out[8] = e.rebuilds; out[9] = (int)e.prec;
    if (!e.n) return;
    e.sync();
    const int n = e.n;
    std::vector<int> cnt(5 * (size_t)n);
    e.d2h(cnt.data(), e.st.nb_cnt, 5 * (size_t)n * sizeof(int));
    for (int i = 0; i < n; ++i) {
      out[4] += cnt[i]; out[5] += cnt[n + i]; out[6] += cnt[2 * (size_t)n + i];
      out[10] += cnt[3 * (size_t)n + i]; out[11] += cnt[4 * (size_t)n + i];
    }
    if (e.qeq_ready && e.qb.cnt && e.qeq_h_valid) {
      std::vector<int> qc(n);
      e.d2h(qc.data(), e.qb.cnt, n * sizeof(int));
      for (int i = 0; i < n; ++i) out[7] += qc[i];
    }
  });
}

}  // extern "C"
