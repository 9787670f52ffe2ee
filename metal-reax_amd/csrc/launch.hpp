// Host-side launchers of the device kernels (one translation unit per kernel family; engine.hip orchestrates them).
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace rxa {

enum class Precision { FP64 = 0, MIXED = 1, FP32 = 2 };

struct BinGrid {
  int nc[3] = {1, 1, 1};
  int ncells() const { return nc[0] * nc[1] * nc[2]; }
};
struct BinBuffers {
  unsigned *keys = nullptr, *skeys = nullptr, *atom_key = nullptr;
  int *idx = nullptr, *sidx = nullptr, *cstart = nullptr, *cend = nullptr;
  void* temp = nullptr;
  size_t temp_bytes = 0;
};

// ---- neighbour search (k_neighbor.hip)
void launch_wrap(const State& st, const Cell& c, hipStream_t s);
void launch_fill_int(int* p, int n, int v, hipStream_t s);
size_t bin_temp_bytes(int n);
void launch_bin(const State& st, const Cell& c, const BinGrid& g, BinBuffers& b, hipStream_t s);
void launch_nbr_cells(const State& st, const Cell& c, const BinGrid& g, const BinBuffers& b, double rs, hipStream_t s);
void launch_nbr_brute(const State& st, const Cell& c, const int sm[3], double rs, hipStream_t s);
void launch_pos4(const State& st, hipStream_t s);                                   // pos = (x, y, z, type)
struct PermuteSet {   // per-atom arrays moved together by launch_permute
  int n = 0, nd = 0, ni = 0, nl = 0;
  double* d[24];
  int* i[4];
  long long* l[2];
};
void launch_permute(const PermuteSet& ps, const int* src, double* sd, int* si, long long* sl, hipStream_t s);
void launch_bin_identity(int n, BinBuffers& b, hipStream_t s);
void launch_positions_in(const State& st, const double* ux, const int* perm, hipStream_t s);
void launch_shift_table(const Cell& c, double4* shv, const int smx[3], hipStream_t s);
void launch_sublists(const State& st, const Globals& g, const Tables& t, double rb, double rh, hipStream_t s);

// ---- bond-order chain and bonded terms (k_bonded.hip)
void launch_bond_build(Precision p, const State& st, const Cell& c, const Globals& g, const Tables& t, hipStream_t s);
void launch_bo_correct(Precision p, const State& st, const Globals& g, const Tables& t, hipStream_t s);
// bond energy / lone pair / over-under (no group size needed), then valence, torsion, H-bonds with the group size for the
// largest bond count bmax
void launch_atom_terms(Precision p, const State& st, const Globals& g, const Tables& t, hipStream_t s);
void launch_angle_terms(Precision p, const State& st, const Globals& g, const Tables& t, int bmax, hipStream_t s);
// BO-derivative pass (CdDelta, then the per-atom bond sums); the final gather adds the non-bonded gradient and the virial
void launch_bond_derivatives(const State& st, hipStream_t s);
void launch_gather(const State& st, hipStream_t s);
void launch_eatom_ref(const State& st, hipStream_t s);   // per-atom energies of the reference's attribution (peratom calls)
void launch_vatom_ref(const State& st, hipStream_t s);   // per-atom virials of the reference's split (peratom calls, before the gather)

// ---- non-bonded terms and the charge-equilibration matrix (k_nonbonded.hip)
void launch_nonbonded(Precision p, const State& st, const Cell& c, const Globals& g, const Tables& t, hipStream_t s);

struct QeqBuffers {
  int* cnt = nullptr;        // n: entries per row
  // the rows are stored in exactly one column format and one value format:
  int* col = nullptr;               // n * maxnb: column (owned atom index), when narrow == 0
  unsigned short* col16 = nullptr;  // n * maxnb: the same in 16 bits, when narrow == 1 (n <= 65535)
  double* val = nullptr;            // n * maxnb: H_ij in FP64 (FP64 mode)
  float* val32 = nullptr;           // n * maxnb: H_ij in FP32 (mixed and FP32 modes)
  // n * maxnb: the FP32 remainder H_ij - (double)val32 (written when refine == 1), so that (double)val32 + (double)lo32
  // carries H_ij to about 48 bits for the FP64 residuals of the iterative refinement
  float* lo32 = nullptr;
  int narrow = 0;
  // FP64 iterative refinement of the FP32-valued solve: initial and verification residuals with the two-float matrix, CG
  // restarted from the true residual while it exceeds the tolerance (mixed mode; FP32 mode with the fix's 'resident' keyword)
  int refine = 0;
  // n * maxnb: how the pair of each entry is reached from the row's atom, for the non-bonded pass over the same rows: bit 15 set
  // if the row's atom is the pair's canonical end, bits 0-14 the index in the shift-vector table of the lattice shift applied to
  // the canonical end's partner. Written when keyed == 1 (shift table of at most 32768 vectors).
  unsigned short* key = nullptr;
  int keyed = 0;
  double* eta = nullptr;     // n: diagonal (fix eta, 2 x ffield)
  double* hinv = nullptr;    // n: 1 / eta
  double* chi = nullptr;     // n
  // CG state, two systems (s: b = -chi, t: b = -1) interleaved per atom: [2*i + sys]
  double *b, *x, *r, *d, *p, *q;
  double* partial;           // 4 * kCgBlocks: per-block partials of the element-wise CG kernels
  double* partial_mv;        // 4 * qeq_mv_partials(n): per-block partials of the matrix-vector product
  double* partial_rs;        // 4 * qeq_mv_partials(n): per-block partials of the refinement's verification residual
  double* partial_qb;        // 4 * qeq_build_blocks(n): per-block partials of the solve's start fused into the matrix build
  // refinement's bound (k_qeq.hip k_check_gate): the FP32 remainders' row sums sum_j |lo_ij| (n), the CG start x0 (2 n, same
  // layout as x) and the gate's per-block partials (3 * kCgBlocks)
  double* lorow = nullptr;
  double* x0 = nullptr;
  double* partial_gate = nullptr;
  int fp32 = 0;              // the matrix values are val32 (mixed and FP32 modes) instead of val
  double* scal;              // device scalars and activity flags per iteration (see k_qeq.hip), own allocation sized by imax
  double *s_hist, *t_hist;   // 4 * n each, [k * n + i], k = 0 most recent
};
constexpr int kCgBlocks = 256;
// doubles per iteration slot of QeqBuffers::scal after its 8-double header (sig of s and t, the activity pair, the restart-record
// pair); the scalars of iterations 0 .. imax + 1 take 8 + kScalSlot (imax + 2) doubles
constexpr int kScalSlot = 4;
inline size_t qeq_scal_doubles(int imax) { return 8 + (size_t)kScalSlot * (size_t)(imax + 2); }
// several independent systems packed in one engine: per-system CG scalars and row scratch (k_qeq.hip)
struct PackedQeq {
  int S = 0;
  const int* sys_start = nullptr;   // S + 1 row offsets
  const int* sysid = nullptr;       // n: system of each row
  double* pk = nullptr;             // 10 S: per solve |b|, r.p, alpha, beta; sums of s and t
  int* pki = nullptr;               // 4 S: per solve state (1 active, 0 stopped, 2 stopped and verified by the refinement
                                    // check), stock loop index
  double* rowv = nullptr;           // 4 n: per-row d.q and r.p of both solves
  int* any = nullptr;               // imax + 2: some system active for iteration it + 1 (set by iteration it; [0] by the start)
};
void launch_pk_qeq_init(const State& st, QeqBuffers& qb, const PackedQeq& pq, int imax, double tol, int use_hist, hipStream_t s);
void launch_pk_qeq_iteration(const State& st, QeqBuffers& qb, const PackedQeq& pq, int it, int imax, double tol, hipStream_t s);
// refinement check after iteration it (qb.refine): every solve that has stopped and is not yet verified gets its true residual
// with the two-float matrix; a solve above tol restarts (beta 0) for iteration it + 1, the others are verified
void launch_pk_qeq_restart(const State& st, QeqBuffers& qb, const PackedQeq& pq, int it, int imax, double tol, hipStream_t s);
void launch_pk_qeq_finish(const State& st, QeqBuffers& qb, const PackedQeq& pq, hipStream_t s);
// packed systems: neighbour search, integrator and energy sums per system (k_packed.hip)
void launch_pk_shift_table(const Cell* cells, int S, double4* shv, const int smx[3], hipStream_t s);
void launch_pk_wrap(const State& st, const Cell* cells, const int* reb, hipStream_t s);
void launch_pk_nbr(const State& st, const Cell* cells, const int* sys_start, const int4* ext, const int* reb, double rs, hipStream_t s);
// fuse = 1: also the charge solve's start from the guess in x (launch_qeq_guess before it), finished by launch_qeq_init_fused
void launch_qeq_build(const State& st, const Cell& c, const Globals& g, const Tables& t, QeqBuffers& qb, int fuse, hipStream_t s);
inline int qeq_build_blocks(int n) { return (int)(((long long)n * 64 + 255) / 256); }
void launch_qeq_guess(const State& st, QeqBuffers& qb, int use_hist, hipStream_t s);
void launch_qeq_init_fused(const State& st, QeqBuffers& qb, int imax, double tol, hipStream_t s);
// census of the non-bonded cut-off decision (st.cen, category kCenNonbonded)
void launch_census_nb(const State& st, const Globals& g, hipStream_t s);
// Non-bonded terms over the charge-matrix rows instead of the skin list: valid while the rows belong to the current positions,
// carry keys, and the matrix cutoff covers nonb_cut.
void launch_nonbonded_rows(Precision p, const State& st, const Globals& g, const Tables& t, const QeqBuffers& qb, hipStream_t s);
void launch_qeq_matvec1(const State& st, const QeqBuffers& qb, const double* x, double* y, hipStream_t s);
void launch_qeq_init(const State& st, QeqBuffers& qb, int imax, double tol, int use_hist, hipStream_t s);
void launch_qeq_mvdir(const State& st, QeqBuffers& qb, int it, int imax, double tol, hipStream_t s);   // beta, d, q = A d, d.q partials
void launch_qeq_update(const State& st, QeqBuffers& qb, int it, hipStream_t s);                        // alpha, x, r, p, r.p partials
// refinement check after update(it): if both systems have stopped, the true residual b - (diag(eta) + H) x with the two-float
// matrix restarts every system whose residual still exceeds tol (slot `it` of the scalars); otherwise it leaves a marker and
// the next product continues the iteration (launch_qeq_mvdir with chk = 1)
void launch_qeq_restart(const State& st, QeqBuffers& qb, int it, int imax, double tol, hipStream_t s);
void launch_qeq_mvdir_chk(const State& st, QeqBuffers& qb, int it, int imax, double tol, hipStream_t s);
int qeq_mv_partials(int n);                                                                             // partials per product
void launch_qeq_finish(const State& st, QeqBuffers& qb, hipStream_t s);

// ---- reductions and integration (k_md.hip)
void launch_energy_reduce(const State& st, double* out /*kTerms + 6 + 1*/, hipStream_t s);
struct MdBuffers {
  double *vx, *vy, *vz, *mass;   // per atom (mass from type)
  double* nh;                    // thermostat state (see k_md.hip)
  double *x0, *y0, *z0;          // positions at the last rebuild
  double* partial;               // reduction scratch
};
void launch_md_half_kick(const State& st, MdBuffers& md, double dtf, hipStream_t s);
void launch_md_drift(const State& st, MdBuffers& md, double dt, hipStream_t s);
void launch_md_ref_positions(const State& st, MdBuffers& md, hipStream_t s);
// fused step kernels: see k_md.hip
void launch_md_step_start(const State& st, MdBuffers& md, double dtf, double dt, bool thermostat, bool pending, hipStream_t s);
int md_disp_partials();   // per-block partials of the kick-drift kernel's largest squared displacement (in md.partial)
void launch_md_gather_step_end(const State& st, MdBuffers& md, double dtf, bool thermostat, hipStream_t s);
void launch_md_scale(const State& st, MdBuffers& md, hipStream_t s);
void launch_pk_ref(const State& st, MdBuffers& md, const int* reb, hipStream_t s);
void launch_pk_disp(const State& st, MdBuffers& md, const int* sys_start, int S, double lim, int* reb, int* any, double* disp,
                    hipStream_t s);
void launch_pk_thermostat(const State& st, MdBuffers& md, const int* sys_start, int S, double* nh_all, hipStream_t s);
void launch_pk_energy(const State& st, const int* sys_start, int S, double* out, hipStream_t s);

void launch_stream_copy(const void* src, void* dst, size_t bytes, hipStream_t s);   // bytes a multiple of 16

}  // namespace rxa
