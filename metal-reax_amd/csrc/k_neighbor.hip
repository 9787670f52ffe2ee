// Neighbour search on the device: wrap into the cell, cell binning (stable radix sort of cell keys), and the full "skin" list of
// every owned atom: all images (j, s) with |x_j + s.L - x_i| <= nonb_cut + skin, in a deterministic order (cells in a fixed stencil
// order, atoms ascending inside a cell). One 64-lane wave per atom; the lanes test 64 candidates at once and compact the hits with a
// ballot, so the list is written in candidate order without atomics.
//
// Reference semantics: the far list of src/neighbor/neighbor.cpp:180-212 (pairs within the row cut-off, measured as x_j - x_i) over
// the image-expanded set of neighbor.cpp:18-88. Here both directions of every pair are kept (a full list), measured from the pair's
// canonical end (common.hpp), and the consumers apply their own cut-offs (nonb_cut, swb, bond_cut, hbond_cut) to that distance.
#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "launch.hpp"

namespace rxa {

namespace {

#pragma clang fp contract(off)
__global__ void k_wrap(State st, Cell c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
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

// lattice shift vectors of all codes in [-smx, smx]^3, accumulated exactly as shift_vec does
__global__ void k_shift_table(Cell c, double4* shv, int smx0, int smx1, int smx2) {
  const int d0 = 2 * smx0 + 1, d1 = 2 * smx1 + 1, d2 = 2 * smx2 + 1;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d0 * d1 * d2) return;
  const int s0 = k % d0 - smx0, s1 = (k / d0) % d1 - smx1, s2 = k / (d0 * d1) - smx2;
  double sx, sy, sz;
  shift_vec(c, shift_code(s0, s1, s2), sx, sy, sz);
  shv[k] = make_double4(sx, sy, sz, 0.0);
}
#pragma clang fp contract(on)

__global__ void k_pos4(State st) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  st.pos[i] = make_double4(st.x[i], st.y[i], st.z[i], pos_w(st, i));
}

__global__ void k_cell_keys(State st, Cell c, int nc0, int nc1, int nc2, unsigned* keys, int* idx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  const int nc[3] = {nc0, nc1, nc2};
  int q[3];
  for (int d = 0; d < 3; ++d) {
    const double f = frac_coord(c, d, st.x[i], st.y[i], st.z[i]);
    int t = (int)floor(f * nc[d]);
    q[d] = t < 0 ? 0 : (t >= nc[d] ? nc[d] - 1 : t);
  }
  keys[i] = (unsigned)((q[2] * nc1 + q[1]) * nc0 + q[0]);
  idx[i] = i;
}

__global__ void k_fill_int(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_cell_ranges(int n, const unsigned* skeys, int* cstart, int* cend, unsigned* atom_key, const int* sidx) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const unsigned k = skeys[p];
  if (p == 0 || skeys[p - 1] != k) cstart[k] = p;
  if (p == n - 1 || skeys[p + 1] != k) cend[k] = p + 1;
  atom_key[sidx[p]] = k;
}

__global__ void __launch_bounds__(256) k_nbr_cells(State st, Cell c, int nc0, int nc1, int nc2, const int* sidx, const int* cstart,
                                                   const int* cend, const unsigned* atom_key, double rs2) {
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= st.n) return;
  const int lane = lane_id();
  const double4 pi = st.pos[i];
  const unsigned key = atom_key[i];
  const int c0 = (int)(key % (unsigned)nc0), c1 = (int)((key / (unsigned)nc0) % (unsigned)nc1), c2 = (int)(key / (unsigned)(nc0 * nc1));
  int count = 0;
  const size_t base = (size_t)i * st.maxnb;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int q0 = c0 + dx, q1 = c1 + dy, q2 = c2 + dz, s0 = 0, s1 = 0, s2 = 0;
        if (q0 < 0) { q0 += nc0; s0 = -1; } else if (q0 >= nc0) { q0 -= nc0; s0 = 1; }
        if (q1 < 0) { q1 += nc1; s1 = -1; } else if (q1 >= nc1) { q1 -= nc1; s1 = 1; }
        if (q2 < 0) { q2 += nc2; s2 = -1; } else if (q2 >= nc2) { q2 -= nc2; s2 = 1; }
        const int cell = (q2 * nc1 + q1) * nc0 + q0;
        const int sc = shift_code(s0, s1, s2);
        // the image's shift vector is the same for the whole cell; the list radius lies a skin beyond every consumer's
        // cut-off, so the test takes the image from this side (no canonical orientation) and only kept entries need the
        // tags that decide their canonical end
        const double4 S = st.shv[pos_sys(pi.w) * st.shv_stride + shift_index(st, sc)];
        const int pe = cend[cell];
        for (int p0 = cstart[cell]; p0 < pe; p0 += 64) {
          const int p = p0 + lane;
          bool keep = false;
          int j = 0;
          if (p < pe) {
            j = sidx[p];
            if (!(j == i && sc == 0)) {
              const double4 pj = st.pos[j];
              const double ex = (pj.x + S.x) - pi.x, ey = (pj.y + S.y) - pi.y, ez = (pj.z + S.z) - pi.z;
              keep = ex * ex + ey * ey + ez * ez <= rs2;
            }
          }
          const unsigned long long m = __ballot(keep);
          if (keep) {
            const int slot = count + lanes_below(m);
            const int e = canonical_first(i, j, sc, st.tag) ? (sc | kCanon) : sc;
            if (slot < st.maxnb) { st.nb_j[base + slot] = j; st.nb_s[base + slot] = e; }
          }
          count += __popcll(m);
        }
      }
  if (lane == 0) {
    st.nb_cnt[i] = count;
    if (count > st.maxnb) atomicMax(&st.flags[0], count);
  }
}

__global__ void __launch_bounds__(256) k_nbr_brute(State st, Cell c, int sm0, int sm1, int sm2, double rs2) {
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= st.n) return;
  const int lane = lane_id();
  const double4 pi = st.pos[i];
  int count = 0;
  const size_t base = (size_t)i * st.maxnb;
  for (int s2 = -sm2; s2 <= sm2; ++s2)
    for (int s1 = -sm1; s1 <= sm1; ++s1)
      for (int s0 = -sm0; s0 <= sm0; ++s0) {
        const int sc = shift_code(s0, s1, s2);
        for (int j0 = 0; j0 < st.n; j0 += 64) {
          const int j = j0 + lane;
          bool keep = false;
          int e = 0;
          if (j < st.n && !(j == i && sc == 0)) {
            e = canonical_first(i, j, sc, st.tag) ? (sc | kCanon) : sc;
            double ddx, ddy, ddz;
            keep = pair_dvec_list(st, pi, j, e, ddx, ddy, ddz) <= rs2;
          }
          const unsigned long long m = __ballot(keep);
          if (keep) {
            const int slot = count + lanes_below(m);
            if (slot < st.maxnb) { st.nb_j[base + slot] = j; st.nb_s[base + slot] = e; }
          }
          count += __popcll(m);
        }
      }
  if (lane == 0) {
    st.nb_cnt[i] = count;
    if (count > st.maxnb) atomicMax(&st.flags[0], count);
  }
}

// Bond and H-bond candidate sub-lists of the skin list (entry order kept): candidates within bond_cut + skin of every typed atom,
// and for H-bond donors (p_hbond 1) the acceptors (p_hbond 2) within hbond_cut + skin. Per-step kernels scan
// these instead of the whole skin list and apply the exact cut-offs to the current distance.
__global__ void __launch_bounds__(256) k_sublists(State st, Globals g, Tables t, double rb2, double rh2) {
  const int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (i >= st.n) return;
  const int lane = lane_id();
  const int ti = st.type[i];
  int nbc = 0, nhc = 0;
  if (ti >= 0) {
    const double4 pi = st.pos[i];
    const int role = g.hbond_cut > 0 ? t.sp[ti].p_hbond : 0;
    const size_t base = (size_t)i * st.maxnb, bb = (size_t)i * st.maxbc, hb = (size_t)i * st.maxhc;
    const int cnt = min(st.nb_cnt[i], st.maxnb);
    for (int k0 = 0; k0 < cnt; k0 += 64) {
      const int k = k0 + lane;
      bool bk = false, hk = false;
      int j = 0, e = 0;
      if (k < cnt) {
        j = st.nb_j[base + k];
        e = st.nb_s[base + k];
        const int tj = st.type[j];
        if (tj >= 0) {
          // list radii with a skin beyond the consumers' cut-offs: the image from this side, no canonical orientation
          const double4 pj = st.pos[j];
          const double4 S = st.shv[pos_sys(pi.w) * st.shv_stride + shift_index(st, entry_code(e))];
          const double ex = (pj.x + S.x) - pi.x, ey = (pj.y + S.y) - pi.y, ez = (pj.z + S.z) - pi.z;
          const double d2 = ex * ex + ey * ey + ez * ez;
          bk = d2 <= rb2;
          if (role == 1 && d2 <= rh2) hk = t.sp[tj].p_hbond == 2;   // donors list acceptors; acceptor rows come from the transpose
        }
      }
      const unsigned long long mb = __ballot(bk), mh = __ballot(hk);
      if (bk) {
        const int p = nbc + lanes_below(mb);
        if (p < st.maxbc) { st.bc_j[bb + p] = j; st.bc_s[bb + p] = e; }
      }
      if (hk) {
        const int p = nhc + lanes_below(mh);
        if (p < st.maxhc) { st.hc_j[hb + p] = j; st.hc_s[hb + p] = e; }
      }
      nbc += __popcll(mb);
      nhc += __popcll(mh);
    }
  }
  if (lane == 0) {
    st.bc_cnt[i] = nbc;
    st.hc_cnt[i] = nhc;
    if (nbc > st.maxbc) atomicMax(&st.flags[4], nbc);
    if (nhc > st.maxhc) atomicMax(&st.flags[5], nhc);
  }
}

// Acceptor lists as the transpose of the donor lists: every donor candidate slot (flat index j * maxhc + h) is appended to its
// acceptor's row in arrival order (an integer atomic claims the position), then each row is put in ascending slot order by
// rank counting (the slot indices are distinct), so the acceptor's later force gather runs in a fixed order.
__global__ void __launch_bounds__(256) k_hc_transpose(State st, Globals g, Tables t) {
  const int j = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (j >= st.n) return;
  const int tj = st.type[j];
  if (tj < 0 || t.sp[tj].p_hbond != 1) return;
  const int nc = min(st.hc_cnt[j], st.maxhc);
  const size_t jb = (size_t)j * st.maxhc;
  for (int h = lane_id(); h < nc; h += 64) {
    const int k = st.hc_j[jb + h];
    const int p = atomicAdd(&st.hc_cnt[k], 1);
    if (p < st.maxhc) st.hc_m[(size_t)k * st.maxhc + p] = (int)(jb + h);
    else atomicMax(&st.flags[5], p + 1);
  }
}

// (kLds: the row's keys staged in the wave's share of dynamic LDS, maxhc ints per wave, so the rank loop reads LDS broadcasts;
// rows too long for LDS read the keys from global memory)
template <bool kLds>
__global__ void __launch_bounds__(256) k_hc_sort(State st, Globals g, Tables t) {
  extern __shared__ int hc_keys[];
  const int k = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  if (k >= st.n) return;
  const int tk = st.type[k];
  if (tk < 0 || t.sp[tk].p_hbond != 2) return;
  const int m = min(st.hc_cnt[k], st.maxhc);
  const size_t kb = (size_t)k * st.maxhc;
  const int* keys = st.hc_m + kb;
  if (kLds) {
    int* lk = hc_keys + (threadIdx.x >> 6) * st.maxhc;
    for (int h = lane_id(); h < m; h += 64) lk[h] = st.hc_m[kb + h];
    __builtin_amdgcn_wave_barrier();
    keys = lk;
  }
  for (int h = lane_id(); h < m; h += 64) {
    const int key = keys[h];
    int rank = 0;
    for (int q = 0; q < m; ++q) rank += keys[q] < key ? 1 : 0;
    st.hc_j[kb + rank] = key;
  }
}

// Gather of every persistent per-atom array into a new atom order: new[i] = old[src[i]] (through a scratch area, one kernel to
// gather, one to copy back). Used to keep atoms in cell order (spatial locality of every neighbour gather) and to carry state
// across a change of the caller's atom order.
__global__ void k_permute_gather(PermuteSet ps, const int* src, double* sd, int* si, long long* sl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ps.n) return;
  const int j = src[i];
  for (int k = 0; k < ps.nd; ++k) sd[(size_t)k * ps.n + i] = ps.d[k][j];
  for (int k = 0; k < ps.ni; ++k) si[(size_t)k * ps.n + i] = ps.i[k][j];
  for (int k = 0; k < ps.nl; ++k) sl[(size_t)k * ps.n + i] = ps.l[k][j];
}

__global__ void k_permute_scatter(PermuteSet ps, const double* sd, const int* si, const long long* sl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ps.n) return;
  for (int k = 0; k < ps.nd; ++k) ps.d[k][i] = sd[(size_t)k * ps.n + i];
  for (int k = 0; k < ps.ni; ++k) ps.i[k][i] = si[(size_t)k * ps.n + i];
  for (int k = 0; k < ps.nl; ++k) ps.l[k][i] = sl[(size_t)k * ps.n + i];
}

// after the atoms were put in the order of their cells: atom p sits at sorted position p
__global__ void k_bin_identity(int n, const unsigned* skeys, int* sidx, unsigned* atom_key) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  sidx[p] = p;
  atom_key[p] = skeys[p];
}

// positions from a caller's device arrays (x[n], y[n], z[n] in the caller's order) into the engine's order
__global__ void k_positions_in(State st, const double* ux, const int* perm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= st.n) return;
  const int u = perm[i];
  const double x = ux[u], y = ux[st.n + u], z = ux[2 * (size_t)st.n + u];
  st.x[i] = x; st.y[i] = y; st.z[i] = z;
  st.pos[i] = make_double4(x, y, z, pos_w(st, i));
}

inline int blocks_for(long long items, int per_block) { return (int)((items + per_block - 1) / per_block); }

}  // namespace

void launch_wrap(const State& st, const Cell& c, hipStream_t s) {
  if (st.n == 0) return;
  hipLaunchKernelGGL(k_wrap, dim3(blocks_for(st.n, 256)), dim3(256), 0, s, st, c);
}

void launch_permute(const PermuteSet& ps, const int* src, double* sd, int* si, long long* sl, hipStream_t s) {
  if (ps.n == 0) return;
  hipLaunchKernelGGL(k_permute_gather, dim3(blocks_for(ps.n, 256)), dim3(256), 0, s, ps, src, sd, si, sl);
  hipLaunchKernelGGL(k_permute_scatter, dim3(blocks_for(ps.n, 256)), dim3(256), 0, s, ps, sd, si, sl);
}

void launch_bin_identity(int n, BinBuffers& b, hipStream_t s) {
  hipLaunchKernelGGL(k_bin_identity, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, b.skeys, b.sidx, b.atom_key);
}

void launch_positions_in(const State& st, const double* ux, const int* perm, hipStream_t s) {
  hipLaunchKernelGGL(k_positions_in, dim3(blocks_for(st.n, 256)), dim3(256), 0, s, st, ux, perm);
}

void launch_pos4(const State& st, hipStream_t s) {
  if (st.n == 0) return;
  hipLaunchKernelGGL(k_pos4, dim3(blocks_for(st.n, 256)), dim3(256), 0, s, st);
}

void launch_shift_table(const Cell& c, double4* shv, const int smx[3], hipStream_t s) {
  const int m = (2 * smx[0] + 1) * (2 * smx[1] + 1) * (2 * smx[2] + 1);
  hipLaunchKernelGGL(k_shift_table, dim3(blocks_for(m, 256)), dim3(256), 0, s, c, shv, smx[0], smx[1], smx[2]);
}

void launch_sublists(const State& st, const Globals& g, const Tables& t, double rb, double rh, hipStream_t s) {
  const double m = 1.0 + 1e-12;
  hipLaunchKernelGGL(k_sublists, dim3(blocks_for((long long)st.n * 64, 256)), dim3(256), 0, s, st, g, t, rb * rb * m, rh * rh * m);
  if (rh > 0) {
    hipLaunchKernelGGL(k_hc_transpose, dim3(blocks_for((long long)st.n * 64, 256)), dim3(256), 0, s, st, g, t);
    const size_t lds = 4 * (size_t)st.maxhc * sizeof(int);   // four waves per block
    if (lds <= 64 * 1024) hipLaunchKernelGGL(k_hc_sort<true>, dim3(blocks_for((long long)st.n * 64, 256)), dim3(256), lds, s, st, g, t);
    else hipLaunchKernelGGL(k_hc_sort<false>, dim3(blocks_for((long long)st.n * 64, 256)), dim3(256), 0, s, st, g, t);
  }
}

void launch_fill_int(int* p, int n, int v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_int, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, n, v);
}

size_t bin_temp_bytes(int n) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned*)nullptr, (unsigned*)nullptr, (const int*)nullptr, (int*)nullptr, n);
  return bytes;
}

void launch_bin(const State& st, const Cell& c, const BinGrid& g, BinBuffers& b, hipStream_t s) {
  const int n = st.n;
  hipLaunchKernelGGL(k_cell_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, st, c, g.nc[0], g.nc[1], g.nc[2], b.keys, b.idx);
  size_t bytes = b.temp_bytes;
  int bits = 1;
  while ((1u << bits) < (unsigned)g.ncells()) ++bits;
  (void)hipcub::DeviceRadixSort::SortPairs(b.temp, bytes, b.keys, b.skeys, b.idx, b.sidx, n, 0, bits, s);
  launch_fill_int(b.cstart, g.ncells(), 0, s);
  launch_fill_int(b.cend, g.ncells(), 0, s);
  hipLaunchKernelGGL(k_cell_ranges, dim3(blocks_for(n, 256)), dim3(256), 0, s, n, b.skeys, b.cstart, b.cend, b.atom_key, b.sidx);
}

void launch_nbr_cells(const State& st, const Cell& c, const BinGrid& g, const BinBuffers& b, double rs, hipStream_t s) {
  hipLaunchKernelGGL(k_nbr_cells, dim3(blocks_for((long long)st.n * 64, 256)), dim3(256), 0, s, st, c, g.nc[0], g.nc[1], g.nc[2], b.sidx,
                     b.cstart, b.cend, b.atom_key, rs * rs);
}

void launch_nbr_brute(const State& st, const Cell& c, const int sm[3], double rs, hipStream_t s) {
  hipLaunchKernelGGL(k_nbr_brute, dim3(blocks_for((long long)st.n * 64, 256)), dim3(256), 0, s, st, c, sm[0], sm[1], sm[2], rs * rs);
}

}  // namespace rxa
