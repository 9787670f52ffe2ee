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
}
#pragma clang fp contract(on)

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
        const int pe = cend[cell];
        for (int p0 = cstart[cell]; p0 < pe; p0 += 64) {
          const int p = p0 + lane;
          bool keep = false;
          int j = 0;
          if (p < pe) {
            j = sidx[p];
            if (!(j == i && sc == 0)) {
              double ddx, ddy, ddz;
              keep = pair_dvec(st, c, i, j, sc, ddx, ddy, ddz) <= rs2;
            }
          }
          const unsigned long long m = __ballot(keep);
          if (keep) {
            const int slot = count + lanes_below(m);
            if (slot < st.maxnb) { st.nb_j[base + slot] = j; st.nb_s[base + slot] = sc; }
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
  int count = 0;
  const size_t base = (size_t)i * st.maxnb;
  for (int s2 = -sm2; s2 <= sm2; ++s2)
    for (int s1 = -sm1; s1 <= sm1; ++s1)
      for (int s0 = -sm0; s0 <= sm0; ++s0) {
        const int sc = shift_code(s0, s1, s2);
        for (int j0 = 0; j0 < st.n; j0 += 64) {
          const int j = j0 + lane;
          bool keep = false;
          if (j < st.n && !(j == i && sc == 0)) {
            double ddx, ddy, ddz;
            keep = pair_dvec(st, c, i, j, sc, ddx, ddy, ddz) <= rs2;
          }
          const unsigned long long m = __ballot(keep);
          if (keep) {
            const int slot = count + lanes_below(m);
            if (slot < st.maxnb) { st.nb_j[base + slot] = j; st.nb_s[base + slot] = sc; }
          }
          count += __popcll(m);
        }
      }
  if (lane == 0) {
    st.nb_cnt[i] = count;
    if (count > st.maxnb) atomicMax(&st.flags[0], count);
  }
}

inline int blocks_for(long long items, int per_block) { return (int)((items + per_block - 1) / per_block); }

}  // namespace

void launch_wrap(const State& st, const Cell& c, hipStream_t s) {
  if (st.n == 0) return;
  hipLaunchKernelGGL(k_wrap, dim3(blocks_for(st.n, 256)), dim3(256), 0, s, st, c);
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
