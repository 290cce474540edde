// Cell binning (counting sort) and pair-list emission shared by the standalone
// GPUCellList and by the SPH step's rebuild.
#pragma once

#include "gclsph_internal.hpp"

namespace gclsph {

// Geometry of a fixed cell grid (src/celllist.jl:41-46): 1-based cell k along
// an axis holds p = (x - off) * (1/cs) in (k-1, k], clamped into [1, grid].
struct CellGeom {
  int dim;
  int64_t grid[3];
  double off64[3];  // offset in Float64 (constructor path)
  float off32[3];   // offset rounded to Float32 (stored-field path, T = Float32)
  double hinv64[3];  // 1/cs computed in T, widened
  float hinv32[3];
  __host__ __device__ int64_t ncells() const { return grid[0] * grid[1] * (dim == 3 ? grid[2] : 1); }
};

template <typename T>
struct PrecTraits;
template <>
struct PrecTraits<double> {
  static __host__ __device__ double off(const CellGeom& g, int k) { return g.off64[k]; }
  static __host__ __device__ double hinv(const CellGeom& g, int k) { return g.hinv64[k]; }
};
template <>
struct PrecTraits<float> {
  static __host__ __device__ float off(const CellGeom& g, int k) { return g.off32[k]; }
  static __host__ __device__ float hinv(const CellGeom& g, int k) { return g.hinv32[k]; }
};

// 1-based cell coordinate along axis k (src/gpukernels_cls.jl:29-47): p = (x - off)·(1/cs) in T
// (or with the Float64 offset), clamped into [1, grid], rounded up.
template <typename T>
__device__ __forceinline__ int32_t cell_coord(T xv, const CellGeom& g, int k, int use_off64) {
#pragma clang fp contract(off)
  double p;
  if (use_off64) {
    p = (double(xv) - g.off64[k]) * double(PrecTraits<T>::hinv(g, k));
  } else {
    T pt = (xv - PrecTraits<T>::off(g, k)) * PrecTraits<T>::hinv(g, k);
    p = double(pt);
  }
  // min(max(p, 1), grid); a NaN coordinate lands in cell 1 so every index
  // stays inside the grid (the reference leaves it undefined)
  double cl = p >= 1.0 ? p : 1.0;
  double gk = double(g.grid[k]);
  cl = cl <= gk ? cl : gk;
  return int32_t(ceil(cl));
}

// Arrival slot of key k in its cell's counter: cnt[k] += 1 for every active
// lane.  Input in (old) cell order makes the lanes of a wave share a few cells,
// so one atomic per distinct cell of the wave serves all its lanes: the peer
// groups are found first (a register-only loop), then the group leaders issue
// their atomics together and the wave waits for one round trip.
__device__ __forceinline__ int32_t wave_aggregated_slot(int32_t* cnt, int32_t k) {
  const int lane = threadIdx.x & 63;
  unsigned long long todo = __ballot(1), mine = 0;
  int leader = lane;
  while (todo) {
    const int l = __ffsll((long long)todo) - 1;
    const int32_t lk = __shfl(k, l, 64);
    const unsigned long long peers = __ballot(k == lk) & todo;
    if (k == lk) {
      mine = peers;
      leader = l;
    }
    todo &= ~peers;
  }
  int32_t base = 0;
  if (lane == leader) base = atomicAdd(&cnt[k], __popcll(mine));
  base = __shfl(base, leader, 64);
  return base + __popcll(mine & ((1ull << lane) - 1));
}

// Scratch for one counting sort of n items into ncells cells.
struct SortBuffers {
  int64_t n = 0, ncells = 0;
  int32_t* key = nullptr;    // cell of each input item
  int32_t* slot = nullptr;   // arrival slot inside its cell (atomic, unordered)
  int32_t* cnt = nullptr;    // ncells + 1 histogram
  int32_t* start = nullptr;  // ncells + 1 exclusive scan of cnt
  int32_t* tmp = nullptr;    // n: input index per unordered slot
  int32_t* perm = nullptr;   // n: sorted position -> input index
  int32_t* skey = nullptr;   // n: cell of each sorted position
  int64_t* scan_tmp = nullptr;
  int32_t* maxcell = nullptr;  // 1: max particles per cell of the last binning
  bool sorted = false;         // perm holds a previous binning's order
  void alloc(int64_t n_, int64_t ncells_);
  void release();
};

// Pair-list buffers: per-sorted-item counts, offsets and the (i,j) pairs.
struct PairBuffers {
  int64_t n = 0;
  int32_t* cnt = nullptr;    // n + 1
  int64_t* off = nullptr;    // n + 1
  int64_t* scan_tmp = nullptr;
  int32_t* pairs = nullptr;  // 2 * capacity
  int64_t capacity = 0;
  // 2D: the neighbour windows the count pass found, read back by the emit
  // pass (entcap per item; 0: the emit pass tests the candidates again)
  uint64_t* ent = nullptr;  // entcap * n
  int32_t* nent = nullptr;  // n
  int entcap = 0;
  void alloc(int64_t n_, int entcap_ = 0);
  void set_capacity(int64_t cap, hipStream_t s);  // zero-filled
  // Zeroes the entries [from, capacity).  A search with fill-zero leaves the
  // pairs in [0, pairsn) and zeros after them: the emit writes the front, so
  // only the tail needs the fill.
  void zero_from(int64_t from, hipStream_t s);
  void release();
};

// Stencil rows of a sorted item in cell c: for each (dy, dz) in {-1,0,1}^(D-1)
// the contiguous sorted range of cells [ix-1, ix+1] clamped to the grid (cells
// are ordered x fastest).  Rows outside the grid are empty (b == e).  The arrays
// have a fixed size and are filled by unrolled code, so they stay in registers.
template <int D>
struct Rows {
  static constexpr int N = D == 2 ? 3 : 9;
  int32_t b[N], e[N];
};

// Grid coordinates of cell c (x fastest).  Cell ids are non-negative int32, so
// 32-bit unsigned division decodes them; in 2D c / gx is already y.
template <int D>
__device__ __forceinline__ void cell_xyz(const CellGeom& g, int32_t c, int64_t& ix, int64_t& iy, int64_t& iz) {
  const uint32_t ugx = uint32_t(g.grid[0]), ugy = uint32_t(g.grid[1]), uc = uint32_t(c);
  const uint32_t rest = uc / ugx;
  ix = int64_t(uc - rest * ugx);
  iy = D == 3 ? int64_t(rest % ugy) : int64_t(rest);
  iz = D == 3 ? int64_t(rest / ugy) : 0;
}

template <int D>
__device__ __forceinline__ Rows<D> stencil_rows(const CellGeom& g, const int32_t* __restrict__ start, int32_t c) {
  Rows<D> r;
  const int64_t gx = g.grid[0], gy = g.grid[1];
  int64_t ix, iy, iz;
  cell_xyz<D>(g, c, ix, iy, iz);
  const int64_t xl = ix > 0 ? ix - 1 : 0, xh = ix + 1 < gx ? ix + 1 : gx - 1;
#pragma unroll
  for (int k = 0; k < Rows<D>::N; ++k) {
    const int64_t y = iy + (k % 3) - 1;
    const int64_t z = D == 3 ? iz + (k / 3) - 1 : 0;
    const bool in = y >= 0 && y < gy && (D == 2 || (z >= 0 && z < g.grid[2]));
    const int64_t row = gx * (y + gy * z);
    r.b[k] = in ? start[row + xl] : 0;
    r.e[k] = in ? start[row + xh + 1] : 0;
  }
  return r;
}

// The union of the stencil rows of `span` consecutive particles from a0 (a
// block or a wave) in one plane (2D: the only plane; 3D: the plane dz ∈ {-1,
// 0, 1}).  The particles a0..a1 occupy one contiguous range of cells [c0, c1],
// so for each row offset k the cells [c0 + o - 1, c1 + o + 1], o = (k-1)·gx +
// dz·gx·gy, hold every candidate of that row of every one of them: three
// contiguous ranges of sorted particles starting at ub[k], laid out one after
// the other from LDS slot uo[k].
struct RowUnion {
  int32_t ub[3], uo[4];
};

__device__ __forceinline__ void row_union(RowUnion& u, const int32_t* start, const int32_t* skey, const CellGeom& G,
                                          int64_t n, int64_t a0, int dz = 0, int span = kBlock) {
  const int64_t a1 = (a0 + span < n ? a0 + span : n) - 1;
  const int64_t c0 = skey[a0], c1 = skey[a1];
  const int64_t gx = G.grid[0], nc = G.ncells();
  const int64_t zo = int64_t(dz) * gx * G.grid[1];
  int32_t o = 0;
  for (int k = 0; k < 3; ++k) {
    int64_t lo = c0 + (k - 1) * gx + zo - 1, hi = c1 + (k - 1) * gx + zo + 1;
    lo = lo < 0 ? 0 : (lo > nc - 1 ? nc - 1 : lo);
    hi = hi < 0 ? 0 : (hi > nc - 1 ? nc - 1 : hi);
    const int32_t b = start[lo], e = start[hi + 1];
    u.ub[k] = b;
    u.uo[k] = o;
    o += e > b ? e - b : 0;
  }
  u.uo[3] = o;
}

// Bins positions `x` (D pointers, input order, element type double or float)
// into cells, and writes the sorted order (perm, skey, start).  The order
// inside a cell is by input index, so the result is deterministic.
// `prec` is 64 or 32; `use_off64` selects the Float64-offset cell map.
// `x_alt`/`sel` (nullable): a second set of position pointers, chosen on the
// device when *sel != 0 (the SPH state is double-buffered on the device).
// `xs` (nullable): D arrays that receive the positions in sorted order (what
// gather_positions would write), moved by the sort's rank kernel.
void bin_sort(const CellGeom& g, int prec, const void* const* x, int64_t n, bool use_off64, SortBuffers& sb,
              const int* flag, hipStream_t s, const void* const* x_alt = nullptr, const int* sel = nullptr,
              void* const* xs = nullptr);

// Gathers x (input order) into xs (sorted order) using perm.
void gather_positions(int dim, int prec, const void* const* x, void* const* xs, const int32_t* perm, int64_t n,
                      hipStream_t s);

// Counts, for every sorted item, its neighbours with a larger sorted index in
// the 3^D stencil and r^2 < dist^2 (exact reference expression, no FMA), and
// scans the counts.  Returns nothing; total lands in pb.off[n] (device).
void pair_count(const CellGeom& g, int prec, const void* const* xs, const int32_t* start, const int32_t* skey,
                int64_t n, double dist, PairBuffers& pb, hipStream_t s);

// Writes the pairs (1-based original indices, (min,max)) after pair_count.
// Pairs past pb.capacity are dropped.  `id` maps sorted position -> 0-based
// original index.  `total` is the count pair_count found: a 2D list with
// kEmitWaveMin or more pairs per item is emitted one wave per item, anything
// else one lane per item; both write the same pairs in the same order.  (One
// wave per item: 0.27 against 0.30 ms for 200,000 uniform points at 79 pairs
// per point; 2.05 against 0.82 ms at 18 pairs per point, and slower in 3D, where
// the 9 stencil rows leave most 64-candidate windows part-empty: 1.31 against
// 1.05 ms at 71 pairs per point.)  The environment variable
// GCLSPH_PAIR_EMIT=lane|wave forces one of them.
constexpr int64_t kEmitWaveMin = 64;
// Neighbour windows per item the 2D count pass keeps for the emit pass (an
// item of the 2D dam break at dist = H has 2-4; an item with more is tested
// again by the emit pass).  48 B per item, kept for cell lists of up to
// 2^27 points (6 GiB); larger ones test the candidates again.
constexpr int kPairEntries2D = 6;
constexpr int64_t kPairEntriesMaxPoints = int64_t(1) << 27;
void pair_emit(const CellGeom& g, int prec, const void* const* xs, const int32_t* start, const int32_t* skey,
               const int32_t* id, int64_t n, double dist, int64_t total, PairBuffers& pb, hipStream_t s);

}  // namespace gclsph
