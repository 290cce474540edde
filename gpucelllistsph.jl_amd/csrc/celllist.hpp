// Cell binning (counting sort) and pair-list emission shared by the standalone
// GPUCellList and by the SPH step's rebuild.
#pragma once

#include "gclsph_internal.hpp"

namespace gclsph {

// Geometry of a fixed cell grid (src/celllist.jl:70-75): 1-based cell k along
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
  void alloc(int64_t n_);
  void set_capacity(int64_t cap, hipStream_t s);  // zero-filled
  void release();
};

// Stencil rows of the sorted item s: for each (dy, dz) the contiguous sorted
// range of cells [ix-1, ix+1] (clamped to the grid).
struct Rows {
  int n;
  int32_t b[9], e[9];
};

template <int D>
__device__ __forceinline__ Rows stencil_rows(const CellGeom& g, const int32_t* __restrict__ start, int32_t c) {
  Rows r;
  r.n = 0;
  int64_t gx = g.grid[0], gy = g.grid[1];
  int64_t ix = c % gx;
  int64_t rest = c / gx;
  int64_t iy = rest % gy;
  int64_t iz = rest / gy;
  int64_t xl = ix > 0 ? ix - 1 : 0, xh = ix + 1 < gx ? ix + 1 : gx - 1;
  int64_t zlo = D == 3 ? (iz > 0 ? iz - 1 : 0) : 0;
  int64_t zhi = D == 3 ? (iz + 1 < g.grid[2] ? iz + 1 : g.grid[2] - 1) : 0;
  for (int64_t z = zlo; z <= zhi; ++z)
    for (int64_t y = (iy > 0 ? iy - 1 : 0); y <= (iy + 1 < gy ? iy + 1 : gy - 1); ++y) {
      int64_t row = gx * (y + gy * z);
      r.b[r.n] = start[row + xl];
      r.e[r.n] = start[row + xh + 1];
      ++r.n;
    }
  return r;
}


// Bins positions `x` (D pointers, input order, element type double or float)
// into cells, and writes the sorted order (perm, skey, start).  The order
// inside a cell is by input index, so the result is deterministic.
// `prec` is 64 or 32; `use_off64` selects the Float64-offset cell map.
// `x_alt`/`sel` (nullable): a second set of position pointers, chosen on the
// device when *sel != 0 (the SPH state is double-buffered on the device).
void bin_sort(const CellGeom& g, int prec, const void* const* x, int64_t n, bool use_off64, SortBuffers& sb,
              const int* flag, hipStream_t s, const void* const* x_alt = nullptr, const int* sel = nullptr);

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
// original index.
void pair_emit(const CellGeom& g, int prec, const void* const* xs, const int32_t* start, const int32_t* skey,
               const int32_t* id, int64_t n, double dist, PairBuffers& pb, hipStream_t s);

}  // namespace gclsph
