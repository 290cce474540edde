// gcl_neighborlist_noalloc: the neighbour list of a point set without any
// device allocation, for callers that cannot allocate (code that must not call
// hipMalloc: stream-captured sequences, tight serving loops, harnesses that
// forbid it).  It runs the kernels of GPUCellList + update! + neighborlist
// (binning counting sort, pair count, scan, emit) with their intermediates in
// `scratch`, a file-scope device array of fixed size, instead of the buffers a
// GPUCellList object allocates.  Inputs (device coordinates) and the output
// pair buffer come from the caller.
//
// Every call on a device shares that device's scratch and keeps the last
// call's contents.  Calls from several host threads on one device are
// serialised by a host mutex; calls that overlap on one device otherwise
// (another process, or work the caller enqueues on the same scratch) race.
#include <array>
#include <cmath>
#include <limits>
#include <mutex>

#include "../../include/gclsph.h"
#include "blockscan.hpp"
#include "sph.hpp"

namespace gclsph {
namespace {

constexpr int64_t kNaMaxPoints = GCLSPH_NOALLOC_MAX_POINTS;
constexpr int64_t kNaMaxCells = GCLSPH_NOALLOC_MAX_CELLS;

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t scan_elems(int64_t n) {  // scan_tmp_elems() of scan.hip, for the fixed maxima
  return size_t((n + kTile - 1) / kTile + 2 + ((n + kTile - 1) / kTile + 1) / 2 + 1);
}

// Layout of the scratch array (each piece 256-byte aligned): the counting
// sort (key, slot, tmp, perm, skey; cell histogram and starts; scan tiles;
// max-per-cell), the sorted coordinates (3 components with the padding the
// pair windows read past the end), and the pair counts, offsets and scan tiles.
struct Layout {
  size_t key, slot, tmp, perm, skey, cnt, start, sort_scan, maxcell, xs[3], pcnt, poff, pair_scan, total;
};
constexpr Layout make_layout() {
  Layout L{};
  size_t o = 0;
  auto take = [&o](size_t bytes) {
    size_t at = o;
    o += align256(bytes);
    return at;
  };
  L.key = take(4 * size_t(kNaMaxPoints));
  L.slot = take(4 * size_t(kNaMaxPoints));
  L.tmp = take(4 * size_t(kNaMaxPoints));
  L.perm = take(4 * size_t(kNaMaxPoints));
  L.skey = take(4 * size_t(kNaMaxPoints));
  L.cnt = take(4 * size_t(kNaMaxCells + 1));
  L.start = take(4 * size_t(kNaMaxCells + 1));
  L.sort_scan = take(8 * scan_elems(kNaMaxCells));
  L.maxcell = take(16);
  for (int k = 0; k < 3; ++k) L.xs[k] = take(8 * size_t(kNaMaxPoints + kPosPad));
  L.pcnt = take(4 * size_t(kNaMaxPoints + 1));
  L.poff = take(8 * size_t(kNaMaxPoints + 1));
  L.pair_scan = take(8 * scan_elems(kNaMaxPoints));
  L.total = o;
  return L;
}
constexpr Layout kLayout = make_layout();

// The scratch memory (38 MB at the default maxima).  In host code this name
// is only a stand-in for the device symbol: the device address comes from
// hipGetSymbolAddress, per device.
__device__ __align__(256) unsigned char scratch[kLayout.total];

std::mutex g_mutex;
std::array<unsigned char*, 64> g_scratch_ptr{};  // device address of `scratch`, per device ordinal

unsigned char* scratch_on_current_device(int dev) {
  GCL_CHECK(dev >= 0 && dev < int(g_scratch_ptr.size()), kErrUnsupported, "device ordinal beyond 63");
  if (!g_scratch_ptr[size_t(dev)]) {
    void* p = nullptr;
    GCL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GCL_CHECK(p != nullptr, kErrHip, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[size_t(dev)] = static_cast<unsigned char*>(p);
  }
  return g_scratch_ptr[size_t(dev)];
}

}  // namespace
}  // namespace gclsph

using namespace gclsph;

extern "C" int gcl_neighborlist_noalloc(int dim, int64_t n, const double* const* pts_dev, const double* cellsize,
                                        double dist, const double* min, const double* max, int32_t* pairs_dev,
                                        int64_t capacity, int64_t* pairsn, void* stream) {
  int rc = GCLSPH_OK;
  try {
    GCL_CHECK(dim == 2 || dim == 3, kErrArgument, "wrong dimention");
    GCL_CHECK(n > 0 && n <= kNaMaxPoints, kErrArgument,
              "number of points must be in [1, GCLSPH_NOALLOC_MAX_POINTS] (524288)");
    GCL_CHECK(pts_dev && cellsize && min && max && pairsn, kErrArgument, "points / cellsize / min / max / pairsn is NULL");
    for (int k = 0; k < dim; ++k) GCL_CHECK(pts_dev[k] != nullptr, kErrArgument, "points component is NULL");
    GCL_CHECK(capacity >= 0 && (capacity == 0 || pairs_dev != nullptr), kErrArgument,
              "pair buffer is NULL or its capacity negative");
    GCL_CHECK(reinterpret_cast<uintptr_t>(pairs_dev) % 8 == 0, kErrArgument,
              "pair buffer must be 8-byte aligned (one (i, j) pair per 8 bytes)");
    GCL_CHECK(dist > 0 && std::isfinite(dist), kErrArgument, "dist must be positive and finite");
    // The reference constructor's grid with MIN and MAX given (src/celllist.jl:27-46):
    // a cell smaller than dist is raised to dist, offset MIN - |(MIN + √eps)·√eps|,
    // grid = ceil((MAX - offset) / cellsize).
    CellGeom g{};
    g.dim = dim;
    for (int k = 0; k < 3; ++k) g.grid[k] = 1;
    const double sq = std::sqrt(std::numeric_limits<double>::epsilon());
    for (int k = 0; k < dim; ++k) {
      const double cs = cellsize[k] < dist ? dist : cellsize[k];
      const double mn = min[k] - std::fabs((min[k] + sq) * sq);
      const double cells = std::ceil((max[k] - mn) / cs);
      GCL_CHECK(cells >= 1 && cells <= double(kNaMaxCells), kErrArgument,
                "cell grid has no cells, or more than GCLSPH_NOALLOC_MAX_CELLS (1048576), along a dimension");
      g.grid[k] = int64_t(cells);
      g.off64[k] = mn;
      g.off32[k] = float(mn);
      g.hinv64[k] = 1.0 / cs;
      g.hinv32[k] = float(g.hinv64[k]);
    }
    GCL_CHECK(g.ncells() <= kNaMaxCells, kErrArgument, "cell grid exceeds GCLSPH_NOALLOC_MAX_CELLS (1048576) cells");

    std::lock_guard<std::mutex> lock(g_mutex);
    int dev = 0;
    GCL_HIP(hipGetDevice(&dev));
    unsigned char* base = scratch_on_current_device(dev);
    hipStream_t s = static_cast<hipStream_t>(stream);

    SortBuffers sb;  // views into the scratch; never allocated or released here
    sb.n = n;
    sb.ncells = g.ncells();
    sb.key = reinterpret_cast<int32_t*>(base + kLayout.key);
    sb.slot = reinterpret_cast<int32_t*>(base + kLayout.slot);
    sb.tmp = reinterpret_cast<int32_t*>(base + kLayout.tmp);
    sb.perm = reinterpret_cast<int32_t*>(base + kLayout.perm);
    sb.skey = reinterpret_cast<int32_t*>(base + kLayout.skey);
    sb.cnt = reinterpret_cast<int32_t*>(base + kLayout.cnt);
    sb.start = reinterpret_cast<int32_t*>(base + kLayout.start);
    sb.scan_tmp = reinterpret_cast<int64_t*>(base + kLayout.sort_scan);
    sb.maxcell = reinterpret_cast<int32_t*>(base + kLayout.maxcell);
    sb.sorted = false;  // visit the points in input order (the scratch holds another call's order)
    PairBuffers pb;
    pb.n = n;
    pb.cnt = reinterpret_cast<int32_t*>(base + kLayout.pcnt);
    pb.off = reinterpret_cast<int64_t*>(base + kLayout.poff);
    pb.scan_tmp = reinterpret_cast<int64_t*>(base + kLayout.pair_scan);
    pb.pairs = pairs_dev;
    pb.capacity = capacity;
    void* xs[3] = {base + kLayout.xs[0], base + kLayout.xs[1], base + kLayout.xs[2]};
    const void* x[3] = {pts_dev[0], pts_dev[1], dim == 3 ? pts_dev[2] : pts_dev[1]};

    // cellmap! + fillcells_naive! with the Float64 MIN, as the constructor
    // bins (src/celllist.jl:55-66), then the pair search (:72-75)
    bin_sort(g, GCLSPH_F64, x, n, true, sb, nullptr, s, nullptr, nullptr, xs);
    pair_count(g, GCLSPH_F64, xs, sb.start, sb.skey, n, dist, pb, s);
    int64_t total = 0;
    GCL_HIP(hipMemcpyAsync(&total, pb.off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    GCL_HIP(hipStreamSynchronize(s));
    pair_emit(g, GCLSPH_F64, xs, sb.start, sb.skey, sb.perm, n, dist, total, pb, s);
    GCL_HIP(hipStreamSynchronize(s));
    *pairsn = total;
  } catch (const Error& e) {
    rc = e.code;
    gclsph::set_last_error(e.what());
  } catch (const std::exception& e) {
    rc = kErrHip;
    gclsph::set_last_error(e.what());
  }
  if (rc == GCLSPH_OK) gclsph::set_last_error("");
  return rc;
}
