// Cell binning and pair-list kernels.
//
// Binning replaces kernel_cellmap! + kernel_fillcells_naive!
// (src/gpukernels_cls.jl:29-101) by a counting sort: a cell histogram with an
// arrival slot per particle, an exclusive scan into cell start offsets, an
// unordered scatter, and a per-cell ranking by input index that makes the order
// deterministic.  The pair search replaces kernel_neib_internal!/external!
// (:309-554): every sorted particle scans the 3^D cells around it (the rows of a
// stencil are contiguous in cell-sorted order) and keeps partners with a larger
// sorted index, so each unordered pair is produced once.
#include <cstdlib>
#include <cstring>

#include "celllist.hpp"

namespace gclsph {
namespace {

// r^2 as the reference writes it, (a1-b1)^2 + (a2-b2)^2 [+ (a3-b3)^2], each
// operation rounded on its own (src/gpukernels_cls.jl:326,390,456,527).
template <typename T, int D>
__device__ __forceinline__ T dist2_exact(const T* const* x, int64_t a, int64_t b) {
#pragma clang fp contract(off)
  T d0 = x[0][a] - x[0][b];
  T d1 = x[1][a] - x[1][b];
  T s = d0 * d0 + d1 * d1;
  if (D == 3) {
    T d2 = x[2][a] - x[2][b];
    s = s + d2 * d2;
  }
  return s;
}

struct XPtrs {
  const void* p[3];
};

// One (i, j) pair of the output list, stored past the caches' normal
// allocation (non-temporal): the list is written once here and read later by
// the caller, so it should not push the binning and search data out of L2.
__device__ __forceinline__ void store_pair(int2* out, int64_t o, int32_t a, int32_t b) {
  using v2i = int __attribute__((ext_vector_type(2)));
  const v2i v = {a < b ? a : b, a < b ? b : a};
  __builtin_nontemporal_store(v, reinterpret_cast<v2i*>(out + o));
}

template <typename T, int D>
__global__ void k_cellmap(XPtrs xp, XPtrs xalt, const int* sel, int64_t n, CellGeom g, int use_off64,
                          int32_t* __restrict__ key, int32_t* __restrict__ slot, int32_t* __restrict__ cnt,
                          const int* flag, int aggregate, const int32_t* __restrict__ order) {
  if (flag && *flag == 0) return;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  // items in the previous binning's order when there is one: neighbouring
  // threads then mostly share a cell, which the wave-aggregated slots need
  const int64_t i = order ? int64_t(order[t]) : t;
  const XPtrs& src = (sel && *sel) ? xalt : xp;
  int32_t c[3] = {1, 1, 1};
#pragma unroll
  for (int k = 0; k < D; ++k) c[k] = cell_coord<T>(static_cast<const T*>(src.p[k])[i], g, k, use_off64);
  int64_t lin = int64_t(c[0] - 1) + g.grid[0] * (int64_t(c[1] - 1) + g.grid[1] * int64_t(c[2] - 1));
  const int32_t k = int32_t(lin);
  // cell and arrival slot in visiting order t (contiguous stores)
  key[t] = k;
  if (!aggregate) {
    slot[t] = atomicAdd(&cnt[lin], 1);
    return;
  }
  slot[t] = wave_aggregated_slot(cnt, k);
}

// Item visited t-th (order as in k_cellmap) to its arrival slot: tmp holds
// the item, skey its cell (every slot of a cell's range holds that cell, so
// the rank below need not write it).
__global__ void k_scatter(const int32_t* __restrict__ key, const int32_t* __restrict__ slot,
                          const int32_t* __restrict__ start, int32_t* __restrict__ tmp, int32_t* __restrict__ skey,
                          const int32_t* __restrict__ order, int64_t n, const int* flag) {
  if (flag && *flag == 0) return;
  int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int32_t c = key[t];
  const int32_t s = start[c] + slot[t];
  tmp[s] = order ? order[t] : int32_t(t);
  skey[s] = c;
}

// Rank each item inside its cell by input index: final order is (cell, input).
// Pos: the item's position also moves to its sorted place in xs (loaded with
// the key, so it arrives while the rank is counted), which replaces a
// separate gather of the positions.
template <typename T, int D, bool Pos>
__global__ void k_rank(const int32_t* __restrict__ start, const int32_t* __restrict__ tmp, int32_t* __restrict__ perm,
                       const int32_t* __restrict__ skey, int64_t n, const int* flag, XPtrs xp, XPtrs xalt,
                       const int* sel, XPtrs xsp) {
  if (flag && *flag == 0) return;
  int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int32_t i = tmp[s];
  T x[3];
  if constexpr (Pos) {
    const XPtrs& src = (sel && *sel) ? xalt : xp;
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = static_cast<const T*>(src.p[k])[i];
  }
  const int32_t c = skey[s];  // the scatter's, the same for every slot of the cell
  int32_t b = start[c], e = start[c + 1];
  // four loads in flight per trip (indices clamped into the cell, which holds
  // at least this item; the clamped ones are not counted)
  int32_t r = 0;
  for (int32_t k = b; k < e; k += 4) {
    int32_t v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = tmp[k + u < e ? k + u : e - 1];
#pragma unroll
    for (int u = 0; u < 4; ++u) r += (k + u < e) & (v[u] < i);
  }
  perm[b + r] = i;
  if constexpr (Pos) {
#pragma unroll
    for (int k = 0; k < D; ++k) static_cast<T*>(const_cast<void*>(xsp.p[k]))[b + r] = x[k];
  }
}

template <typename T, int D>
__global__ void k_gather_pos(XPtrs xp, XPtrs xsp, const int32_t* __restrict__ perm, int64_t n) {
  int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int32_t i = perm[s];
#pragma unroll
  for (int k = 0; k < D; ++k)
    const_cast<T*>(static_cast<const T*>(xsp.p[k]))[s] = static_cast<const T*>(xp.p[k])[i];
}

// Partner bits of the window [lo, hi) (hi - lo <= 32) of item s: bit j is set
// when r^2 < dist2max for candidate lo + j, with the expression of dist2_exact
// (no FMA).  Candidates are read four at a time so their loads are in flight
// together; reads run up to 3 past hi (the position arrays carry a padding of
// kPosPad elements), and those bits are dropped.
template <typename T, int D>
__device__ __forceinline__ uint32_t pair_window(const T* const* x, T xa0, T xa1, T xa2, int32_t lo, int32_t hi,
                                                T dist2max) {
#pragma clang fp contract(off)
  const int32_t len = hi - lo;
  uint32_t mk = 0;
  for (int32_t j = 0; j < len; j += 4) {
    T c0[4], c1[4], c2[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c0[u] = x[0][lo + j + u];
      c1[u] = x[1][lo + j + u];
      if (D == 3) c2[u] = x[2][lo + j + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const T d0 = xa0 - c0[u];
      const T d1 = xa1 - c1[u];
      T r2 = d0 * d0 + d1 * d1;
      if (D == 3) {
        const T d2 = xa2 - c2[u];
        r2 = r2 + d2 * d2;
      }
      mk |= uint32_t(r2 < dist2max) << (j + u);
    }
  }
  return len < 32 ? mk & ((1u << len) - 1u) : mk;
}

// Neighbours t > s of sorted item s, stencil row by row in 32-candidate
// windows (ascending t inside a row): count pass, then the emit pass writes
// them in the same order.
template <typename T, int D>
__global__ void k_pair_count(XPtrs xsp, const int32_t* __restrict__ start, const int32_t* __restrict__ skey,
                             int64_t n, CellGeom g, T dist2max, int32_t* __restrict__ cnt) {
  int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const T* x[3] = {static_cast<const T*>(xsp.p[0]), static_cast<const T*>(xsp.p[1]),
                   static_cast<const T*>(xsp.p[D == 3 ? 2 : 1])};
  const T xa0 = x[0][s], xa1 = x[1][s], xa2 = D == 3 ? x[2][s] : T(0);
  const Rows<D> r = stencil_rows<D>(g, start, skey[s]);
  int32_t c = 0;
#pragma unroll
  for (int k = 0; k < Rows<D>::N; ++k) {
    const int32_t b = r.b[k] > int32_t(s) + 1 ? r.b[k] : int32_t(s) + 1;
    for (int32_t w0 = b; w0 < r.e[k]; w0 += 32) {
      const int32_t hi = r.e[k] < w0 + 32 ? r.e[k] : w0 + 32;
      c += __builtin_popcount(pair_window<T, D>(x, xa0, xa1, xa2, w0, hi, dist2max));
    }
  }
  cnt[s] = c;
}

// 2D count pass that also stores the windows holding neighbours as
// (base << 32 | bits) entries, entry e of item s at ent[e*n + s], nent[s] of
// them (-1 when more than entcap), so the emit pass reads them instead of
// testing the candidates again.  Each wave stages the positions of rows
// 1 and 2 of its stencil-row union (row_union over its 64 items; the half
// stencil never reaches row 0, whose items all sort before the wave's) in its
// LDS slice, coalesced, and every lane tests its candidates there; a wave
// whose union exceeds kCountWave reads them from global memory.  The test is
// pair_window's expression either way, so the bits are the same.
constexpr int kCountWave = 512;

template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};

template <typename T>
__global__ __launch_bounds__(256) void k_pair_count2(XPtrs xsp, const int32_t* __restrict__ start,
                                                     const int32_t* __restrict__ skey, int64_t n, CellGeom g,
                                                     T dist2max, int32_t* __restrict__ cnt,
                                                     uint64_t* __restrict__ ent, int32_t* __restrict__ nent,
                                                     int entcap) {
#pragma clang fp contract(off)
  using T2 = typename Vec2<T>::type;
  __shared__ T2 lds[kBlock / 64][kCountWave + 4];  // +4: window reads run up to 3 past a row's end
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int lane = int(threadIdx.x & 63);
  const int64_t s0 = int64_t(blockIdx.x) * kBlock + 64 * w;
  if (s0 >= n) return;  // whole wave; no block barrier follows
  const int64_t s = s0 + lane;
  const T* x[3] = {static_cast<const T*>(xsp.p[0]), static_cast<const T*>(xsp.p[1]),
                   static_cast<const T*>(xsp.p[1])};
  RowUnion u;
  row_union(u, start, skey, g, n, s0, 0, 64);
  const int32_t total = u.uo[3] - u.uo[1];  // rows 1 and 2
  const bool staged = total <= kCountWave;   // uniform over the wave
  T2* wl = lds[w];
  if (staged) {
    const int32_t d1 = u.ub[1] - u.uo[1], d2 = u.ub[2] - u.uo[2];
    for (int32_t i = lane; i < total; i += 64) {
      const int32_t j = i + u.uo[1];
      const int32_t b = j + (j >= u.uo[2] ? d2 : d1);
      T2 v;
      v.x = x[0][b];
      v.y = x[1][b];
      wl[i] = v;
    }
    // the wave's LDS operations complete in issue order; the fence keeps the
    // compiler from moving the reads below ahead of these writes
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
  if (s >= n) return;
  const T xa0 = x[0][s], xa1 = x[1][s];
  const Rows<2> r = stencil_rows<2>(g, start, skey[s]);
  int32_t c = 0;
  int ne = 0;
#pragma unroll
  for (int k = 1; k < 3; ++k) {  // row 0 holds no candidate t > s
    const T2* lrow = wl + (u.uo[k] - u.uo[1] - u.ub[k]);  // candidate t at lrow[t]
    const int32_t b = r.b[k] > int32_t(s) + 1 ? r.b[k] : int32_t(s) + 1;
    for (int32_t w0 = b; w0 < r.e[k]; w0 += 32) {
      const int32_t hi = r.e[k] < w0 + 32 ? r.e[k] : w0 + 32;
      uint32_t mk;
      if (staged) {
        const int32_t len = hi - w0;
        mk = 0;
        for (int32_t j = 0; j < len; j += 4) {
          T2 cv[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) cv[q] = lrow[w0 + j + q];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const T d0 = xa0 - cv[q].x;
            const T d1 = xa1 - cv[q].y;
            const T r2 = d0 * d0 + d1 * d1;
            mk |= uint32_t(r2 < dist2max) << (j + q);
          }
        }
        if (len < 32) mk &= (1u << len) - 1u;
      } else {
        mk = pair_window<T, 2>(x, xa0, xa1, T(0), w0, hi, dist2max);
      }
      c += __builtin_popcount(mk);
      if (mk != 0) {
        if (ne < entcap)
          ent[int64_t(ne) * n + s] = (uint64_t(uint32_t(w0)) << 32) | mk;
        ++ne;
      }
    }
  }
  cnt[s] = c;
  nent[s] = ne <= entcap ? ne : -1;
}

// Pairs staged in LDS per wave before the wave writes them out: a wave's
// items are consecutive, so its pairs fill one contiguous range of the output
// [off[s0], off[s0 + 64]); staged, that range goes out in whole-line stores
// instead of one scattered 8-byte store per pair, and the partners' ids are
// looked up there too, a lane per pair (consecutive pairs share an item and
// have ascending partners, so the lookups of a wave touch few lines), two
// pairs per lane in 16-byte stores.  At 18 pairs per point on 4 M points: 0.25
// ms, against 0.29 with each item's lane looking up its partners' ids and
// 8-byte stores.  A wave with more pairs writes the rest directly.  40 KiB
// per block.
constexpr int kEmitStage = 1280;

template <typename T, int D, bool Stored>
__global__ __launch_bounds__(256) void k_pair_emit(XPtrs xsp, const int32_t* __restrict__ start,
                                                   const int32_t* __restrict__ skey, const int32_t* __restrict__ id,
                                                   int64_t n, CellGeom g, T dist2max, const int64_t* __restrict__ off,
                                                   int32_t* __restrict__ pairs, int64_t capacity,
                                                   const uint64_t* __restrict__ ent, const int32_t* __restrict__ nent) {
  __shared__ int2 stage[kBlock / 64][kEmitStage];
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int64_t s0 = int64_t(blockIdx.x) * kBlock + 64 * w;
  if (s0 >= n) return;  // whole wave; only wave-level ordering follows
  const int64_t s = s0 + (threadIdx.x & 63);
  const int64_t o0 = off[s0];
  const int64_t oend = off[s0 + 64 < n ? s0 + 64 : n];
  int2* st = stage[w];
  int2* out = reinterpret_cast<int2*>(pairs);
  if (s < n) {
    int64_t o = off[s];
    const int32_t is = id[s] + 1;
    // the neighbours in window [w0, w0 + 32) marked by mk, in ascending order;
    // staged as (is, sorted index of the neighbour): the ids are looked up in
    // the copy-out below, where consecutive lanes hold consecutive pairs
    auto emit = [&](int32_t w0, uint32_t mk) {
      for (; mk != 0 && o < capacity; mk &= mk - 1, ++o) {
        const int32_t t = w0 + __builtin_ctz(mk);
        if (o - o0 < kEmitStage) {
          st[o - o0] = make_int2(is, t);
        } else {
          const int32_t it = id[t] + 1;
          store_pair(out, o, is, it);
        }
      }
    };
    const int ne = Stored ? nent[s] : -1;
    if (ne >= 0) {
      for (int e = 0; e < ne && o < capacity; ++e) {
        const uint64_t en = ent[int64_t(e) * n + s];
        emit(int32_t(en >> 32), uint32_t(en));
      }
    } else {
      const T* x[3] = {static_cast<const T*>(xsp.p[0]), static_cast<const T*>(xsp.p[1]),
                       static_cast<const T*>(xsp.p[D == 3 ? 2 : 1])};
      const T xa0 = x[0][s], xa1 = x[1][s], xa2 = D == 3 ? x[2][s] : T(0);
      const Rows<D> r = stencil_rows<D>(g, start, skey[s]);
#pragma unroll
      for (int k = 0; k < Rows<D>::N; ++k) {
        const int32_t b = r.b[k] > int32_t(s) + 1 ? r.b[k] : int32_t(s) + 1;
        for (int32_t w0 = b; w0 < r.e[k] && o < capacity; w0 += 32) {
          const int32_t hi = r.e[k] < w0 + 32 ? r.e[k] : w0 + 32;
          emit(w0, pair_window<T, D>(x, xa0, xa1, xa2, w0, hi, dist2max));
        }
      }
    }
  }
  // the wave's LDS operations complete in issue order, so its lanes read what
  // the others staged; the fence keeps the compiler from hoisting the reads
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  int64_t ns = oend - o0 < kEmitStage ? oend - o0 : kEmitStage;
  if (ns > capacity - o0) ns = capacity - o0;
  auto pair_at = [&](int64_t i) {
    const int2 q = st[i];
    const int32_t it = id[q.y] + 1;
    return q.x < it ? make_int2(q.x, it) : make_int2(it, q.x);
  };
  // two pairs per lane in 16-byte stores from the first 16-byte aligned output
  // slot (pair buffers are 8-byte aligned); a first slot off that alignment
  // and an odd last pair go out on their own
  const int lane = int(threadIdx.x & 63);
  const int64_t h = (reinterpret_cast<uintptr_t>(out + o0) & 15) != 0 && ns > 0 ? 1 : 0;
  if (h && lane == 0) {
    const int2 q = pair_at(0);
    store_pair(out, o0, q.x, q.y);
  }
  const int64_t body = (ns - h) / 2;
  int4* out4 = reinterpret_cast<int4*>(out + o0 + h);
  for (int64_t j = lane; j < body; j += 64) {
    const int2 p0 = pair_at(h + 2 * j), p1 = pair_at(h + 2 * j + 1);
    using v4i = int __attribute__((ext_vector_type(4)));
    const v4i v = {p0.x, p0.y, p1.x, p1.y};
    __builtin_nontemporal_store(v, reinterpret_cast<v4i*>(out4 + j));  // the list is not read back here
  }
  if (h + 2 * body < ns && lane == 0) {
    const int2 q = pair_at(ns - 1);
    store_pair(out, o0 + ns - 1, q.x, q.y);
  }
}

// Dense lists: one wave per sorted item s.  The 64 lanes test 64 consecutive
// candidates of a stencil row at a time; the hits go out as one contiguous run
// at off[s] + (hits in lower lanes), so the stores coalesce.  Rows, windows and
// candidates are visited in ascending order, as in k_pair_emit, so both kernels
// write the same pairs to the same slots.
template <typename T, int D>
__global__ __launch_bounds__(256) void k_pair_emit_wave(XPtrs xsp, const int32_t* __restrict__ start,
                                                        const int32_t* __restrict__ skey,
                                                        const int32_t* __restrict__ id, int64_t n, CellGeom g,
                                                        T dist2max, const int64_t* __restrict__ off,
                                                        int32_t* __restrict__ pairs, int64_t capacity) {
#pragma clang fp contract(off)
  const int64_t s = int64_t(blockIdx.x) * (kBlock / 64) + __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  if (s >= n) return;  // whole wave
  const int32_t lane = int32_t(threadIdx.x & 63);
  const T* x[3] = {static_cast<const T*>(xsp.p[0]), static_cast<const T*>(xsp.p[1]),
                   static_cast<const T*>(xsp.p[D == 3 ? 2 : 1])};
  const T xa0 = x[0][s], xa1 = x[1][s], xa2 = D == 3 ? x[2][s] : T(0);
  const Rows<D> r = stencil_rows<D>(g, start, skey[s]);
  const int32_t is = id[s] + 1;
  int64_t o = off[s];
  int2* out = reinterpret_cast<int2*>(pairs);
#pragma unroll
  for (int k = 0; k < Rows<D>::N; ++k) {
    const int32_t b = r.b[k] > int32_t(s) + 1 ? r.b[k] : int32_t(s) + 1;
    for (int32_t w0 = b; w0 < r.e[k]; w0 += 64) {
      const int32_t j = w0 + lane;
      bool hit = false;
      if (j < r.e[k]) {
        const T d0 = xa0 - x[0][j];
        const T d1 = xa1 - x[1][j];
        T r2 = d0 * d0 + d1 * d1;
        if (D == 3) {
          const T d2 = xa2 - x[2][j];
          r2 = r2 + d2 * d2;
        }
        hit = r2 < dist2max;
      }
      const uint64_t m = __ballot(hit);
      if (hit) {
        const int64_t slot =
            o + __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
        if (slot < capacity) {
          const int32_t it = id[j] + 1;
          store_pair(out, slot, is, it);
        }
      }
      o += __builtin_popcountll(m);
    }
  }
}

template <typename T>
void bin_sort_t(const CellGeom& g, const void* const* x, int64_t n, bool use_off64, SortBuffers& sb, const int* flag,
                hipStream_t s, const void* const* x_alt, const int* sel, void* const* xs) {
  XPtrs xp{{x[0], x[1], g.dim == 3 ? x[2] : x[1]}};
  XPtrs xa = xp;
  if (x_alt) xa = XPtrs{{x_alt[0], x_alt[1], g.dim == 3 ? x_alt[2] : x_alt[1]}};
  int64_t nc = g.ncells();
  fill_i32(sb.cnt, nc + 1, 0, flag, s);
  if (g.dim == 2)
    k_cellmap<T, 2><<<grid_for(n), kBlock, 0, s>>>(xp, xa, sel, n, g, use_off64 ? 1 : 0, sb.key, sb.slot, sb.cnt, flag, 1,
                                                   sb.sorted ? sb.perm : nullptr);
  else
    k_cellmap<T, 3><<<grid_for(n), kBlock, 0, s>>>(xp, xa, sel, n, g, use_off64 ? 1 : 0, sb.key, sb.slot, sb.cnt, flag, 1,
                                                   sb.sorted ? sb.perm : nullptr);
  GCL_HIP(hipGetLastError());
  fill_i32(sb.maxcell, 1, 0, flag, s);
  exclusive_scan_i32(sb.cnt, sb.start, nc, sb.scan_tmp, flag, s, sb.maxcell);
  k_scatter<<<grid_for(n), kBlock, 0, s>>>(sb.key, sb.slot, sb.start, sb.tmp, sb.skey, sb.sorted ? sb.perm : nullptr, n,
                                           flag);
  GCL_HIP(hipGetLastError());
  const XPtrs xsp = xs ? XPtrs{{xs[0], xs[1], g.dim == 3 ? xs[2] : xs[1]}} : XPtrs{{nullptr, nullptr, nullptr}};
  if (!xs)
    k_rank<T, 2, false><<<grid_for(n), kBlock, 0, s>>>(sb.start, sb.tmp, sb.perm, sb.skey, n, flag, xp, xa, sel,
                                                        xsp);
  else if (g.dim == 2)
    k_rank<T, 2, true><<<grid_for(n), kBlock, 0, s>>>(sb.start, sb.tmp, sb.perm, sb.skey, n, flag, xp, xa, sel,
                                                       xsp);
  else
    k_rank<T, 3, true><<<grid_for(n), kBlock, 0, s>>>(sb.start, sb.tmp, sb.perm, sb.skey, n, flag, xp, xa, sel,
                                                       xsp);
  GCL_HIP(hipGetLastError());
  if (!flag) sb.sorted = true;  // perm now holds a permutation of [0, n) (a flagged sort may not have run)
}

}  // namespace

void SortBuffers::alloc(int64_t n_, int64_t ncells_) {
  release();
  sorted = false;
  n = n_;
  ncells = ncells_;
  GCL_HIP(hipMalloc(&key, sizeof(int32_t) * size_t(n)));
  GCL_HIP(hipMalloc(&slot, sizeof(int32_t) * size_t(n)));
  GCL_HIP(hipMalloc(&tmp, sizeof(int32_t) * size_t(n)));
  GCL_HIP(hipMalloc(&perm, sizeof(int32_t) * size_t(n)));
  GCL_HIP(hipMalloc(&skey, sizeof(int32_t) * size_t(n)));
  GCL_HIP(hipMalloc(&cnt, sizeof(int32_t) * size_t(ncells + 1)));
  GCL_HIP(hipMalloc(&start, sizeof(int32_t) * size_t(ncells + 1)));
  GCL_HIP(hipMalloc(&scan_tmp, sizeof(int64_t) * scan_tmp_elems(ncells)));
  GCL_HIP(hipMalloc(&maxcell, sizeof(int32_t) * 4));
}

void SortBuffers::release() {
  for (void* p : {(void*)key, (void*)slot, (void*)tmp, (void*)perm, (void*)skey, (void*)cnt, (void*)start,
                  (void*)scan_tmp, (void*)maxcell})
    if (p) (void)hipFree(p);
  key = slot = tmp = perm = skey = cnt = start = maxcell = nullptr;
  scan_tmp = nullptr;
}

void PairBuffers::alloc(int64_t n_, int entcap_) {
  release();
  n = n_;
  GCL_HIP(hipMalloc(&cnt, sizeof(int32_t) * size_t(n + 1)));
  GCL_HIP(hipMalloc(&off, sizeof(int64_t) * size_t(n + 1)));
  GCL_HIP(hipMalloc(&scan_tmp, sizeof(int64_t) * scan_tmp_elems(n)));
  entcap = entcap_;
  if (entcap > 0) {
    GCL_HIP(hipMalloc(&ent, sizeof(uint64_t) * size_t(entcap) * size_t(n)));
    GCL_HIP(hipMalloc(&nent, sizeof(int32_t) * size_t(n)));
  }
}

void PairBuffers::set_capacity(int64_t cap, hipStream_t s) {
  if (pairs) GCL_HIP(hipFree(pairs));
  pairs = nullptr;
  capacity = cap;
  if (cap > 0) {
    GCL_HIP(hipMalloc(&pairs, sizeof(int32_t) * 2 * size_t(cap)));
    GCL_HIP(hipMemsetAsync(pairs, 0, sizeof(int32_t) * 2 * size_t(cap), s));
  }
}

void PairBuffers::zero_from(int64_t from, hipStream_t s) {
  if (from < 0) from = 0;
  if (from < capacity)
    GCL_HIP(hipMemsetAsync(pairs + 2 * size_t(from), 0, sizeof(int32_t) * 2 * size_t(capacity - from), s));
}

void PairBuffers::release() {
  for (void* p : {(void*)cnt, (void*)off, (void*)scan_tmp, (void*)pairs, (void*)ent, (void*)nent})
    if (p) (void)hipFree(p);
  cnt = nullptr;
  off = nullptr;
  scan_tmp = nullptr;
  pairs = nullptr;
  ent = nullptr;
  nent = nullptr;
  entcap = 0;
  capacity = 0;
}

void bin_sort(const CellGeom& g, int prec, const void* const* x, int64_t n, bool use_off64, SortBuffers& sb,
              const int* flag, hipStream_t s, const void* const* x_alt, const int* sel, void* const* xs) {
  if (prec == 32) bin_sort_t<float>(g, x, n, use_off64, sb, flag, s, x_alt, sel, xs);
  else bin_sort_t<double>(g, x, n, use_off64, sb, flag, s, x_alt, sel, xs);
}

void gather_positions(int dim, int prec, const void* const* x, void* const* xs, const int32_t* perm, int64_t n,
                      hipStream_t s) {
  XPtrs xp{{x[0], x[1], dim == 3 ? x[2] : x[1]}};
  XPtrs xsp{{xs[0], xs[1], dim == 3 ? xs[2] : xs[1]}};
  if (prec == 32) {
    if (dim == 2) k_gather_pos<float, 2><<<grid_for(n), kBlock, 0, s>>>(xp, xsp, perm, n);
    else k_gather_pos<float, 3><<<grid_for(n), kBlock, 0, s>>>(xp, xsp, perm, n);
  } else {
    if (dim == 2) k_gather_pos<double, 2><<<grid_for(n), kBlock, 0, s>>>(xp, xsp, perm, n);
    else k_gather_pos<double, 3><<<grid_for(n), kBlock, 0, s>>>(xp, xsp, perm, n);
  }
  GCL_HIP(hipGetLastError());
}

void pair_count(const CellGeom& g, int prec, const void* const* xs, const int32_t* start, const int32_t* skey,
                int64_t n, double dist, PairBuffers& pb, hipStream_t s) {
  XPtrs xsp{{xs[0], xs[1], g.dim == 3 ? xs[2] : xs[1]}};
  const bool store = g.dim == 2 && pb.entcap > 0;
  auto run = [&](auto d2) {
    using T = decltype(d2);
    if (g.dim == 3)
      k_pair_count<T, 3><<<grid_for(n), kBlock, 0, s>>>(xsp, start, skey, n, g, d2, pb.cnt);
    else if (store)
      k_pair_count2<T><<<grid_for(n), kBlock, 0, s>>>(xsp, start, skey, n, g, d2, pb.cnt, pb.ent, pb.nent, pb.entcap);
    else
      k_pair_count<T, 2><<<grid_for(n), kBlock, 0, s>>>(xsp, start, skey, n, g, d2, pb.cnt);
  };
  if (prec == 32) {
    float d = float(dist);
    run(d * d);
  } else {
    run(dist * dist);
  }
  GCL_HIP(hipGetLastError());
  exclusive_scan_i64(pb.cnt, pb.off, n, pb.scan_tmp, nullptr, s);
}

namespace {
template <typename T, int D>
void launch_emit(bool wave, const XPtrs& xsp, const int32_t* start, const int32_t* skey, const int32_t* id,
                 int64_t n, const CellGeom& g, T d2, PairBuffers& pb, hipStream_t s) {
  if (wave)
    k_pair_emit_wave<T, D><<<grid_for(n, kBlock / 64), kBlock, 0, s>>>(xsp, start, skey, id, n, g, d2, pb.off,
                                                                       pb.pairs, pb.capacity);
  else if (D == 2 && pb.entcap > 0)
    k_pair_emit<T, D, true><<<grid_for(n), kBlock, 0, s>>>(xsp, start, skey, id, n, g, d2, pb.off, pb.pairs,
                                                           pb.capacity, pb.ent, pb.nent);
  else
    k_pair_emit<T, D, false><<<grid_for(n), kBlock, 0, s>>>(xsp, start, skey, id, n, g, d2, pb.off, pb.pairs,
                                                            pb.capacity, nullptr, nullptr);
}
}  // namespace

void pair_emit(const CellGeom& g, int prec, const void* const* xs, const int32_t* start, const int32_t* skey,
               const int32_t* id, int64_t n, double dist, int64_t total, PairBuffers& pb, hipStream_t s) {
  if (pb.capacity <= 0 || n <= 0) return;
  bool wave = g.dim == 2 && total >= kEmitWaveMin * n;
  if (const char* e = std::getenv("GCLSPH_PAIR_EMIT")) {
    if (std::strcmp(e, "wave") == 0) wave = true;
    else if (std::strcmp(e, "lane") == 0) wave = false;
  }
  XPtrs xsp{{xs[0], xs[1], g.dim == 3 ? xs[2] : xs[1]}};
  if (prec == 32) {
    float d = float(dist);
    float d2 = d * d;
    if (g.dim == 2) launch_emit<float, 2>(wave, xsp, start, skey, id, n, g, d2, pb, s);
    else launch_emit<float, 3>(wave, xsp, start, skey, id, n, g, d2, pb, s);
  } else {
    double d2 = dist * dist;
    if (g.dim == 2) launch_emit<double, 2>(wave, xsp, start, skey, id, n, g, d2, pb, s);
    else launch_emit<double, 3>(wave, xsp, start, skey, id, n, g, d2, pb, s);
  }
  GCL_HIP(hipGetLastError());
}

}  // namespace gclsph
