// SPH step kernels other than the two pair passes (sph_pass.hip): the rebuild
// (counting sort, reorder, membership store), the corrections after the update
// (DPC, XSPH, CSPM), the Δt / time / rebuild-flag commit, the diagnostics,
// and the host launchers.  Conventions as in sph_pass.hip: one thread per
// particle of the cell-sorted state, gathers over the stored neighbour-list
// membership, FP64 in the reference's operation order without FMA contraction
// (R = double) or FP32 pair terms with FP64 sums (R = float, FP32-split).
#include "sph_device.hpp"

namespace gclsph {

// acc·16 + the four bits (r_u < dist2), candidate 0 the highest: four FP64
// compares into SGPR lane masks, then four add-with-carry steps acc+acc+bit.
// Each mask is read three instructions after it is written (gfx950 needs two
// wait states between a VALU write of an SGPR and its use as a carry-in).
__device__ __forceinline__ uint32_t push_lt4(uint32_t acc, double r0, double r1, double r2, double r3,
                                             double dist2) {
  uint64_t c0, c1, c2, c3, co;
  asm("v_cmp_gt_f64_e64 %1, %6, %7\n\t"
      "v_cmp_gt_f64_e64 %2, %6, %8\n\t"
      "v_cmp_gt_f64_e64 %3, %6, %9\n\t"
      "v_cmp_gt_f64_e64 %4, %6, %10\n\t"
      "v_addc_co_u32_e64 %0, %5, %0, %0, %1\n\t"
      "v_addc_co_u32_e64 %0, %5, %0, %0, %2\n\t"
      "v_addc_co_u32_e64 %0, %5, %0, %0, %3\n\t"
      "v_addc_co_u32_e64 %0, %5, %0, %0, %4"
      : "+v"(acc), "=&s"(c0), "=&s"(c1), "=&s"(c2), "=&s"(c3), "=&s"(co)
      : "s"(dist2), "v"(r0), "v"(r1), "v"(r2), "v"(r3));
  return acc;
}

// Window mask from the pushed bits: len candidates pushed in groups of four
// (candidate t at bit nc-1-t, nc = len rounded up to 4) become bit t, the
// candidates past len and the particle itself cleared.
__device__ __forceinline__ uint32_t window_mask(uint32_t acc, int32_t len, int64_t a, int32_t lo) {
  const int32_t nc = (len + 3) & ~3;
  uint32_t mk = __builtin_bitreverse32(acc << (32 - nc));
  if (len < 32) mk &= (1u << len) - 1u;
  const int64_t self = a - lo;
  if (self >= 0 && self < len) mk &= ~(1u << self);
  return mk;
}

// window_bits over candidate positions staged in LDS (2D: (x, y) pairs): lp
// points at candidate lo.  Same arithmetic, so the same bits.
__device__ __forceinline__ uint32_t window_bits_lds(const double2* lp, double xa0, double xa1, int64_t a,
                                                    int32_t lo, int32_t hi, double dist2) {
#pragma clang fp contract(off)
  const int32_t len = hi - lo;
  uint32_t acc = 0;
  for (int32_t j = 0; j < len; j += 4) {
    double2 c[4];
    double r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = lp[j + u];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double d0 = xa0 - c[u].x;
      const double d1 = xa1 - c[u].y;
      r[u] = d0 * d0 + d1 * d1;
    }
    acc = push_lt4(acc, r[0], r[1], r[2], r[3], dist2);
  }
  return window_mask(acc, len, a, lo);
}

// 3D: the same over positions staged as three arrays (x, y, z).
__device__ __forceinline__ uint32_t window_bits_lds3(const double* lx, const double* ly, const double* lz, double xa0,
                                                     double xa1, double xa2, int64_t a, int32_t lo, int32_t hi,
                                                     double dist2) {
#pragma clang fp contract(off)
  const int32_t len = hi - lo;
  uint32_t acc = 0;
  for (int32_t j = 0; j < len; j += 4) {
    double c0[4], c1[4], c2[4];
    double r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c0[u] = lx[j + u];
      c1[u] = ly[j + u];
      c2[u] = lz[j + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double d0 = xa0 - c0[u];
      const double d1 = xa1 - c1[u];
      const double d2 = xa2 - c2[u];
      double r2 = d0 * d0 + d1 * d1;
      r[u] = r2 + d2 * d2;
    }
    acc = push_lt4(acc, r[0], r[1], r[2], r[3], dist2);
  }
  return window_mask(acc, len, a, lo);
}

// The candidates of a stencil row that can pass the membership test, where
// the rank orders every cell by x (so a row, three consecutive cells of one
// y and z, is in x order): [lo, hi) narrowed by bisection over the staged
// x (lx[b] of candidate b) to the first b with xa - x_b < dist and, from
// there, the first b with xa - x_b <= -dist.  Every candidate outside has
// |fl(xa - x_b)| >= dist, the same FP64 difference the test squares, so its
// r^2 rounds to at least fl(dist^2) = dist2 and the test fails.  A NaN
// coordinate (never a member: its r^2 is NaN) compares false both ways; the
// bisection still only drops candidates whose real x lies beyond an x it
// found outside the range, since the real x are in order.
__device__ __forceinline__ void x_range(const double* lx, double xa, double dist, int32_t& lo, int32_t& hi) {
  int32_t n = hi - lo;
  while (n > 0) {
    const int32_t h = n >> 1;
    if (xa - lx[lo + h] >= dist) {
      lo += h + 1;
      n -= h + 1;
    } else {
      n = h;
    }
  }
  int32_t b = lo;
  n = hi - lo;
  while (n > 0) {
    const int32_t h = n >> 1;
    if (xa - lx[b + h] <= -dist) {
      n = h;
    } else {
      b += h + 1;
      n -= h + 1;
    }
  }
  hi = b;
}

// Row windows as stored: up to 32 candidates from a first member, the member
// bits of the consecutive test windows of one row merged into the pending
// window while they fall inside it (fewer, fuller windows for the pair
// loops).  add() in row order, flush() at the end of the row.
struct RowWindows {
  uint64_t* out;
  int64_t n;
  int cap;
  int& cnt;
  int32_t pb = 0;
  uint32_t pm = 0;
  __device__ __forceinline__ void flush() {
    if (pm != 0) {
      if (cnt < cap) out[int64_t(cnt) * n] = (uint64_t(uint32_t(pb)) << 32) | pm;
      ++cnt;
      pm = 0;
    }
  }
  __device__ __forceinline__ void add(int32_t w0, uint32_t mk) {
    uint64_t rest = mk;
    while (rest != 0) {
      const int t0 = __builtin_ctzll(rest);
      const int32_t b0 = w0 + t0;
      if (pm == 0 || b0 >= pb + 32) {
        flush();
        pb = b0;
      }
      const uint64_t lm = (uint64_t(1) << (pb + 32 - b0)) - 1u;  // 1..32 bits
      pm |= uint32_t(((rest >> t0) & lm) << (b0 - pb));
      rest &= ~(lm << t0);
    }
  }
};

// Candidates staged in LDS: 2D 400 positions per wave (the union of a wave's
// three stencil rows holds ~300 on a uniform lattice), 26 KiB per block, six
// blocks per CU; 3D 1280 per block and plane of three stencil rows, 32 KiB.
constexpr int kNmaskWave = 400;
constexpr int kNmaskPlane = 1280;

// Neighbour-list membership of the last rebuild, stored per particle as
// (mask, base) entries: one window of candidates [base, base + len) with its
// member bits in the low word.  Windows of up to 32 candidates of one
// stencil row, or in 3D FP32-split of up to kWin3 candidates inside one
// stencil cell with the cell's offset coded above the member bits
// (win_code3), which is what the cell frames of its partner records need.
// Each window starts at a member (the first one not covered by the window
// before it), so with the rows in x order a particle's partners fill few
// windows; stored in stencil order.  Entry e of particle a is NE[e * n + a]; NC[a]
// is the entry count, or -1 when the particle needs more than necap entries
// (its partners are then found by scanning the stencil).
//
// k_nmask is the 3D store (the 2D one runs inside k_reorder_nmask2 below):
// the block stages the candidates' positions plane by plane (three stencil
// rows at a time, ~1000 candidates per particle in all) and every lane tests
// there the candidates of each row within dist of it in x (x_range, ~46 % of
// them); a block whose union exceeds the LDS budget reads the whole rows from
// global memory.  The tests are the same expression in every path, so the
// bits are the same.
template <int D, typename R, typename SS>
__global__ __launch_bounds__(256) void k_nmask(SS ss) {
  static_assert(D == 3, "the 2D membership store is part of k_reorder_nmask2");
  GCLSPH_SCENE(ss);
  const StepCtl* ctl = B.ctl;
  if (!ctl->rebuilt) return;
  const int64_t a0 = int64_t(m.y) * kBlock;
  if (a0 >= B.n) return;  // whole block
  const bool active = a < B.n;
  const double dist2 = C.dist2;
  const int p = ctl->parity;  // the rebuild positions are the current ones
  const double* XM[3] = {B.X[p][0], B.X[p][1], B.X[p][D == 3 ? 2 : 1]};
  const int cap = B.necap;
  uint64_t* out = B.NE + a;
  __shared__ double lx[kNmaskPlane + 4], ly[kNmaskPlane + 4], lz[kNmaskPlane + 4];  // +4: window reads run up to 3 past a row's end
  __shared__ RowUnion u;
  double xa0 = 0.0, xa1 = 0.0, xa2 = 0.0;
  int64_t ix = 0, iy = 0, iz = 0;
  if (active) {
    xa0 = XM[0][a];
    xa1 = XM[1][a];
    xa2 = XM[2][a];
    cell_xyz<3>(G, B.skey[a], ix, iy, iz);
  }
  const int64_t gx = G.grid[0], gy = G.grid[1], gz = G.grid[2], nce = G.ncells();
  int cnt = 0;
#pragma unroll
  for (int pl = 0; pl < 3; ++pl) {  // rows 3·pl .. 3·pl+2 lie in the plane dz = pl - 1
    // the cell boundaries of the particle's three rows in this plane, loaded
    // together before the staging: cell ox = j - 1 of row kk is [sb[kk][j],
    // sb[kk][j + 1]) when it lies in the grid (indices clamped for the load)
    const int64_t z = iz + pl - 1;
    int32_t sb[3][4];
#pragma unroll
    for (int kk = 0; kk < 3; ++kk) {
      const int64_t y = iy + kk - 1;
      const int64_t row = gx * (y + gy * z);
      const bool in = active && y >= 0 && y < gy && z >= 0 && z < gz;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        int64_t i = row + ix - 1 + t;
        i = i < 0 ? 0 : (i > nce ? nce : i);
        sb[kk][t] = in ? B.start[i] : 0;
      }
    }
    if (threadIdx.x == 0) row_union(u, B.start, B.skey, G, B.n, a0, pl - 1);
    __syncthreads();
    const int32_t total = u.uo[3];
    const bool staged = total <= kNmaskPlane;  // uniform over the block
    if (staged) {
      for (int32_t i = threadIdx.x; i < total; i += kBlock) {
        const int k = (i >= u.uo[1]) + (i >= u.uo[2]);
        const int32_t b = u.ub[k] + (i - u.uo[k]);
        lx[i] = XM[0][b];
        ly[i] = XM[1][b];
        lz[i] = XM[2][b];
      }
    }
    __syncthreads();
    if (active) {
#pragma unroll
      for (int kk = 0; kk < 3; ++kk) {
        const int64_t y = iy + kk - 1;
        if (y < 0 || y >= gy || z < 0 || z >= gz) continue;
        const int32_t off = u.uo[kk] - u.ub[kk];  // candidate b at l?[b + off]
        auto test = [&](int32_t w0, int32_t hi) {
          return staged ? window_bits_lds3(lx + (w0 + off), ly + (w0 + off), lz + (w0 + off), xa0, xa1, xa2, a, w0, hi,
                                           dist2)
                        : window_bits<D>(XM[0], XM[1], XM[2], xa0, xa1, xa2, a, w0, hi, dist2);
        };
        // the row's candidates in the grid, narrowed to |x_a - x_b| < dist
        // where they are staged, tested in 32-candidate windows (one trip
        // count for every lane of the wave, whatever its x in its cell)
        int32_t rb = ix >= 1 ? sb[kk][0] : sb[kk][1];
        int32_t re = ix + 1 < gx ? sb[kk][3] : sb[kk][2];
        if (staged) x_range(lx + off, xa0, C.dist, rb, re);
        // in 3D FP32-split the member bits are stored as windows of one cell
        // (the cell frames of the records need it), merged across the test
        // windows: from a first member, up to kWin3 candidates of its cell
        // (ox = pj - 1); pm holds the pending window's bits
        int32_t pb = 0;
        int pj = 0;
        uint32_t pm = 0;
        RowWindows rw{out, B.n, cap, cnt};  // FP64: 32-candidate row windows
        auto flush = [&]() {
          if (pm != 0) {
            const uint64_t code = uint64_t(win_code3(pj - 1, kk - 1, pl - 1)) << kWin3;
            if (cnt < cap) out[int64_t(cnt) * B.n] = (uint64_t(uint32_t(pb)) << 32) | code | pm;
            ++cnt;
            pm = 0;
          }
        };
        for (int32_t w0 = rb; w0 < re; w0 += 32) {
          const int32_t hi = re < w0 + 32 ? re : w0 + 32;
          const uint32_t mk = test(w0, hi);
          if constexpr (cell_windows<D, R>()) {
            uint64_t rest = mk;
            while (rest != 0) {
              const int t0 = __builtin_ctzll(rest);
              const int32_t b0 = w0 + t0;
              const int j = (b0 >= sb[kk][1]) + (b0 >= sb[kk][2]);
              if (pm == 0 || j != pj || b0 >= pb + int32_t(kWin3)) {
                flush();
                pb = b0;
                pj = j;
              }
              // this test window's bits from b0 to the end of the pending
              // window or of the cell, whichever comes first (1..kWin3 of them)
              const int32_t ce = sb[kk][j + 1], we = pb + int32_t(kWin3);
              const uint64_t lm = (uint64_t(1) << ((ce < we ? ce : we) - b0)) - 1u;
              pm |= uint32_t(((rest >> t0) & lm) << (b0 - pb));
              rest &= ~(lm << t0);
            }
          } else {
            rw.add(w0, mk);
          }
        }
        if constexpr (cell_windows<D, R>()) flush();
        else rw.flush();
      }
    }
    __syncthreads();  // the next plane overwrites the staged positions
  }
  if (active) B.NC[a] = cnt <= cap ? cnt : -1;
  return;
}

// ------------------------------------------------------------------ kernels
// Runs after the (conditional) rebuild kernels: commits the reorder.
template <typename R, typename SS>
__global__ void k_post_rebuild(SS ss) {
  if (threadIdx.x != 0) return;
  const SceneDev<R>& sd = scene_at<R>(ss, blockIdx.x);
  StepCtl* ctl = sd.B.ctl;
  const int32_t* maxcell = sd.maxcell;
  const int write_xrb = sd.F.write_xrb;
  if (ctl->rebuild) {
    ctl->parity ^= 1;
    ctl->pparity ^= 1;
    ctl->rebuilt = 1;
    ctl->updaten += 1;
    long long mp = *maxcell;
    if (mp > ctl->mppcell) ctl->mppcell = (long long)ceil(double(mp) * 1.05 + 1);  // src/celllist.jl:105-106
    ctl->rebuild = 0;
    ctl->xrb_cur = write_xrb ? 0 : 1;
  } else {
    ctl->rebuilt = 0;
  }
}

// New sorted position s takes the state of old position perm[s] (its
// position is already there: the rank kernel moves the positions, and moving
// the other fields there too measured slower: its stores are scattered where
// the loads here are); returns the particle's position (xs) for a membership
// test that follows.
template <int D, typename R>
__device__ __forceinline__ void reorder_one(const SphBufs<R>& B, const SphConst& C, const CellGeom& G,
                                            const StepFlags& F, const StepCtl* ctl, int64_t s, double* xs) {
  const int p = ctl->parity, q = p ^ 1, pp = ctl->pparity, pq = pp ^ 1;
  const int32_t src = B.perm[s];
  R vs[3];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double xv = B.X[q][k][s];
    const R vv = B.V[p][k][src];
    xs[k] = xv;
    vs[k] = vv;
    B.V[q][k][s] = vv;
    if (F.write_xrb) B.XRB[k][s] = xv;
    if (!F.synced_prev) {  // previous half-step state moves with its particle
      B.XH[k][s] = B.XH2[k][src];
      B.VH[k][s] = B.VH2[k][src];
    }
  }
  const R rv = B.RHO[p][src];
  const int32_t pt = B.PT[pp][src];
  B.RHO[q][s] = rv;
  typename Records<D, R>::N rec{};
  double fr[3];
  particle_frame<D, R>(B, C, G, s, fr);  // s's cell in the new order
#pragma unroll
  for (int k = 0; k < D; ++k) {
    put_x(rec, k, xs[k], fr);
    rec.v[k] = vs[k];
  }
  rec.rho = rv;
  if constexpr (std::is_same<R, double>::value) {
    const double P = eos_d(rv, pt, C);
    B.PN[s] = P;
    rec.P = P;
    rec.pt = pt;
  } else {
    const float P = eos_f(rv, pt, C);
    B.PN[s] = P;
    set_pt(rec, pt, P);
  }
  st_rec(B.PRN, s, rec, B.n);
  if (!F.synced_prev) B.RH[s] = B.RH2[src];
  B.PT[pq][s] = pt;
  B.ID[pq][s] = B.ID[pp][src];
}

template <int D, typename R, typename SS>
__global__ void k_reorder(SS ss) {
  GCLSPH_SCENE(ss);
  const StepCtl* ctl = B.ctl;
  if (ctl->rebuild == 0) return;
  if (a >= B.n) return;
  double xs[3];
  reorder_one<D, R>(B, C, G, F, ctl, a, xs);
}

// Membership store (k_nmask's 2D path) of particle a, from the rebuild
// positions in the new order, which the rank kernel wrote (the other fields
// move in this kernel, which needs no other block's writes).  xa: a's
// rebuild position.  Wave-level only (no block barrier).
template <typename R>
__device__ __forceinline__ void nmask2_new(const SphBufs<R>& B, const SphConst& C, const CellGeom& G,
                                           const StepCtl* ctl, int64_t a, int64_t a0, const double* xa,
                                           double2 (*lds)[kNmaskWave + 4]) {
  const bool active = a < B.n;
  const double dist2 = C.dist2;
  const int q = ctl->parity ^ 1;  // the new order (committed after this kernel)
  const double* XN[2] = {B.X[q][0], B.X[q][1]};
  const int cap = B.necap;
  uint64_t* out = B.NE + a;
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int lane = int(threadIdx.x & 63);
  const int64_t aw = a0 + 64 * w;
  if (aw >= B.n) return;  // whole wave
  RowUnion u;
  row_union(u, B.start, B.skey, G, B.n, aw, 0, 64);
  const int32_t total = u.uo[3];
  int cnt = 0;
  if (total <= kNmaskWave) {  // uniform over the wave
    double2* wl = lds[w];
    const int32_t d0 = u.ub[0] - u.uo[0], d1 = u.ub[1] - u.uo[1], d2 = u.ub[2] - u.uo[2];
    // four slots per lane per trip, their loads in flight together; a slot
    // past the union repeats the last valid one's loads and is not written
    for (int32_t i0 = lane; i0 < total; i0 += 4 * 64) {
      double2 v[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int32_t i = i0 + 64 * t < total ? i0 + 64 * t : total - 1;
        const int32_t b = i + (i >= u.uo[2] ? d2 : (i >= u.uo[1] ? d1 : d0));
        v[t] = make_double2(XN[0][b], XN[1][b]);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (i0 + 64 * t < total) wl[i0 + 64 * t] = v[t];
    }
    // the wave's LDS operations complete in issue order; the fence keeps the
    // compiler from moving the reads below ahead of these writes
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (active) {
      const Rows<2> rows = stencil_rows<2>(G, B.start, B.skey[a]);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double2* lrow = wl + (u.uo[k] - u.ub[k]);  // candidate b at lrow[b]
        // (a row of ~35 candidates: bounding it to |x_a - x_b| < dist by
        // bisection, as in 3D, costs more than the tests it saves)
        RowWindows rw{out, B.n, cap, cnt};
        for (int32_t w0 = rows.b[k]; w0 < rows.e[k]; w0 += 32) {
          const int32_t hi = rows.e[k] < w0 + 32 ? rows.e[k] : w0 + 32;
          rw.add(w0, window_bits_lds(lrow + w0, xa[0], xa[1], a, w0, hi, dist2));
        }
        rw.flush();
      }
    }
  } else if (active) {
    const Rows<2> rows = stencil_rows<2>(G, B.start, B.skey[a]);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      RowWindows rw{out, B.n, cap, cnt};
      for (int32_t w0 = rows.b[k]; w0 < rows.e[k]; w0 += 32) {
        const int32_t hi = rows.e[k] < w0 + 32 ? rows.e[k] : w0 + 32;
        rw.add(w0, window_bits<2>(XN[0], XN[1], XN[1], xa[0], xa[1], 0.0, a, w0, hi, dist2));
      }
      rw.flush();
    }
  }
  if (active) B.NC[a] = cnt <= cap ? cnt : -1;
}

// 2D rebuild tail in one kernel: the reorder of the state and the membership
// store of the same particles, which are independent of each other (one
// launch instead of two; both parts are bound by vector-memory traffic, so
// they add up rather than overlap).  Same tests, same bits as k_nmask.
template <typename R, typename SS>
__global__ __launch_bounds__(256) void k_reorder_nmask2(SS ss) {
  GCLSPH_SCENE(ss);
  const StepCtl* ctl = B.ctl;
  if (ctl->rebuild == 0) return;
  const int64_t a0 = int64_t(m.y) * kBlock;
  if (a0 >= B.n) return;  // whole block
  __shared__ double2 lds[kBlock / 64][kNmaskWave + 4];  // +4: window reads run up to 3 past a row's end
  double xa[3] = {0.0, 0.0, 0.0};
  if (a < B.n) reorder_one<2, R>(B, C, G, F, ctl, a, xa);
  nmask2_new<R>(B, C, G, ctl, a, a0, xa, lds);
}

// ---- corrections after the update (src/sphproblem.jl:303-322) --------------
// dpcreg! (src/gpukernels_sph.jl:1500-1553): ∑Δv_dpc into DV; P = EOS(ρ^{n+1}).
template <int D, typename R, typename SS>
__global__ __launch_bounds__(256) void k_dpc_sum(SS ss) {
#pragma clang fp contract(off)
  GCLSPH_SCENE(ss);
  if (!F.dpc_on) return;
  const StepCtl* ctl = B.ctl;
  const int p = ctl->parity, q = p ^ 1, pp = ctl->pparity, rebuilt = ctl->rebuilt;
  const double dt = ctl->dt;
  if (a >= B.n) return;
  const double* X[3] = {B.X[q][0], B.X[q][1], B.X[q][D == 3 ? 2 : 1]};
  const R* V[3] = {B.V[q][0], B.V[q][1], B.V[q][D == 3 ? 2 : 1]};
  const R* RHO = B.RHO[q];
  const int32_t* PT = B.PT[pp];
  const double* XM[3];
  membership_ptrs<D>(B, F, p, rebuilt, XM);
  const int32_t pta = PT[a];
  const int nc = B.NC[a];
  double xa[3];
  R va[3];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    xa[d] = X[d][a];
    va[d] = V[d][a];
  }
  // partners from the records, which hold the updated state (pass 2)
  using Nb = typename Records<D, R>::N;
  auto fetch = [&](const int32_t b) { return ld_rec<Nb>(B.PRN, b, B.n); };
  const R Pa = B.PN[a];
  if (F.want_out) B.PBUF[a] = Pa;
  if (pta <= 0) {  // every DPC pair needs ptype > 0 on both sides
#pragma unroll
    for (int d = 0; d < D; ++d) B.DV[d][a] = R(0);
    return;
  }
  if constexpr (std::is_same<R, double>::value) {
    const double rhoa = RHO[a];
    double sum[3] = {0.0, 0.0, 0.0};
    for_each_partner_fetch<true, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& n) {
      if (n.pt <= 0) return;
      double dx[3], dv[3];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        dx[d] = xa[d] - n.x[d];
        dv[d] = va[d] - n.v[d];
      }
      const double rhob = n.rho;
      const double r2 = dotd<D>(dx, dx);
      const double r = sqrt(r2);
      const double u = r * C.dpc_l0inv;
      const double w = W_d(u, C.dpc_wnorm);
      const double chi = sqrt(w * C.dpc_whinv);
      const double kk = u < C.dpc_lim ? 1.0 : chi;
      double lo = C.dpc_lambda * fabs(Pa + n.P), cap = C.dpc_lambda * C.dpc_pmax;
      double pt = lo < cap ? lo : cap;
      pt = pt > C.dpc_pmin ? pt : C.dpc_pmin;
      const double Pb_dpc = chi * pt;
      const double vr = dotd<D>(dv, dx);
      if (vr < 0) {
        double vrdr = vr / (r2 + C.dpc_eta2);
#pragma unroll
        for (int d = 0; d < D; ++d) sum[d] += kk * (-vrdr * dx[d]);
      } else {
        double tvar = 2 * dt * Pb_dpc / ((rhoa + rhob) * (r2 + C.dpc_eta2));
#pragma unroll
        for (int d = 0; d < D; ++d) sum[d] += tvar * dx[d];
      }
    });
#pragma unroll
    for (int d = 0; d < D; ++d) B.DV[d][a] = sum[d];
  } else {
#pragma clang fp contract(fast)
    const float rhoa = C.f_rho0 + RHO[a];
    const float fdt = float(dt);
    Sum<float> sum[3];
    const FrameK3 kt = particle_k<D, R>(B, C, G, a, xa);
    for_each_partner_fetch<true, D, R, true>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& n,
                                                                         const auto&... win) {
      if (rec_pt(n) <= 0) return;
      float dx[3], dv[3];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        dx[d] = sep_x(xa, n, d, win...);
        dv[d] = va[d] - n.v[d];
      }
      const float rhob = C.f_rho0 + n.rho;
      const float r2 = dotf<D>(dx, dx);
      const float u = __builtin_amdgcn_sqrtf(r2) * C.f_dpc_l0inv;
      const float w = W_f(u, C.f_dpc_wnorm);
      const float chi = __builtin_amdgcn_sqrtf(w * C.f_dpc_whinv);
      const float kk = u < C.f_dpc_lim ? 1.0f : chi;
      const float lo = C.f_dpc_lambda * fabsf(Pa + rec_P(n, C)), cap = C.f_dpc_lambda * C.f_dpc_pmax;
      float pt = lo < cap ? lo : cap;
      pt = pt > C.f_dpc_pmin ? pt : C.f_dpc_pmin;
      const float Pb_dpc = chi * pt;
      const float vr = dotf<D>(dv, dx);
      const float den = __builtin_amdgcn_rcpf(r2 + C.f_dpc_eta2);
      if (vr < 0.0f) {
        const float vrdr = vr * den;
#pragma unroll
        for (int d = 0; d < D; ++d) sum[d].add(kk * (-vrdr * dx[d]));
      } else {
        const float tvar = 2.0f * fdt * Pb_dpc * __builtin_amdgcn_rcpf(rhoa + rhob) * den;
#pragma unroll
        for (int d = 0; d < D; ++d) sum[d].add(tvar * dx[d]);
      }
    }, &kt);
#pragma unroll
    for (int d = 0; d < D; ++d) B.DV[d][a] = sum[d].get();
  }
}

// update_dpcreg! (src/gpukernels_sph.jl:1617-1632)
template <int D, typename R, typename SS>
__global__ void k_dpc_apply(SS ss) {
#pragma clang fp contract(off)
  GCLSPH_SCENE(ss);
  if (!F.dpc_on) return;
  const StepCtl* ctl = B.ctl;
  const int q = ctl->parity ^ 1, pp = ctl->pparity;
  const double dt = ctl->dt;
  if (a >= B.n || B.PT[pp][a] <= 1) return;
  double fr[3];
  particle_frame<D, R>(B, C, G, a, fr);
  using N = typename Records<D, R>::N;
  N rec = ld_rec<N>(B.PRN, a, B.n);
  const int32_t pt = B.PT[pp][a];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    R dv = B.DV[d][a];
    const R nv = B.V[q][d][a] + dv;
    const double nx = B.X[q][d][a] + double(dv) * dt;
    B.V[q][d][a] = nv;
    B.X[q][d][a] = nx;
    rec.v[d] = nv;
    put_x(rec, d, nx, fr);
  }
  if constexpr (!std::is_same<R, double>::value) set_pt(rec, pt, B.PN[a]);  // the 2D type class sits in y
  st_rec(B.PRN, a, rec, B.n);
}

// Partner records of the XSPH pass in the Records::N layout: xⁿ (the
// positions XSPH weighs with), v and ρ after the update and the DPC, ptype.
// They go over the half-step records, which pass 2 was the last to read.
template <int D, typename R, typename SS>
__global__ void k_xsph_rec(SS ss) {
  GCLSPH_SCENE(ss);
  if (!F.xsph_on) return;
  const StepCtl* ctl = B.ctl;
  const int p = ctl->parity, q = p ^ 1, pp = ctl->pparity;
  if (a >= B.n) return;
  typename Records<D, R>::N rec{};
  double fr[3];
  particle_frame<D, R>(B, C, G, a, fr);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    put_x(rec, d, B.X[p][d][a], fr);
    rec.v[d] = B.V[q][d][a];
  }
  rec.rho = B.RHO[q][a];
  if constexpr (std::is_same<R, double>::value) rec.pt = B.PT[pp][a];
  else set_pt(rec, B.PT[pp][a], B.PN[a]);
  st_rec(B.PRH, a, rec, B.n);
}

// xsphcorr! + update_xsphcorr! (src/gpukernels_sph.jl:1757-1838): W at xⁿ.
template <int D, typename R, typename SS>
__global__ __launch_bounds__(256) void k_xsph(SS ss) {
#pragma clang fp contract(off)
  GCLSPH_SCENE(ss);
  if (!F.xsph_on) return;
  const StepCtl* ctl = B.ctl;
  const int p = ctl->parity, q = p ^ 1, pp = ctl->pparity, rebuilt = ctl->rebuilt;
  const double dt = ctl->dt;
  if (a >= B.n) return;
  const int32_t* PT = B.PT[pp];
  const int32_t pta = PT[a];
  if (pta <= 1) return;
  const double* XN[3] = {B.X[p][0], B.X[p][1], B.X[p][D == 3 ? 2 : 1]};
  const R* V[3] = {B.V[q][0], B.V[q][1], B.V[q][D == 3 ? 2 : 1]};
  const R* RHO = B.RHO[q];
  const double* XM[3];
  membership_ptrs<D>(B, F, p, rebuilt, XM);
  const int nc = B.NC[a];
  double xa[3];
  R va[3];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    xa[d] = XN[d][a];
    va[d] = V[d][a];
  }
  // xⁿ, v, ρ and ptype of a partner from the records of k_xsph_rec
  using Nb = typename Records<D, R>::N;
  auto fetch = [&](const int32_t b) { return ld_rec<Nb>(B.PRH, b, B.n); };
  double out[3];
  if constexpr (std::is_same<R, double>::value) {
    const double rhoa = RHO[a];
    double sum[3] = {0.0, 0.0, 0.0};
    for_each_partner_fetch<true, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& n) {
      if (n.pt <= 1) return;
      double dx[3];
#pragma unroll
      for (int d = 0; d < D; ++d) dx[d] = xa[d] - n.x[d];
      const double r = sqrt(dotd<D>(dx, dx));
      const double w = W_d(r * C.Hinv, C.wnorm);
      const double xs = C.xsph_c * w / (rhoa + n.rho);
#pragma unroll
      for (int d = 0; d < D; ++d) sum[d] += (n.v[d] - va[d]) * xs;
    });
#pragma unroll
    for (int d = 0; d < D; ++d) out[d] = sum[d];
  } else {
#pragma clang fp contract(fast)
    const float rhoa = C.f_rho0 + RHO[a];
    Sum<float> sum[3];
    const FrameK3 kt = particle_k<D, R>(B, C, G, a, xa);
    for_each_partner_fetch<true, D, R, true>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& n,
                                                                         const auto&... win) {
      if (rec_pt(n) <= 1) return;
      float dx[3];
#pragma unroll
      for (int d = 0; d < D; ++d) dx[d] = sep_x(xa, n, d, win...);
      const float r = __builtin_amdgcn_sqrtf(dotf<D>(dx, dx));
      const float w = W_f(r * C.f_Hinv, C.f_wnorm);
      const float xs = C.f_xsph_c * w * __builtin_amdgcn_rcpf(rhoa + (C.f_rho0 + n.rho));
#pragma unroll
      for (int d = 0; d < D; ++d) sum[d].add((n.v[d] - va[d]) * xs);
    }, &kt);
#pragma unroll
    for (int d = 0; d < D; ++d) out[d] = double(sum[d].get());
  }
  double fr[3];
  particle_frame<D, R>(B, C, G, a, fr);
  using N = typename Records<D, R>::N;
  N rec = ld_rec<N>(B.PRN, a, B.n);
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double nx = B.X[q][d][a] + out[d] * dt;
    B.X[q][d][a] = nx;
    put_x(rec, d, nx, fr);
  }
  if constexpr (!std::is_same<R, double>::value) set_pt(rec, pta, B.PN[a]);  // the 2D type class sits in y
  st_rec(B.PRN, a, rec, B.n);
}

// cspmcorr! (src/gpukernels_sph.jl:1693-1719): ρ = Σ m0 W / Σ m0 W / ρ_b, W at xⁿ.
template <int D, typename R, typename SS>
__global__ __launch_bounds__(256) void k_cspm_sum(SS ss) {
#pragma clang fp contract(off)
  GCLSPH_SCENE(ss);
  if (!F.cspm_fire) return;
  const StepCtl* ctl = B.ctl;
  const int p = ctl->parity, q = p ^ 1, pp = ctl->pparity, rebuilt = ctl->rebuilt;
  if (a >= B.n) return;
  const int32_t* PT = B.PT[pp];
  const int32_t pta = PT[a];
  const double* XN[3] = {B.X[p][0], B.X[p][1], B.X[p][D == 3 ? 2 : 1]};
  const R* RHO = B.RHO[q];
  const double* XM[3];
  membership_ptrs<D>(B, F, p, rebuilt, XM);
  double s1 = 0.0, s2 = 0.0;
  if (pta > 0) {
    const int nc = B.NC[a];
    double xa[3];
#pragma unroll
    for (int d = 0; d < D; ++d) xa[d] = XN[d][a];
    struct Nb {
      double x[3];
      R rho;
      int32_t pt;
    };
    auto fetch = [&](const int32_t b) {
      const uint32_t o8 = uint32_t(b) * 8u, oR = uint32_t(b) * uint32_t(sizeof(R)), o4 = uint32_t(b) * 4u;
      Nb n;
#pragma unroll
      for (int d = 0; d < D; ++d) n.x[d] = ldo(XN[d], o8);
      n.rho = ldo(RHO, oR);
      n.pt = ldo(PT, o4);
      return n;
    };
    if constexpr (std::is_same<R, double>::value) {
      for_each_partner_fetch<true, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& n) {
        if (n.pt <= 0) return;
        double dx[3];
#pragma unroll
        for (int d = 0; d < D; ++d) dx[d] = xa[d] - n.x[d];
        const double w = W_d(sqrt(dotd<D>(dx, dx)) * C.Hinv, C.wnorm);
        s1 += C.m0 * w;
        s2 += w * C.m0 / n.rho;
      });
    } else {
      // on δρ: with e = δρ_b/ρ0, ρ' - ρ0 = ρ0 Σ m0W e/(1+e) / Σ m0W/(1+e), which
      // is s1/s2 - ρ0 without the cancellation against ρ0
#pragma clang fp contract(fast)
      for_each_partner_fetch<true, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& n) {
        if (n.pt <= 0) return;
        float dx[3];
#pragma unroll
        for (int d = 0; d < D; ++d) dx[d] = float(xa[d] - n.x[d]);
        const float w = W_f(__builtin_amdgcn_sqrtf(dotf<D>(dx, dx)) * C.f_Hinv, C.f_wnorm);
        const float e = n.rho * (1.0f / C.f_rho0);
        const float mw = C.f_m0 * w * __builtin_amdgcn_rcpf(1.0f + e);
        s1 += double(mw * e);
        s2 += double(mw);
      });
      s1 *= C.rho0;
    }
  }
  B.TMP[a] = s1 / s2;  // ρ (FP64 path) or ρ - ρ0 (FP32-split); NaN without partners
}

template <int D, typename R, typename SS>
__global__ void k_cspm_apply(SS ss) {
  GCLSPH_SCENE(ss);
  if (!F.cspm_fire) return;
  const StepCtl* ctl = B.ctl;
  const int q = ctl->parity ^ 1, pp = ctl->pparity;
  if (a >= B.n) return;
  double nr = B.TMP[a];
  if (nr != nr) return;
  const int32_t pt = B.PT[pp][a];
  if constexpr (std::is_same<R, double>::value) {
    B.RHO[q][a] = nr;
    const double P = eos_d(nr, pt, C);
    B.PN[a] = P;
    using N = typename Records<D, R>::N;
    N rec = ld_rec<N>(B.PRN, a, B.n);
    rec.rho = nr;
    rec.P = P;
    st_rec(B.PRN, a, rec, B.n);
  } else {
    const float d = float(nr);
    B.RHO[q][a] = d;
    const float P = eos_f(d, pt, C);
    B.PN[a] = P;
    using N = typename Records<D, R>::N;
    N rec = ld_rec<N>(B.PRN, a, B.n);
    rec.rho = d;
    set_pt(rec, pt, P);
    st_rec(B.PRN, a, rec, B.n);
  }
}

// PN = EOS(current density) for every particle (after construction).
template <typename R, typename SS>
__global__ void k_init_pn(SS ss) {
  GCLSPH_SCENE(ss);
  const StepCtl* ctl = B.ctl;
  const int p = ctl->parity, pp = ctl->pparity;
  if (a >= B.n) return;
  const R rv = B.RHO[p][a];
  const int32_t pt = B.PT[pp][a];
  if constexpr (std::is_same<R, double>::value) B.PN[a] = eos_d(rv, pt, C);
  else B.PN[a] = eos_f(rv, pt, C);
}

// Δt viscous rate over the corrected state (src/gpukernels_sph.jl:1325-1334).
template <int D, typename R, typename SS>
__global__ void k_dt_visc(SS ss) {
  GCLSPH_SCENE(ss);
  if (!(F.ts && !F.visc_in_pass2)) return;  // the whole block belongs to one problem
  StepCtl* ctl = B.ctl;
  const int q = ctl->parity ^ 1;
  const bool active = a < B.n;
  double v = 0.0;
  if (active) {
    double vn[3], xn[3];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      vn[d] = double(B.V[q][d][a]);
      xn[d] = B.X[q][d][a];
    }
    v = fabs(C.h * dotd<D>(vn, xn) / (dotd<D>(xn, xn) + C.dt_eta2));
  }
  block_max(v, active, B.RED + 2 * kRedSlots, m.y);
}

__device__ __forceinline__ double jl_min(double a, double b) { return (a != a || b != b) ? a + b : (a < b ? a : b); }
__device__ __forceinline__ double jl_max(double a, double b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }

// End of step: commits the new state, rebuild flag, time and Δt
// (src/sphproblem.jl:325-335, src/gpukernels_sph.jl:1293-1323).
template <typename R, typename SS>
__global__ void k_finalize(SS ss) {
  // one wave per problem: lane l reads slot l of each reduction and clears it
  // for the next step
  const SceneDev<R>& sd = scene_at<R>(ss, blockIdx.x);
  const StepFlags F = step_flags(sd);
  const SphConst& C = sd.C;
  StepCtl* ctl = sd.B.ctl;
  double* red = sd.B.RED;
  const int ts = F.ts, corr_on = F.corr_on;
  const int l = threadIdx.x;
  double tot[3];
  for (int k = 0; k < 3; ++k) {
    const double v = red[k * kRedSlots + l];
    red[k * kRedSlots + l] = 0.0;
    tot[k] = wave_max(v);
  }
  if (l != 0) return;
  double cdx = tot[0];
  if (cdx > 0.9 * C.nui) ctl->rebuild = 1;
  ctl->maxcdx_out = jl_max(ctl->maxcdx_out, cdx);
  ctl->etime += ctl->dt;
  if (ts) {
    double amax = tot[1];
    double vmax = tot[2];
    double dt1 = sqrt(C.h / (3 * amax));
    double dt2 = C.h / (C.c0 + vmax);
    double dt = C.CFL * jl_min(dt1, dt2);
    dt = jl_min(jl_max(dt, C.tl0), C.tl1);
    ctl->dt = dt;
  }
  ctl->red_cdx = 0;
  ctl->red_acc = 0;
  ctl->red_visc = 0;
  ctl->parity ^= 1;
  ctl->synced = corr_on ? 0 : 1;
  ctl->xrb_cur = 0;
  ctl->iter += 1;
}

// sph∑W! + sph∑∇W! over the current positions (src/gpukernels_sph.jl:139-156, :332-359).
template <int D, typename R, typename SS>
__global__ __launch_bounds__(256) void k_diag(SS ss, int xm_mode, int use_masks) {
#pragma clang fp contract(off)
  GCLSPH_SCENE(ss);
  const StepCtl* ctl = B.ctl;
  const int p = ctl->parity, pp = ctl->pparity;
  if (a >= B.n) return;
  const double* X[3] = {B.X[p][0], B.X[p][1], B.X[p][D == 3 ? 2 : 1]};
  // membership positions: 0 = stored rebuild positions, 1 = the step's start
  // positions (other buffer), 2 = the current positions
  const int mp = xm_mode == 1 ? (p ^ 1) : p;
  const double* XM[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) XM[k] = xm_mode == 0 ? B.XRB[k < D ? k : 1] : B.X[mp][k < D ? k : 1];
  const int32_t* PT = B.PT[pp];
  const bool pa = PT[a] > 0;
  double sw = 0.0, sg[3] = {0.0, 0.0, 0.0};
  const int nc = use_masks ? B.NC[a] : -1;
  for_each_partner<D>(B, a, nc, G, XM, C.dist2, [&](const int32_t b) {
      if (pa != (PT[b] > 0)) return;
      double dx[3];
#pragma unroll
      for (int d = 0; d < D; ++d) dx[d] = X[d][a] - X[d][b];
      const double r = sqrt(dotd<D>(dx, dx));
      const double u = r * C.Hinv;
      sw += W_d(u, C.wnorm);
      const double dwk_r = dW_d(u, C.dwnorm) / r;
#pragma unroll
      for (int d = 0; d < D; ++d) sg[d] += dx[d] * dwk_r;
  });
  B.SUMW[a] = sw;
#pragma unroll
  for (int d = 0; d < D; ++d) B.SUMGW[d][a] = sg[d];
}

// ------------------------------------------------- rebuild: counting sort
// The problem's cell-sorted order is rebuilt by a counting sort over its cells
// (src/gpukernels_cls.jl:29-101 bins particles with atomics into a dense table):
// histogram with per-particle arrival slots, exclusive scan into cell starts,
// scatter, and a per-cell rank by input index (the previous sorted position),
// which makes the order deterministic.  Every kernel returns at once for a
// problem whose control block does not ask for a rebuild.

// cnt[0..ncells] = 0 and maxcell = 0 (cell tiles).
template <typename R, typename SS>
__global__ void kb_clear_cells(SS ss) {
  int2 m;
  const SceneDev<R>& sd = tile_scene<R>(ss, m);
  if (sd.B.ctl->rebuild == 0) return;
  const int64_t base = int64_t(m.y) * kTile;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + int64_t(k) * kScanThreads + threadIdx.x;
    if (i <= sd.ncells) sd.cnt[i] = 0;
  }
  if (m.y == 0 && threadIdx.x == 0) *sd.maxcell = 0;
}

// Cell of every particle (current positions) and its arrival slot.
template <int D, typename R, typename SS>
__global__ void kb_cellmap(SS ss) {
  GCLSPH_SCENE(ss);
  const StepCtl* ctl = B.ctl;
  if (ctl->rebuild == 0 || a >= B.n) return;
  const int p = ctl->parity;
  int32_t c[3] = {1, 1, 1};
#pragma unroll
  for (int k = 0; k < D; ++k) c[k] = cell_coord<double>(B.X[p][k][a], G, k, 0);
  const int64_t lin = int64_t(c[0] - 1) + G.grid[0] * (int64_t(c[1] - 1) + G.grid[1] * int64_t(c[2] - 1));
  const int32_t key = int32_t(lin);
  sd.key[a] = key;
  sd.slot[a] = wave_aggregated_slot(sd.cnt, key);
}

// Tile sums (and tile maxima, stored after the sums) of the cell counts.
template <typename R, typename SS>
__global__ void kb_tile_sums(SS ss) {
  int2 m;
  const SceneDev<R>& sd = tile_scene<R>(ss, m);
  if (sd.B.ctl->rebuild == 0) return;
  const int64_t n = sd.ncells, nt = (n + kTile - 1) / kTile;
  if (m.y >= nt) return;
  const int64_t base = int64_t(m.y) * kTile;
  int64_t s = 0;
  int32_t mx = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + int64_t(k) * kScanThreads + threadIdx.x;
    if (i < n) {
      const int32_t v = sd.cnt[i];
      s += v;
      mx = v > mx ? v : mx;
    }
  }
  int64_t tot;
  block_exclusive_scan<int64_t>(s, &tot);
  if (threadIdx.x == 0) sd.scan_tmp[m.y] = tot;
  __shared__ int32_t wmx[kScanThreads / 64];
  for (int o = 32; o > 0; o >>= 1) {
    const int32_t t = __shfl_xor(mx, o, 64);
    mx = t > mx ? t : mx;
  }
  if ((threadIdx.x & 63) == 0) wmx[threadIdx.x >> 6] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t v = wmx[0];
    for (int w = 1; w < kScanThreads / 64; ++w) v = wmx[w] > v ? wmx[w] : v;
    reinterpret_cast<int32_t*>(sd.scan_tmp + nt + 2)[m.y] = v;
  }
}

// One block per problem: exclusive scan of the tile sums in place (total at
// [nt]) and the maximum cell count.
template <typename R, typename SS>
__global__ void kb_scan_tile_sums(SS ss) {
  const SceneDev<R>& sd = scene_at<R>(ss, blockIdx.x);
  if (sd.B.ctl->rebuild == 0) return;
  int64_t* tsum = sd.scan_tmp;
  const int64_t nt = (sd.ncells + kTile - 1) / kTile;
  const int32_t* tm = reinterpret_cast<const int32_t*>(tsum + nt + 2);
  int32_t mx = 0;
  for (int64_t i = threadIdx.x; i < nt; i += kScanThreads) mx = tm[i] > mx ? tm[i] : mx;
  for (int o = 32; o > 0; o >>= 1) {
    const int32_t t = __shfl_xor(mx, o, 64);
    mx = t > mx ? t : mx;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(sd.maxcell, mx);
  const int64_t per = (nt + kScanThreads - 1) / kScanThreads;
  const int64_t b = int64_t(threadIdx.x) * per, e = b + per < nt ? b + per : nt;
  int64_t s = 0;
  for (int64_t i = b; i < e; ++i) s += tsum[i];
  int64_t tot;
  int64_t pre = block_exclusive_scan<int64_t>(s, &tot);
  for (int64_t i = b; i < e; ++i) {
    const int64_t v = tsum[i];
    tsum[i] = pre;
    pre += v;
  }
  if (threadIdx.x == 0) tsum[nt] = tot;
}

// Cell starts: tile scans plus the tile prefixes; start[ncells] = n.
template <typename R, typename SS>
__global__ void kb_tile_scan(SS ss) {
  int2 m;
  const SceneDev<R>& sd = tile_scene<R>(ss, m);
  if (sd.B.ctl->rebuild == 0) return;
  const int64_t n = sd.ncells, nt = (n + kTile - 1) / kTile;
  if (m.y >= nt) return;
  const int64_t base = int64_t(m.y) * kTile + int64_t(threadIdx.x) * kScanItems;
  static_assert(kScanItems == 4, "one 16-byte word of counts per thread");
  const bool whole = base + kScanItems <= n;  // the last tile's tail goes element by element
  int32_t v[kScanItems];
  if (whole) {
    const int4 w = *reinterpret_cast<const int4*>(sd.cnt + base);
    v[0] = w.x, v[1] = w.y, v[2] = w.z, v[3] = w.w;
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) v[k] = base + k < n ? sd.cnt[base + k] : 0;
  }
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) s += v[k];
  int64_t tot;
  int64_t pre = block_exclusive_scan<int64_t>(s, &tot) + sd.scan_tmp[m.y];
  int32_t o[kScanItems];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    o[k] = int32_t(pre);
    pre += v[k];
  }
  if (whole) {
    *reinterpret_cast<int4*>(sd.start + base) = make_int4(o[0], o[1], o[2], o[3]);
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
      if (base + k < n) sd.start[base + k] = o[k];
  }
  if (m.y == 0 && threadIdx.x == 0) sd.start[n] = int32_t(sd.scan_tmp[nt]);
}

// Sort key of an x coordinate: its bit pattern made unsigned-monotone (order
// of the key = order of the value for every non-NaN x; -0 before +0).
__device__ __forceinline__ uint64_t x_order_key(double x) {
  const uint64_t u = uint64_t(__double_as_longlong(x));
  return (u >> 63) ? ~u : (u | (uint64_t(1) << 63));
}

template <int D, typename R, typename SS>
__global__ void kb_scatter(SS ss) {
  GCLSPH_SCENE(ss);
  if (B.ctl->rebuild == 0 || a >= B.n) return;
  sd.tmp[sd.start[sd.key[a]] + sd.slot[a]] = int32_t(a);
}

// Rank inside the cell by (x, input index): the final order is (cell, x,
// input), so the candidates of a stencil row (three consecutive cells of one
// y, and z) are in x order (the cell of a coordinate is a non-decreasing
// function of it): a particle's partners in a row are then near each other,
// so they fill fewer stored windows, and in 3D the membership store bounds
// each particle's candidates in a row to |x_a - x_b| < dist by bisection.
// The order is the same for every run (ties in x by input index).  The
// particle's position (loaded first: its x is the key) moves to its new place
// here, so the membership store after it reads the candidates of a stencil
// row as contiguous ranges of the new position arrays instead of through the
// permutation.
template <int D, typename R, typename SS>
__global__ __launch_bounds__(256) void kb_rank(SS ss) {
  GCLSPH_SCENE(ss);
  if (B.ctl->rebuild == 0) return;
  const int64_t a0 = int64_t(m.y) * kBlock;
  if (a0 >= B.n) return;  // whole block
  // the (x key, input) of the block's slots, which hold the whole cell of
  // most of its particles, are shared through LDS; a cell across the block's
  // edge reads its members' x from global memory
  __shared__ uint64_t lk[kBlock];
  __shared__ int32_t li[kBlock];
  const int p = B.ctl->parity;
  const bool active = a < B.n;
  int32_t i = 0;
  double x[3] = {0.0, 0.0, 0.0};
  if (active) {
    i = sd.tmp[a];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = B.X[p][k][i];
  }
  const uint64_t ki = x_order_key(x[0]);
  lk[threadIdx.x] = ki;
  li[threadIdx.x] = i;
  __syncthreads();
  if (!active) return;
  const int32_t c = sd.key[i];
  const int32_t b = sd.start[c], e = sd.start[c + 1];
  const int32_t s0 = int32_t(a0), s1 = int32_t(a0 + kBlock < B.n ? a0 + kBlock : B.n);
  // four compares per trip (indices clamped into the cell, which holds at
  // least this particle; the clamped ones are not counted)
  int32_t r = 0;
  if (b >= s0 && e <= s1) {
    for (int32_t k = b; k < e; k += 4) {
      int32_t v[4];
      uint64_t kv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int32_t j = (k + u < e ? k + u : e - 1) - s0;
        v[u] = li[j];
        kv[u] = lk[j];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) r += (k + u < e) & ((kv[u] < ki) | ((kv[u] == ki) & (v[u] < i)));
    }
  } else {
    const double* X0 = B.X[p][0];
    for (int32_t k = b; k < e; k += 4) {
      int32_t v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = sd.tmp[k + u < e ? k + u : e - 1];
      uint64_t kv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) kv[u] = x_order_key(X0[v[u]]);
#pragma unroll
      for (int u = 0; u < 4; ++u) r += (k + u < e) & ((kv[u] < ki) | ((kv[u] == ki) & (v[u] < i)));
    }
  }
  sd.perm[b + r] = i;
  sd.skey[b + r] = c;
#pragma unroll
  for (int k = 0; k < D; ++k) B.X[p ^ 1][k][b + r] = x[k];
}

// ------------------------------------------------------------ host launchers
template <int D, typename R>
void launch_rebuild(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) {
    using SS = decltype(ss);
    kb_clear_cells<R, SS><<<L.ctiles, kScanThreads, 0, s>>>(ss);
    kb_cellmap<D, R, SS><<<L.pblocks, kBlock, 0, s>>>(ss);
    kb_tile_sums<R, SS><<<L.ctiles, kScanThreads, 0, s>>>(ss);
    kb_scan_tile_sums<R, SS><<<L.count, kScanThreads, 0, s>>>(ss);
    kb_tile_scan<R, SS><<<L.ctiles, kScanThreads, 0, s>>>(ss);
    kb_scatter<D, R, SS><<<L.pblocks, kBlock, 0, s>>>(ss);
    kb_rank<D, R, SS><<<L.pblocks, kBlock, 0, s>>>(ss);
    if constexpr (D == 2) {
      k_reorder_nmask2<R, SS><<<L.pblocks, kBlock, 0, s>>>(ss);
      k_post_rebuild<R, SS><<<L.count, 64, 0, s>>>(ss);
    } else {
      k_reorder<D, R, SS><<<L.pblocks, kBlock, 0, s>>>(ss);
      k_post_rebuild<R, SS><<<L.count, 64, 0, s>>>(ss);
      k_nmask<D, R, SS><<<L.pblocks, kBlock, 0, s>>>(ss);
    }
  });
}
template <int D, typename R>
void launch_init_pn(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) { k_init_pn<R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss); });
}
template <int D, typename R>
void launch_dpc(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) {
    k_dpc_sum<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss);
    k_dpc_apply<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss);
  });
}
template <int D, typename R>
void launch_xsph(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) {
    k_xsph_rec<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss);
    k_xsph<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss);
  });
}
template <int D, typename R>
void launch_cspm(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) {
    k_cspm_sum<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss);
    k_cspm_apply<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss);
  });
}
template <int D, typename R>
void launch_dt_visc(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) { k_dt_visc<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss); });
}
template <int D, typename R>
void launch_finalize(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) { k_finalize<R, decltype(ss)><<<L.count, kRedSlots, 0, s>>>(ss); });
}
template <int D, typename R>
void launch_diag(const LaunchSet& L, int xm_mode, int use_masks, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) {
    k_diag<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss, xm_mode, use_masks);
  });
}

#define GCLSPH_INSTANTIATE(D, R)                                                         \
  template void launch_rebuild<D, R>(const LaunchSet&, hipStream_t);                     \
  template void launch_init_pn<D, R>(const LaunchSet&, hipStream_t);                     \
  template void launch_dpc<D, R>(const LaunchSet&, hipStream_t);                         \
  template void launch_xsph<D, R>(const LaunchSet&, hipStream_t);                        \
  template void launch_cspm<D, R>(const LaunchSet&, hipStream_t);                        \
  template void launch_dt_visc<D, R>(const LaunchSet&, hipStream_t);                     \
  template void launch_finalize<D, R>(const LaunchSet&, hipStream_t);                    \
  template void launch_diag<D, R>(const LaunchSet&, int, int, hipStream_t);

GCLSPH_INSTANTIATE(2, double)
GCLSPH_INSTANTIATE(3, double)
GCLSPH_INSTANTIATE(2, float)
GCLSPH_INSTANTIATE(3, float)

}  // namespace gclsph
