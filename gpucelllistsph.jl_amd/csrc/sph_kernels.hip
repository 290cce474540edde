// Particle-centric SPH step kernels.
//
// Every pair kernel of the reference (src/gpukernels_sph.jl) runs one thread per
// *pair* and scatters both halves with FP64 atomics.  Here one thread owns one
// particle `a` of the cell-sorted state and gathers over the 3^D cells around
// it; the pairs it visits are exactly the reference's neighbour list (same
// cells, r^2 < dist^2 at the positions of the last rebuild).  For the pair
// (a, b) it evaluates the reference's per-pair expression for a's side — with
// Δx = x_a - x_b, which is the reference's Δx or its exact negation — so the
// per-pair values match the reference's and only the summation order differs.
// Sums are per-thread and deterministic; there are no floating-point atomics.
//
// Pass 1 (tⁿ) fuses sphW!/sph∇W!, ∂ρ∂tDDT!, pressure!, ∂v∂t!, ∂v∂t_av!,
// ∂v∂t_visc!, ∂v∂t_addgrav!, fbmolforce!, ∂v∂tpF! and the three half-step
// updates (src/sphproblem.jl:244-274).  Pass 2 (t^{n+1/2}) fuses the second
// evaluation with symplectic_update! and the Δt / rebuild reductions
// (src/sphproblem.jl:277-335).
//
// R = double: the Float64 path.  Every expression is evaluated in the
// reference's operation order with FMA contraction off, so branch decisions
// (r < r0, cond < 0, clamps) see the same values as the reference would.
// R = float: the FP32-split path (density stored as ρ-ρ0, FP32 pair terms
// summed in FP64, Float64 positions and Δt).
#include <type_traits>

#include "blockscan.hpp"
#include "sph.hpp"

namespace gclsph {

// ---------------------------------------------------------------- math helpers
__device__ __forceinline__ double powi_d(double x, int p) {
#pragma clang fp contract(off)
  // binary powering, same operation sequence as the oracle's powi
  double r = 1.0, b = x;
  bool first = true;
  while (p > 0) {
    if (p & 1) {
      r = first ? b : r * b;
      first = false;
    }
    p >>= 1;
    if (p) b = b * b;
  }
  return r;
}

// powfancy7th (src/gpukernels_sph.jl:499-515) for γ == 7; x^(1/γ) otherwise.
__device__ __forceinline__ double pf7(double x, int gamma, double ginv) {
#pragma clang fp contract(off)
  if (gamma == 7) {
    unsigned long long bits = (unsigned long long)__double_as_longlong(fabs(x));
    unsigned long long tb = 0x36cd000000000000ull + bits / 7ull;
    double t = copysign(__longlong_as_double((long long)tb), x);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
      double t2 = t * t;
      double t3 = t2 * t;
      double t4 = t2 * t2;
      double xot4 = x / t4;
      t = t - t * (t3 - xot4) / (4.0 * t3 + 3.0 * xot4);
    }
    return t;
  }
  return pow(x, ginv);
}

// WendlandC2 (SPHKernels.jl): W = ((1+4u)(1-u)^4) σ H⁻ᴰ, dW = -20u(1-u)^3 σ H⁻⁽ᴰ⁺¹⁾.
__device__ __forceinline__ double W_d(double u, double wnorm) {
#pragma clang fp contract(off)
  if (u < 1.0) {
    double um1 = 1.0 - u;
    um1 = um1 * um1;
    um1 = um1 * um1;
    return ((1.0 + 4.0 * u) * um1) * wnorm;
  }
  return 0.0;
}
__device__ __forceinline__ double dW_d(double u, double dwnorm) {
#pragma clang fp contract(off)
  if (u < 1.0) {
    double um1 = 1.0 - u;
    double um13 = um1 * um1 * um1;
    return -20.0 * u * um13 * dwnorm;
  }
  return 0.0;
}
__device__ __forceinline__ float W_f(float u, float wnorm) {
  if (u < 1.0f) {
    float um1 = 1.0f - u;
    um1 = um1 * um1;
    um1 = um1 * um1;
    return ((1.0f + 4.0f * u) * um1) * wnorm;
  }
  return 0.0f;
}
__device__ __forceinline__ float dW_f(float u, float dwnorm) {
  if (u < 1.0f) {
    float um1 = 1.0f - u;
    return -20.0f * u * (um1 * um1 * um1) * dwnorm;
  }
  return 0.0f;
}

// Tait EOS (src/gpukernels_sph.jl:691-697): 0 if ptype <= 1 and ρ < ρ0.
__device__ __forceinline__ double eos_d(double rho, int32_t pt, const SphConst& C) {
#pragma clang fp contract(off)
  if (pt <= 1 && rho < C.rho0) return 0.0;
  return C.P0 * (powi_d(rho / C.rho0, C.gamma) - 1);
}
// FP32-split EOS on δρ: P0((1+e)^7 - 1) = P0 e (7 + 21e + 35e² + 35e³ + 21e⁴ + 7e⁵ + e⁶).
__device__ __forceinline__ float eos_f(float drho, int32_t pt, const SphConst& C) {
  if (pt <= 1 && drho < 0.0f) return 0.0f;
  float e = drho / C.f_rho0;
  float poly = 7.0f + e * (21.0f + e * (35.0f + e * (35.0f + e * (21.0f + e * (7.0f + e)))));
  return C.f_P0 * (e * poly);
}
// ρ0((1+z)^(1/7) - 1) without cancellation (FP32-split form of
// ρ0*powfancy7th(1+z) - ρ0): binomial series, exact function for |z| > 1e-2.
__device__ __forceinline__ float hydro_drho_f(float z, float rho0) {
  if (fabsf(z) < 1e-2f) {
    float t = z * (1.0f / 7.0f + z * (-3.0f / 49.0f + z * (13.0f / 343.0f + z * (-65.0f / 2401.0f))));
    return rho0 * t;
  }
  return rho0 * expm1f(log1pf(z) * (1.0f / 7.0f));
}

template <typename R>
struct Sum {  // plain sum
  R s = R(0);
  __device__ __forceinline__ void add(R v) { s += v; }
  __device__ __forceinline__ R get() const { return s; }
};
template <>
struct Sum<float> {  // FP32 pair terms accumulated in FP64 (2 VALU per term, exact for ~30 terms)
  double s = 0.0;
  __device__ __forceinline__ void add(float v) { s += double(v); }
  __device__ __forceinline__ float get() const { return float(s); }
};

template <int D>
__device__ __forceinline__ bool member(const double* const* XM, int64_t a, int64_t b, double dist2) {
#pragma clang fp contract(off)
  double d0 = XM[0][a] - XM[0][b];
  double d1 = XM[1][a] - XM[1][b];
  double s = d0 * d0 + d1 * d1;
  if (D == 3) {
    double d2 = XM[2][a] - XM[2][b];
    s = s + d2 * d2;
  }
  return s < dist2;
}

template <int D>
__device__ __forceinline__ double dotd(const double* a, const double* b) {
#pragma clang fp contract(off)
  double s = a[0] * b[0] + a[1] * b[1];
  if (D == 3) s = s + a[2] * b[2];
  return s;
}
template <int D>
__device__ __forceinline__ float dotf(const float* a, const float* b) {
  float s = a[0] * b[0] + a[1] * b[1];
  if (D == 3) s = s + a[2] * b[2];
  return s;
}

// Positions of the pair-list membership test: the current positions when this
// step rebuilt (or when no later step can skip a rebuild), else the stored
// rebuild positions.
template <int D, typename R>
__device__ __forceinline__ void membership_ptrs(const SphBufs<R>& B, const StepFlags& F, int p, int rebuilt,
                                                const double** XM) {
  bool cur = rebuilt || !F.need_xrb;
#pragma unroll
  for (int k = 0; k < 3; ++k) XM[k] = cur ? B.X[p][k < D ? k : 1] : B.XRB[k < D ? k : 1];
}

// Logical number of this block with an XCD-aware order: workgroups are dealt
// to the 8 XCDs round robin (block b runs on XCD b % 8), and renumbering so
// that each XCD walks one contiguous range of logical blocks (cell-sorted
// particles) keeps the neighbourhoods shared by consecutive blocks, and by the
// stencil rows above and below, inside one XCD's L2.
__device__ __forceinline__ unsigned xcd_block() {
  const unsigned G = gridDim.x, b = blockIdx.x;
  const unsigned q = G / 8, r = G % 8, xcd = b % 8, i = b / 8;
  return xcd < r ? xcd * (q + 1) + i : r * (q + 1) + (xcd - r) * q + i;
}

// Where a kernel finds its problem.  One problem: the descriptor is a kernel
// argument (constant memory the compiler may reload at will, which keeps
// register pressure as low as plain arguments).  Several problems: an array of
// descriptors in device memory and maps from blocks to (problem, block of the
// problem) or (problem, cell tile).
template <typename R>
struct OneScene {
  SceneDev<R> d;
};
template <typename R>
struct ManyScenes {
  const SceneDev<R>* __restrict__ S;
  const int2* __restrict__ pmap;
  const int2* __restrict__ cmap;
};
template <typename R>
__device__ __forceinline__ const SceneDev<R>& particle_scene(const OneScene<R>& s, int2& m) {
  m = make_int2(0, int(xcd_block()));
  return s.d;
}
template <typename R>
__device__ __forceinline__ const SceneDev<R>& particle_scene(const ManyScenes<R>& s, int2& m) {
  m = s.pmap[xcd_block()];
  return s.S[m.x];
}
template <typename R>
__device__ __forceinline__ const SceneDev<R>& tile_scene(const OneScene<R>& s, int2& m) {
  m = make_int2(0, int(blockIdx.x));
  return s.d;
}
template <typename R>
__device__ __forceinline__ const SceneDev<R>& tile_scene(const ManyScenes<R>& s, int2& m) {
  m = s.cmap[blockIdx.x];
  return s.S[m.x];
}
template <typename R>
__device__ __forceinline__ const SceneDev<R>& scene_at(const OneScene<R>& s, unsigned) {
  return s.d;
}
template <typename R>
__device__ __forceinline__ const SceneDev<R>& scene_at(const ManyScenes<R>& s, unsigned i) {
  return s.S[i];
}

// Per-step switches of a problem, from its per-call switches and control block.
template <typename R>
__device__ __forceinline__ StepFlags step_flags(const SceneDev<R>& sd) {
  StepFlags F = sd.F;
  const StepCtl* c = sd.B.ctl;
  const int it = c->iter;
  F.synced_prev = c->synced;
  F.cspm_fire = F.cspmn > 0 && it % F.cspmn == 0;
  F.corr_on = F.dpc_on || F.xsph_on || F.cspm_fire;
  F.want_out = it == c->nsteps;
  return F;
}

// Binds the problem of this block: B, C, G, F and the particle index a.
#define GCLSPH_SCENE(ss)                                       \
  int2 m;                                                      \
  const SceneDev<R>& sd = particle_scene<R>(ss, m);            \
  const SphBufs<R>& B = sd.B;                                  \
  const SphConst& C = sd.C;                                    \
  const CellGeom& G = sd.G;                                    \
  const StepFlags F = step_flags(sd);                          \
  const int64_t a = int64_t(m.y) * kBlock + threadIdx.x;       \
  (void)C;                                                     \
  (void)G;                                                     \
  (void)F

// Member bits of one window [lo, hi) (hi - lo <= 32) of a stencil row: bit j is
// set when candidate lo + j is a partner of particle a, i.e. b != a and
// r^2 < dist^2 with the reference's expression (no FMA).  Candidates are read
// four at a time so their loads are in flight together; reads run up to 3 past
// hi, which the position arrays' padding keeps in bounds, and those bits are
// dropped.
template <int D>
__device__ __forceinline__ uint32_t window_bits(const double* const* XM, double xa0, double xa1, double xa2,
                                                int64_t a, int32_t lo, int32_t hi, double dist2) {
#pragma clang fp contract(off)
  const double* p0 = XM[0] + lo;
  const double* p1 = XM[1] + lo;
  const double* p2 = XM[2] + lo;
  const int32_t len = hi - lo;
  uint32_t mk = 0;
  for (int32_t j = 0; j < len; j += 4) {
    double c0[4], c1[4], c2[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      c0[u] = p0[j + u];
      c1[u] = p1[j + u];
      if (D == 3) c2[u] = p2[j + u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double d0 = xa0 - c0[u];
      const double d1 = xa1 - c1[u];
      double r2 = d0 * d0 + d1 * d1;
      if (D == 3) {
        const double d2 = xa2 - c2[u];
        r2 = r2 + d2 * d2;
      }
      mk |= uint32_t(r2 < dist2) << (j + u);
    }
  }
  if (len < 32) mk &= (1u << len) - 1u;
  const int64_t self = a - lo;
  if (self >= 0 && self < len) mk &= ~(1u << self);
  return mk;
}

// Sorted range [b, e) of stencil row k (dy = k%3 - 1, dz = k/3 - 1) of cell c;
// empty when the row lies outside the grid.
template <int D>
__device__ __forceinline__ void row_range(const CellGeom& g, const int32_t* __restrict__ start, int32_t c, int k,
                                          int32_t& b, int32_t& e) {
  const int64_t gx = g.grid[0], gy = g.grid[1];
  const int64_t ix = c % gx, rest = c / gx;
  const int64_t iy = rest % gy, iz = rest / gy;
  const int64_t xl = ix > 0 ? ix - 1 : 0, xh = ix + 1 < gx ? ix + 1 : gx - 1;
  const int64_t y = iy + (k % 3) - 1;
  const int64_t z = D == 3 ? iz + (k / 3) - 1 : 0;
  const bool in = y >= 0 && y < gy && (D == 2 || (z >= 0 && z < g.grid[2]));
  const int64_t row = gx * (y + gy * z);
  b = in ? start[row + xl] : 0;
  e = in ? start[row + xh + 1] : 0;
}

// window_bits over candidate positions staged in LDS (2D: (x, y) pairs): lp
// points at candidate lo.  Same arithmetic, so the same bits.
__device__ __forceinline__ uint32_t window_bits_lds(const double2* lp, double xa0, double xa1, int64_t a,
                                                    int32_t lo, int32_t hi, double dist2) {
#pragma clang fp contract(off)
  const int32_t len = hi - lo;
  uint32_t mk = 0;
  for (int32_t j = 0; j < len; j += 4) {
    double2 c[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) c[u] = lp[j + u];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const double d0 = xa0 - c[u].x;
      const double d1 = xa1 - c[u].y;
      const double r2 = d0 * d0 + d1 * d1;
      mk |= uint32_t(r2 < dist2) << (j + u);
    }
  }
  if (len < 32) mk &= (1u << len) - 1u;
  const int64_t self = a - lo;
  if (self >= 0 && self < len) mk &= ~(1u << self);
  return mk;
}

// Candidates staged in LDS per block (2D): 2048 positions, 32 KiB.
constexpr int kNmaskLds = 2048;

// Neighbour-list membership of the last rebuild, stored per particle as
// (mask, base) entries: one 32-candidate window [base, base+32) of one stencil
// row with its member bits in the low word.  Only windows with members are
// stored, in row order.  Entry e of particle a is NE[e * n + a]; NC[a] is the
// entry count, or -1 when the particle needs more than necap entries (its
// partners are then found by scanning the stencil).
//
// 2D: the block's 256 consecutive particles lie in one contiguous range of
// cells, so the union of their stencil rows is, per row offset, one contiguous
// range of sorted particles.  The block stages those positions in LDS once
// (coalesced) and every lane tests its candidates there, instead of each lane
// fetching each of its ~100 candidates through the vector L1.  A block whose
// union exceeds the LDS budget reads the candidates from global memory.
template <int D, typename R, typename SS>
__global__ __launch_bounds__(256) void k_nmask(SS ss) {
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
  if constexpr (D == 2) {
    __shared__ double2 lds[kNmaskLds + 4];  // +4: window reads run up to 3 past a row's end
    __shared__ int32_t ub[3], uo[4];
    if (threadIdx.x == 0) {
      const int64_t a1 = (a0 + kBlock < B.n ? a0 + kBlock : B.n) - 1;
      const int64_t c0 = B.skey[a0], c1 = B.skey[a1];
      const int64_t gx = G.grid[0], nc = G.ncells();
      int32_t o = 0;
      for (int k = 0; k < 3; ++k) {
        int64_t lo = c0 + (k - 1) * gx - 1, hi = c1 + (k - 1) * gx + 1;
        lo = lo < 0 ? 0 : (lo > nc - 1 ? nc - 1 : lo);
        hi = hi < 0 ? 0 : (hi > nc - 1 ? nc - 1 : hi);
        const int32_t b = B.start[lo], e = B.start[hi + 1];
        ub[k] = b;
        uo[k] = o;
        o += e > b ? e - b : 0;
      }
      uo[3] = o;
    }
    __syncthreads();
    const int32_t total = uo[3];
    if (total <= kNmaskLds) {  // uniform over the block
      for (int32_t i = threadIdx.x; i < total; i += kBlock) {
        const int k = (i >= uo[1]) + (i >= uo[2]);
        const int32_t b = ub[k] + (i - uo[k]);
        lds[i] = make_double2(XM[0][b], XM[1][b]);
      }
      __syncthreads();
      if (!active) return;
      const double xa0 = XM[0][a], xa1 = XM[1][a];
      const Rows<D> rows = stencil_rows<D>(G, B.start, B.skey[a]);
      int cnt = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double2* lrow = lds + (uo[k] - ub[k]);  // candidate b at lrow[b]
        for (int32_t w0 = rows.b[k]; w0 < rows.e[k]; w0 += 32) {
          const int32_t hi = rows.e[k] < w0 + 32 ? rows.e[k] : w0 + 32;
          const uint32_t mk = window_bits_lds(lrow + w0, xa0, xa1, a, w0, hi, dist2);
          if (mk != 0) {
            if (cnt < cap) out[int64_t(cnt) * B.n] = (uint64_t(uint32_t(w0)) << 32) | mk;
            ++cnt;
          }
        }
      }
      B.NC[a] = cnt <= cap ? cnt : -1;
      return;
    }
  }
  if (!active) return;
  const double xa0 = XM[0][a], xa1 = XM[1][a], xa2 = D == 3 ? XM[2][a] : 0.0;
  const Rows<D> rows = stencil_rows<D>(G, B.start, B.skey[a]);
  int cnt = 0;
#pragma unroll
  for (int k = 0; k < Rows<D>::N; ++k) {
    for (int32_t w0 = rows.b[k]; w0 < rows.e[k]; w0 += 32) {
      const int32_t hi = rows.e[k] < w0 + 32 ? rows.e[k] : w0 + 32;
      const uint32_t mk = window_bits<D>(XM, xa0, xa1, xa2, a, w0, hi, dist2);
      if (mk != 0) {
        if (cnt < cap) out[int64_t(cnt) * B.n] = (uint64_t(uint32_t(w0)) << 32) | mk;
        ++cnt;
      }
    }
  }
  B.NC[a] = cnt <= cap ? cnt : -1;
}

// Element of a per-particle array at a 32-bit byte offset: the load then uses
// the scalar base + 32-bit vector offset addressing form (no 64-bit address
// arithmetic per partner).  Arrays stay below 4 GiB (the problem size limit
// checked at construction).
template <typename T>
__device__ __forceinline__ T ldo(const T* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Neighbour-list partners b of particle a in a fixed order (stencil rows in
// order, ascending sorted index inside a row).  nc >= 0: the nc stored entries
// of k_nmask, read one ahead so the next load overlaps the current window's
// pair bodies.  nc < 0: the windows are computed here from the membership
// positions XM.  Each lane runs its own sequence, so all lanes of a wave
// execute the (long) pair body together.
template <int D, typename R>
struct Partners {
  const SphBufs<R>& B;
  const CellGeom& G;
  const double* const* XM;
  const double dist2;
  const int64_t a;
  const int nc;
  const uint64_t* pe;
  uint64_t nxt;
  int e = 1, k = -1;
  int32_t w0 = 0, re = 0, cb = 0;
  uint32_t c = 0;

  __device__ __forceinline__ Partners(const SphBufs<R>& B_, int64_t a_, int nc_, const CellGeom& G_,
                                      const double* const* XM_, double dist2_)
      : B(B_), G(G_), XM(XM_), dist2(dist2_), a(a_), nc(nc_), pe(B_.NE + a_), nxt(nc_ > 0 ? *pe : 0ull) {}

  __device__ __forceinline__ bool next(int32_t& b) {
    if (c == 0) {
      if (nc >= 0) {
        if (nxt == 0) return false;  // stored entries always have member bits
        c = uint32_t(nxt);
        cb = int32_t(nxt >> 32);
        pe += B.n;
        nxt = e < nc ? *pe : 0ull;
        ++e;
      } else {
        while (c == 0) {
          if (w0 >= re) {
            if (++k >= Rows<D>::N) break;
            row_range<D>(G, B.start, B.skey[a], k, w0, re);
            continue;
          }
          const int32_t hi = re < w0 + 32 ? re : w0 + 32;
          c = window_bits<D>(XM, XM[0][a], XM[1][a], D == 3 ? XM[2][a] : 0.0, a, w0, hi, dist2);
          cb = w0;
          w0 = hi;
        }
        if (c == 0) return false;
      }
    }
    b = cb + __builtin_ctz(c);
    c &= c - 1;
    return true;
  }
};

template <int D, typename R, typename F>
__device__ __forceinline__ void for_each_partner(const SphBufs<R>& B, int64_t a, int nc, const CellGeom& G,
                                                 const double* const* XM, double dist2, F&& body) {
  Partners<D, R> it(B, a, nc, G, XM, dist2);
  int32_t b;
  while (it.next(b)) body(b);
}

// As for_each_partner, with the partner data split off: fetch(b) loads what
// the body needs of partner b.  Pipe: the fetch of the next partner is issued
// before the body of the current one runs, so the loads overlap the pair
// arithmetic (at the cost of the registers holding both partners' data).
template <bool Pipe, int D, typename R, typename Fetch, typename F>
__device__ __forceinline__ void for_each_partner_fetch(const SphBufs<R>& B, int64_t a, int nc, const CellGeom& G,
                                                       const double* const* XM, double dist2, Fetch&& fetch,
                                                       F&& body) {
  Partners<D, R> it(B, a, nc, G, XM, dist2);
  int32_t b0, b1;
  if constexpr (!Pipe) {
    while (it.next(b0)) body(b0, fetch(b0));
    return;
  }
  // two partners in flight, alternating between two sets of registers; a
  // missing next partner fetches the current one again (a valid address)
  if (!it.next(b0)) return;
  auto f0 = fetch(b0);
  for (;;) {
    b1 = b0;
    const bool m1 = it.next(b1);
    auto f1 = fetch(b1);
    body(b0, f0);
    if (!m1) break;
    b0 = b1;
    const bool m0 = it.next(b0);
    f0 = fetch(b0);
    body(b1, f1);
    if (!m0) break;
  }
}

__device__ __forceinline__ double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }

constexpr int kRedSlots = 64;

// Block maximum (NaN wins) folded into one of kRedSlots slots by an atomic max
// on its bit pattern (non-negative doubles order like their bits, NaN above
// +Inf); spreading blocks over many slots avoids one hot address.  Every
// thread of the block must call it.
__device__ __forceinline__ void block_max(double v, bool active, double* slots, int blk) {
  __shared__ double wm[kBlock / 64];
  double m = wave_max(active ? v : 0.0);
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = wm[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) r = nanmax(r, wm[w]);
    atomic_max_nonneg(reinterpret_cast<unsigned long long*>(slots) + (blk % kRedSlots), r);
  }
  __syncthreads();
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
    if (mp > ctl->mppcell) ctl->mppcell = (long long)ceil(double(mp) * 1.05 + 1);  // src/celllist.jl:134-135
    ctl->rebuild = 0;
    ctl->xrb_cur = write_xrb ? 0 : 1;
  } else {
    ctl->rebuilt = 0;
  }
}

// New sorted position s takes the state of old position perm[s].
template <int D, typename R, typename SS>
__global__ void k_reorder(SS ss) {
  GCLSPH_SCENE(ss);
  const StepCtl* ctl = B.ctl;
  if (ctl->rebuild == 0) return;
  const int64_t s = a;
  if (s >= B.n) return;
  const int p = ctl->parity, q = p ^ 1, pp = ctl->pparity, pq = pp ^ 1;
  const int32_t src = B.perm[s];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double xv = B.X[p][k][src];
    B.X[q][k][s] = xv;
    B.V[q][k][s] = B.V[p][k][src];
    if (F.write_xrb) B.XRB[k][s] = xv;
    if (!F.synced_prev) {  // previous half-step state moves with its particle
      B.XH[k][s] = B.XH2[k][src];
      B.VH[k][s] = B.VH2[k][src];
    }
  }
  const R rv = B.RHO[p][src];
  const int32_t pt = B.PT[pp][src];
  B.RHO[q][s] = rv;
  if constexpr (std::is_same<R, double>::value) B.PN[s] = eos_d(rv, pt, C);
  else B.PN[s] = eos_f(rv, pt, C);
  if (!F.synced_prev) B.RH[s] = B.RH2[src];
  B.PT[pq][s] = pt;
  B.ID[pq][s] = B.ID[pp][src];
}

// ---- pass 1 ---------------------------------------------------------------
template <int D, typename R, typename SS>
__global__ __launch_bounds__(256) void k_pass1(SS ss) {
#pragma clang fp contract(off)
  GCLSPH_SCENE(ss);
  const StepCtl* ctl = B.ctl;
  const int p = ctl->parity, pp = ctl->pparity, rebuilt = ctl->rebuilt;
  const double dt = ctl->dt;
  if (a >= B.n) return;
  const double* X[3] = {B.X[p][0], B.X[p][1], B.X[p][D == 3 ? 2 : 1]};
  const R* V[3] = {B.V[p][0], B.V[p][1], B.V[p][D == 3 ? 2 : 1]};
  const R* RHO = B.RHO[p];
  const int32_t* PT = B.PT[pp];
  const double* XM[3];
  membership_ptrs<D>(B, F, p, rebuilt, XM);

  double xa[3];
  R va[3];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    xa[k] = X[k][a];
    va[k] = V[k][a];
  }
  const R sa = RHO[a];
  const int32_t pta = PT[a];
  Sum<R> acc[3], drho;

  const int nc = B.NC[a];
  if constexpr (std::is_same<R, double>::value) {
    const double rhoa = sa;
    const double Pa = B.PN[a];
    struct Nb {
      double x[3], v[3], rho, P;
      int32_t pt;
    };
    auto fetch = [&](const int32_t b) {
      const uint32_t o8 = uint32_t(b) * 8u, oR = uint32_t(b) * uint32_t(sizeof(R)), o4 = uint32_t(b) * 4u;
      Nb q;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        q.x[d] = ldo(X[d], o8);
        q.v[d] = ldo(V[d], oR);
      }
      q.rho = ldo(RHO, oR);
      q.P = ldo<R>(B.PN, oR);
      q.pt = ldo(PT, o4);
      return q;
    };
    for_each_partner_fetch<false, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& q) {
      const int32_t ptb = q.pt;
      double dx[3], dv[3], gw[3];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        dx[d] = xa[d] - q.x[d];
        dv[d] = va[d] - q.v[d];
      }
      const double r2 = dotd<D>(dx, dx);
      const double r = sqrt(r2);
      const double u = r * C.Hinv;
      const double dwk_r = dW_d(u, C.dwnorm) / r;
#pragma unroll
      for (int d = 0; d < D; ++d) gw[d] = dx[d] * dwk_r;
      const double rhob = q.rho;
      const bool both = pta > 0 && ptb > 0;
      double ax[3] = {0.0, 0.0, 0.0};
      if (both) {  // ∂ρ∂tDDT! (:579-625)
        double dot3 = -dotd<D>(dx, gw);
        double s = C.m0 * dotd<D>(dv, gw);
        if (pta > 1) {
          double drhop = C.rho0 * pf7(1 + C.DDTgz * dx[1], C.gamma, C.ginv) - C.rho0;
          double vd = C.DDTkh * (rhob - rhoa - drhop) / (r2 + C.eta2);
          s += vd * dot3 * C.m0 / rhob;
        }
        drho.add(s);
      }
      if ((long long)pta * (long long)ptb > 0) {  // ∂v∂t! (:755-772)
        double Pfac = (Pa + q.P) / (rhoa * rhob);
#pragma unroll
        for (int d = 0; d < D; ++d) ax[d] += -C.m0 * Pfac * gw[d];
      }
      if (both) {  // ∂v∂t_av! (:869-897)
        double cond = dotd<D>(dv, dx);
        if (cond < 0) {
          double dmu = C.h * cond / (r2 + C.eta2);
          double dPi = (-C.alpha * C.c0 * dmu) / ((rhoa + rhob) * 0.5);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += -dPi * C.m0 * gw[d];
        }
        if (F.nu_on) {  // ∂v∂t_visc! (:967-980)
          double term = 4 * C.nu * C.m0 * dotd<D>(dx, gw) / ((rhoa + rhob) * (r2 + C.eta2));
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += term * dv[d];
        }
        if (F.s_on && r < C.st_2h) {  // ∂v∂tpF! (:1387-1397)
          double scos = C.s * cos(C.st_c * r / C.st_2h) / (r + C.st_01h);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += scos * dx[d] / C.m0;
        }
      }
      if ((pta > 0) != (ptb > 0) && r < C.fb_r0[0]) {  // fbmolforce! (:1891-1901)
        double q1 = C.fb_r0[0] / (r + C.fb_eta2[0]);
        double Fc = C.fb_d[0] * (powi_d(q1, 12) - powi_d(q1, 4)) / (r2 + C.fb_eta2[0]);
#pragma unroll
        for (int d = 0; d < D; ++d) ax[d] += dx[d] * Fc;
      }
#pragma unroll
      for (int d = 0; d < D; ++d) acc[d].add(ax[d]);
    });
    // ∂v∂t_addgrav!, then the half-step updates (:270-274)
    double av[3];
#pragma unroll
    for (int d = 0; d < D; ++d) av[d] = acc[d].get();
    av[1] = av[1] - C.g;
    double rh0, vh0[3], xh0[3];
    if (F.synced_prev) {
      rh0 = rhoa;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        vh0[d] = va[d];
        xh0[d] = xa[d];
      }
    } else {
      const bool mv = rebuilt;  // the reorder moved the old half-step state to XH/VH/RH
      rh0 = mv ? B.RH[a] : B.RH2[a];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        vh0[d] = mv ? B.VH[d][a] : B.VH2[d][a];
        xh0[d] = mv ? B.XH[d][a] : B.XH2[d][a];
      }
    }
    const double hdt = dt * 0.5;
    double rv = rh0 + drho.get() * hdt;
    if (rv < C.rho0 && pta < 0) rv = C.rho0;
    B.RH[a] = rv;
    B.PH[a] = eos_d(rv, pta, C);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double vv = vh0[d];
      if (pta > 1) vv += av[d] * hdt;
      B.VH[d][a] = vv;
      B.XH[d][a] = xh0[d] + vv * hdt;
    }
  } else {
    // ---------------- FP32-split
#pragma clang fp contract(fast)
    const float dra = sa;
    const float rhoa = C.f_rho0 + dra;
    const float inv_rhoa = __builtin_amdgcn_rcpf(rhoa);
    const float Pa = B.PN[a];
    struct Nb {
      double x[3];
      float v[3], rho, P;
      int32_t pt;
    };
    auto fetch = [&](const int32_t b) {
      const uint32_t o8 = uint32_t(b) * 8u, oR = uint32_t(b) * uint32_t(sizeof(R)), o4 = uint32_t(b) * 4u;
      Nb q;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        q.x[d] = ldo(X[d], o8);
        q.v[d] = ldo(V[d], oR);
      }
      q.rho = ldo(RHO, oR);
      q.P = ldo<R>(B.PN, oR);
      q.pt = ldo(PT, o4);
      return q;
    };
    for_each_partner_fetch<true, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& q) {
      const int32_t ptb = q.pt;
      float dx[3], dv[3], gw[3];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        dx[d] = float(xa[d] - q.x[d]);
        dv[d] = va[d] - q.v[d];
      }
      const float r2 = dotf<D>(dx, dx);
      const float rinv = __builtin_amdgcn_rsqf(r2);
      const float r = r2 * rinv;
      const float u = r * C.f_Hinv;
      const float dwk_r = dW_f(u, C.f_dwnorm) * rinv;
      const float inv_den = __builtin_amdgcn_rcpf(r2 + C.f_eta2);
#pragma unroll
      for (int d = 0; d < D; ++d) gw[d] = dx[d] * dwk_r;
      const float drb = q.rho;
      const float rhob = C.f_rho0 + drb;
      const float inv_rhob = __builtin_amdgcn_rcpf(rhob);
      const bool both = pta > 0 && ptb > 0;
      float ax[3] = {0.0f, 0.0f, 0.0f};
      if (both) {
        float dot3 = -dotf<D>(dx, gw);
        float s = C.f_m0 * dotf<D>(dv, gw);
        if (pta > 1) {
          float drhop = hydro_drho_f(C.f_DDTgz * dx[1], C.f_rho0);
          float vd = C.f_DDTkh * ((drb - dra) - drhop) * inv_den;
          s += vd * dot3 * C.f_m0 * inv_rhob;
        }
        drho.add(s);
      }
      if ((long long)pta * (long long)ptb > 0) {
        float Pfac = (Pa + q.P) * (inv_rhoa * inv_rhob);
#pragma unroll
        for (int d = 0; d < D; ++d) ax[d] += -C.f_m0 * Pfac * gw[d];
      }
      if (both) {
        float cond = dotf<D>(dv, dx);
        if (cond < 0.0f) {
          float dmu = C.f_h * cond * inv_den;
          float dPi = (-C.f_alpha * C.f_c0 * dmu) * (2.0f * __builtin_amdgcn_rcpf(rhoa + rhob));
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += -dPi * C.f_m0 * gw[d];
        }
        if (F.nu_on) {
          float term = 4.0f * C.f_nu * C.f_m0 * dotf<D>(dx, gw) * (__builtin_amdgcn_rcpf(rhoa + rhob) * inv_den);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += term * dv[d];
        }
        if (F.s_on && r < C.f_st_2h) {
          float scos = C.f_s * cosf(C.f_st_c * r / C.f_st_2h) / (r + C.f_st_01h);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += scos * dx[d] * (1.0f / C.f_m0);
        }
      }
      if ((pta > 0) != (ptb > 0) && r < C.f_fb_r0[0]) {
        // wall force in FP64: (r0/r)^12 overflows FP32 for near-coincident pairs
        double q1 = C.fb_r0[0] / (double(r) + C.fb_eta2[0]);
        double q2 = q1 * q1, q4 = q2 * q2, q8 = q4 * q4;
        double Fc = C.fb_d[0] * (q4 * q8 - q4) / (double(r2) + C.fb_eta2[0]);
#pragma unroll
        for (int d = 0; d < D; ++d) ax[d] += float(double(dx[d]) * Fc);
      }
#pragma unroll
      for (int d = 0; d < D; ++d) acc[d].add(ax[d]);
    });
    float av[3];
#pragma unroll
    for (int d = 0; d < D; ++d) av[d] = acc[d].get();
    av[1] = av[1] - C.f_g;
    float rh0, vh0[3];
    double xh0[3];
    if (F.synced_prev) {
      rh0 = dra;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        vh0[d] = va[d];
        xh0[d] = xa[d];
      }
    } else {
      const bool mv = rebuilt;
      rh0 = mv ? B.RH[a] : B.RH2[a];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        vh0[d] = mv ? B.VH[d][a] : B.VH2[d][a];
        xh0[d] = mv ? B.XH[d][a] : B.XH2[d][a];
      }
    }
    const float hdt = float(dt * 0.5);
    float rv = rh0 + drho.get() * hdt;
    if (rv < 0.0f && pta < 0) rv = 0.0f;
    B.RH[a] = rv;
    B.PH[a] = eos_f(rv, pta, C);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float vv = vh0[d];
      if (pta > 1) vv += av[d] * hdt;
      B.VH[d][a] = vv;
      B.XH[d][a] = xh0[d] + double(vv) * (dt * 0.5);
    }
  }
}

// ---- pass 2 + symplectic update ------------------------------------------
template <int D, typename R, typename SS>
__global__ __launch_bounds__(256) void k_pass2(SS ss) {
#pragma clang fp contract(off)
  GCLSPH_SCENE(ss);
  StepCtl* ctl = B.ctl;
  const int p = ctl->parity, q = p ^ 1, pp = ctl->pparity, rebuilt = ctl->rebuilt;
  const double dt = ctl->dt;
  const bool active = a < B.n;
  double red_cdx = 0.0, red_acc = 0.0, red_visc = 0.0;
  if (active) {
    const double* X[3] = {B.X[p][0], B.X[p][1], B.X[p][D == 3 ? 2 : 1]};
    const double* XH[3] = {B.XH[0], B.XH[1], B.XH[D == 3 ? 2 : 1]};
    const R* VH[3] = {B.VH[0], B.VH[1], B.VH[D == 3 ? 2 : 1]};
    const R* RH = B.RH;
    const int32_t* PT = B.PT[pp];
    const double* XM[3];
    membership_ptrs<D>(B, F, p, rebuilt, XM);
    double xa[3], xha[3];
    R vha[3];
#pragma unroll
    for (int k = 0; k < D; ++k) {
      xa[k] = X[k][a];
      xha[k] = XH[k][a];
      vha[k] = VH[k][a];
    }
    const R sha = RH[a];
    const int32_t pta = PT[a];
    Sum<R> acc[3], drho;
    const int nc = B.NC[a];
    R av[3];
    R Pha, drho_v;
    if constexpr (std::is_same<R, double>::value) {
      const double rhoa = sha;
      Pha = B.PH[a];
      struct Nb {
        double xn[3], xh[3], vh[3], rh, P;
        int32_t pt;
      };
      auto fetch = [&](const int32_t b) {
        const uint32_t o8 = uint32_t(b) * 8u, oR = uint32_t(b) * uint32_t(sizeof(R)), o4 = uint32_t(b) * 4u;
        Nb q;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          q.xn[d] = ldo(X[d], o8);
          q.xh[d] = ldo(XH[d], o8);
          q.vh[d] = ldo(VH[d], oR);
        }
        q.rh = ldo(RH, oR);
        q.P = ldo<R>(B.PH, oR);
        q.pt = ldo(PT, o4);
        return q;
      };
      for_each_partner_fetch<true, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& q) {
        const int32_t ptb = q.pt;
        double dxn[3], dx[3], dv[3], gw[3];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          dxn[d] = xa[d] - q.xn[d];
          dx[d] = xha[d] - q.xh[d];
          dv[d] = vha[d] - q.vh[d];
        }
        // ∇W from pass 1 (positions xⁿ, src/sphproblem.jl:246,281)
        const double r2n = dotd<D>(dxn, dxn);
        const double rn = sqrt(r2n);
        const double dwk_r = dW_d(rn * C.Hinv, C.dwnorm) / rn;
#pragma unroll
        for (int d = 0; d < D; ++d) gw[d] = dxn[d] * dwk_r;
        const double r2 = dotd<D>(dx, dx);
        const double rhob = q.rh;
        const bool both = pta > 0 && ptb > 0;
        double ax[3] = {0.0, 0.0, 0.0};
        if (both) {
          double dot3 = -dotd<D>(dx, gw);
          double s = C.m0 * dotd<D>(dv, gw);
          if (pta > 1) {
            double drhop = C.rho0 * pf7(1 + C.DDTgz * dx[1], C.gamma, C.ginv) - C.rho0;
            double vd = C.DDTkh * (rhob - rhoa - drhop) / (r2 + C.eta2);
            s += vd * dot3 * C.m0 / rhob;
          }
          drho.add(s);
        }
        if ((long long)pta * (long long)ptb > 0) {
          double Pfac = (Pha + q.P) / (rhoa * rhob);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += -C.m0 * Pfac * gw[d];
        }
        if (both) {
          double cond = dotd<D>(dv, dx);
          if (cond < 0) {
            double dmu = C.h * cond / (r2 + C.eta2);
            double dPi = (-C.alpha * C.c0 * dmu) / ((rhoa + rhob) * 0.5);
#pragma unroll
            for (int d = 0; d < D; ++d) ax[d] += -dPi * C.m0 * gw[d];
          }
          if (F.nu_on) {
            double term = 4 * C.nu * C.m0 * dotd<D>(dx, gw) / ((rhoa + rhob) * (r2 + C.eta2));
#pragma unroll
            for (int d = 0; d < D; ++d) ax[d] += term * dv[d];
          }
          if (F.s_on) {
            double r = sqrt(r2);
            if (r < C.st_2h) {
              double scos = C.s * cos(C.st_c * r / C.st_2h) / (r + C.st_01h);
#pragma unroll
              for (int d = 0; d < D; ++d) ax[d] += scos * dx[d] / C.m0;
            }
          }
        }
        // fbmolforce! at xⁿ with (bound_D, bound_l) (src/sphproblem.jl:295)
        if ((pta > 0) != (ptb > 0) && rn < C.fb_r0[1]) {
          double q1 = C.fb_r0[1] / (rn + C.fb_eta2[1]);
          double Fc = C.fb_d[1] * (powi_d(q1, 12) - powi_d(q1, 4)) / (r2n + C.fb_eta2[1]);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += dxn[d] * Fc;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d].add(ax[d]);
      });
#pragma unroll
      for (int d = 0; d < D; ++d) av[d] = acc[d].get();
      av[1] = av[1] - C.g;
      drho_v = drho.get();
      // symplectic_update! (src/gpukernels_sph.jl:1204-1245)
      const double epsi = -(drho_v / rhoa) * dt;
      double rv = B.RHO[p][a] * (2 - epsi) / (2 + epsi);
      if (rv < C.rho0 && pta <= 1) rv = C.rho0;
      B.RHO[q][a] = rv;
      B.PN[a] = eos_d(rv, pta, C);
      if (F.corr_on) B.RH2[a] = rv;
      const double ml = pta > 1 ? 1.0 : 0.0;
      double xn[3], vn[3];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        double vv = B.V[p][d][a];
        double nv = vv + av[d] * dt * ml;
        double dd = (vv + nv) * 0.5 * dt;
        double cd = dd;
        if (F.need_cdx) {
          cd = (rebuilt ? 0.0 : B.CDX[d][a]) + dd;
          B.CDX[d][a] = cd;
        }
        red_cdx = fmax(red_cdx, fabs(cd)) + (cd != cd ? cd : 0.0);
        double xv = xa[d] + dd;
        xn[d] = xv;
        vn[d] = nv;
        B.V[q][d][a] = nv;
        B.X[q][d][a] = xv;
        if (F.corr_on) {
          B.VH2[d][a] = nv;
          B.XH2[d][a] = xv;
        }
      }
      if (F.ts) {
        red_acc = sqrt(dotd<D>(av, av));
        if (F.visc_in_pass2) red_visc = fabs(C.h * dotd<D>(vn, xn) / (dotd<D>(xn, xn) + C.dt_eta2));
      }
    } else {
#pragma clang fp contract(fast)
      const float dra = sha;
      const float rhoa = C.f_rho0 + dra;
      const float inv_rhoa = __builtin_amdgcn_rcpf(rhoa);
      Pha = B.PH[a];
      struct Nb {
        double xn[3], xh[3];
        float vh[3], rh, P;
        int32_t pt;
      };
      auto fetch = [&](const int32_t b) {
        const uint32_t o8 = uint32_t(b) * 8u, oR = uint32_t(b) * uint32_t(sizeof(R)), o4 = uint32_t(b) * 4u;
        Nb q;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          q.xn[d] = ldo(X[d], o8);
          q.xh[d] = ldo(XH[d], o8);
          q.vh[d] = ldo(VH[d], oR);
        }
        q.rh = ldo(RH, oR);
        q.P = ldo<R>(B.PH, oR);
        q.pt = ldo(PT, o4);
        return q;
      };
      for_each_partner_fetch<true, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& q) {
        const int32_t ptb = q.pt;
        float dxn[3], dx[3], dv[3], gw[3];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          dxn[d] = float(xa[d] - q.xn[d]);
          dx[d] = float(xha[d] - q.xh[d]);
          dv[d] = vha[d] - q.vh[d];
        }
        const float r2n = dotf<D>(dxn, dxn);
        const float rninv = __builtin_amdgcn_rsqf(r2n);
        const float rn = r2n * rninv;
        const float dwk_r = dW_f(rn * C.f_Hinv, C.f_dwnorm) * rninv;
#pragma unroll
        for (int d = 0; d < D; ++d) gw[d] = dxn[d] * dwk_r;
        const float r2 = dotf<D>(dx, dx);
        const float inv_den = __builtin_amdgcn_rcpf(r2 + C.f_eta2);
        const float drb = q.rh;
        const float rhob = C.f_rho0 + drb;
        const float inv_rhob = __builtin_amdgcn_rcpf(rhob);
        const bool both = pta > 0 && ptb > 0;
        float ax[3] = {0.0f, 0.0f, 0.0f};
        if (both) {
          float dot3 = -dotf<D>(dx, gw);
          float s = C.f_m0 * dotf<D>(dv, gw);
          if (pta > 1) {
            float drhop = hydro_drho_f(C.f_DDTgz * dx[1], C.f_rho0);
            float vd = C.f_DDTkh * ((drb - dra) - drhop) * inv_den;
            s += vd * dot3 * C.f_m0 * inv_rhob;
          }
          drho.add(s);
        }
        if ((long long)pta * (long long)ptb > 0) {
          float Pfac = (Pha + q.P) * (inv_rhoa * inv_rhob);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += -C.f_m0 * Pfac * gw[d];
        }
        if (both) {
          float cond = dotf<D>(dv, dx);
          if (cond < 0.0f) {
            float dmu = C.f_h * cond * inv_den;
            float dPi = (-C.f_alpha * C.f_c0 * dmu) * (2.0f * __builtin_amdgcn_rcpf(rhoa + rhob));
#pragma unroll
            for (int d = 0; d < D; ++d) ax[d] += -dPi * C.f_m0 * gw[d];
          }
          if (F.nu_on) {
            float term = 4.0f * C.f_nu * C.f_m0 * dotf<D>(dx, gw) * (__builtin_amdgcn_rcpf(rhoa + rhob) * inv_den);
#pragma unroll
            for (int d = 0; d < D; ++d) ax[d] += term * dv[d];
          }
          if (F.s_on) {
            float r = sqrtf(r2);
            if (r < C.f_st_2h) {
              float scos = C.f_s * cosf(C.f_st_c * r / C.f_st_2h) / (r + C.f_st_01h);
#pragma unroll
              for (int d = 0; d < D; ++d) ax[d] += scos * dx[d] * (1.0f / C.f_m0);
            }
          }
        }
        if ((pta > 0) != (ptb > 0) && rn < C.f_fb_r0[1]) {
          // wall force in FP64: (r0/r)^12 overflows FP32 for near-coincident pairs
          double q1 = C.fb_r0[1] / (double(rn) + C.fb_eta2[1]);
          double q2 = q1 * q1, q4 = q2 * q2, q8 = q4 * q4;
          double Fc = C.fb_d[1] * (q4 * q8 - q4) / (double(r2n) + C.fb_eta2[1]);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += float(double(dxn[d]) * Fc);
        }
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d].add(ax[d]);
      });
#pragma unroll
      for (int d = 0; d < D; ++d) av[d] = acc[d].get();
      av[1] = av[1] - C.f_g;
      drho_v = drho.get();
      // symplectic update on δρ: ρ(2-ε)/(2+ε) - ρ0 = (2δρ - ε(2ρ0 + δρ)) / (2 + ε)
      const float fdt = float(dt);
      const float epsi = -(drho_v / rhoa) * fdt;
      const float dr_n = B.RHO[p][a];
      float rv = (2.0f * dr_n - epsi * (2.0f * C.f_rho0 + dr_n)) / (2.0f + epsi);
      if (rv < 0.0f && pta <= 1) rv = 0.0f;
      B.RHO[q][a] = rv;
      B.PN[a] = eos_f(rv, pta, C);
      if (F.corr_on) B.RH2[a] = rv;
      const float ml = pta > 1 ? 1.0f : 0.0f;
      double xn[3], vn[3];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        float vv = B.V[p][d][a];
        float nv = vv + av[d] * fdt * ml;
        double dd = (double(vv) + double(nv)) * 0.5 * dt;
        double cd = dd;
        if (F.need_cdx) {
          cd = (rebuilt ? 0.0 : B.CDX[d][a]) + dd;
          B.CDX[d][a] = cd;
        }
        red_cdx = fmax(red_cdx, fabs(cd)) + (cd != cd ? cd : 0.0);
        double xv = xa[d] + dd;
        xn[d] = xv;
        vn[d] = nv;
        B.V[q][d][a] = nv;
        B.X[q][d][a] = xv;
        if (F.corr_on) {
          B.VH2[d][a] = nv;
          B.XH2[d][a] = xv;
        }
      }
      if (F.ts) {
        float a2 = dotf<D>(av, av);
        red_acc = sqrt(double(a2));
        if (F.visc_in_pass2) red_visc = fabs(C.h * dotd<D>(vn, xn) / (dotd<D>(xn, xn) + C.dt_eta2));
      }
    }
    if (F.want_out) {
#pragma unroll
      for (int d = 0; d < D; ++d) B.ACC[d][a] = av[d];
      B.DRHO[a] = drho_v;
      B.PBUF[a] = Pha;
    }
  }
  block_max(red_cdx, active, B.RED, m.y);
  if (F.ts) {
    block_max(red_acc, active, B.RED + kRedSlots, m.y);
    if (F.visc_in_pass2) block_max(red_visc, active, B.RED + 2 * kRedSlots, m.y);
  }
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
  struct Nb {
    double x[3];
    R v[3], rho, P;
    int32_t pt;
  };
  auto fetch = [&](const int32_t b) {
    const uint32_t o8 = uint32_t(b) * 8u, oR = uint32_t(b) * uint32_t(sizeof(R)), o4 = uint32_t(b) * 4u;
    Nb n;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      n.x[d] = ldo(X[d], o8);
      n.v[d] = ldo(V[d], oR);
    }
    n.rho = ldo(RHO, oR);
    n.P = ldo<R>(B.PN, oR);
    n.pt = ldo(PT, o4);
    return n;
  };
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
    for_each_partner_fetch<true, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& n) {
      if (n.pt <= 0) return;
      float dx[3], dv[3];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        dx[d] = float(xa[d] - n.x[d]);
        dv[d] = va[d] - n.v[d];
      }
      const float rhob = C.f_rho0 + n.rho;
      const float r2 = dotf<D>(dx, dx);
      const float u = __builtin_amdgcn_sqrtf(r2) * C.f_dpc_l0inv;
      const float w = W_f(u, C.f_dpc_wnorm);
      const float chi = __builtin_amdgcn_sqrtf(w * C.f_dpc_whinv);
      const float kk = u < C.f_dpc_lim ? 1.0f : chi;
      const float lo = C.f_dpc_lambda * fabsf(Pa + n.P), cap = C.f_dpc_lambda * C.f_dpc_pmax;
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
    });
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
#pragma unroll
  for (int d = 0; d < D; ++d) {
    R dv = B.DV[d][a];
    B.V[q][d][a] = B.V[q][d][a] + dv;
    B.X[q][d][a] = B.X[q][d][a] + double(dv) * dt;
  }
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
  struct Nb {
    double x[3];
    R v[3], rho;
    int32_t pt;
  };
  auto fetch = [&](const int32_t b) {
    const uint32_t o8 = uint32_t(b) * 8u, oR = uint32_t(b) * uint32_t(sizeof(R)), o4 = uint32_t(b) * 4u;
    Nb n;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      n.x[d] = ldo(XN[d], o8);
      n.v[d] = ldo(V[d], oR);
    }
    n.rho = ldo(RHO, oR);
    n.pt = ldo(PT, o4);
    return n;
  };
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
    for_each_partner_fetch<true, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& n) {
      if (n.pt <= 1) return;
      float dx[3];
#pragma unroll
      for (int d = 0; d < D; ++d) dx[d] = float(xa[d] - n.x[d]);
      const float r = __builtin_amdgcn_sqrtf(dotf<D>(dx, dx));
      const float w = W_f(r * C.f_Hinv, C.f_wnorm);
      const float xs = C.f_xsph_c * w * __builtin_amdgcn_rcpf(rhoa + (C.f_rho0 + n.rho));
#pragma unroll
      for (int d = 0; d < D; ++d) sum[d].add((n.v[d] - va[d]) * xs);
    });
#pragma unroll
    for (int d = 0; d < D; ++d) out[d] = double(sum[d].get());
  }
#pragma unroll
  for (int d = 0; d < D; ++d) B.X[q][d][a] = B.X[q][d][a] + out[d] * dt;
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

template <typename R, typename SS>
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
    B.PN[a] = eos_d(nr, pt, C);
  } else {
    const float d = float(nr);
    B.RHO[q][a] = d;
    B.PN[a] = eos_f(d, pt, C);
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
  int32_t v[kScanItems];
  int64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    v[k] = i < n ? sd.cnt[i] : 0;
    s += v[k];
  }
  int64_t tot;
  int64_t pre = block_exclusive_scan<int64_t>(s, &tot) + sd.scan_tmp[m.y];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (i < n) sd.start[i] = int32_t(pre);
    pre += v[k];
  }
  if (m.y == 0 && threadIdx.x == 0) sd.start[n] = int32_t(sd.scan_tmp[nt]);
}

template <int D, typename R, typename SS>
__global__ void kb_scatter(SS ss) {
  GCLSPH_SCENE(ss);
  if (B.ctl->rebuild == 0 || a >= B.n) return;
  sd.tmp[sd.start[sd.key[a]] + sd.slot[a]] = int32_t(a);
}

// Rank inside the cell by input index: the final order is (cell, input).
template <int D, typename R, typename SS>
__global__ void kb_rank(SS ss) {
  GCLSPH_SCENE(ss);
  if (B.ctl->rebuild == 0 || a >= B.n) return;
  const int32_t i = sd.tmp[a];
  const int32_t c = sd.key[i];
  const int32_t b = sd.start[c], e = sd.start[c + 1];
  int32_t r = 0;
  for (int32_t k = b; k < e; ++k) r += sd.tmp[k] < i;
  sd.perm[b + r] = i;
  sd.skey[b + r] = c;
}

// ------------------------------------------------------------ host launchers
namespace {
// Calls f with the scene source of L: the descriptor by value for one problem,
// the device descriptor array and maps for several.
template <typename R, typename Fn>
void with_scenes(const LaunchSet& L, Fn&& f) {
  if (L.one) f(OneScene<R>{*static_cast<const SceneDev<R>*>(L.one)});
  else f(ManyScenes<R>{static_cast<const SceneDev<R>*>(L.scenes), L.pmap, L.cmap});
}
}  // namespace

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
    k_reorder<D, R, SS><<<L.pblocks, kBlock, 0, s>>>(ss);
    k_post_rebuild<R, SS><<<L.count, 64, 0, s>>>(ss);
    k_nmask<D, R, SS><<<L.pblocks, kBlock, 0, s>>>(ss);
  });
}
template <int D, typename R>
void launch_init_pn(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) { k_init_pn<R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss); });
}
template <int D, typename R>
void launch_pass1(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) { k_pass1<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss); });
}
template <int D, typename R>
void launch_pass2(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) { k_pass2<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss); });
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
  with_scenes<R>(L, [&](auto ss) { k_xsph<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss); });
}
template <int D, typename R>
void launch_cspm(const LaunchSet& L, hipStream_t s) {
  with_scenes<R>(L, [&](auto ss) {
    k_cspm_sum<D, R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss);
    k_cspm_apply<R, decltype(ss)><<<L.pblocks, kBlock, 0, s>>>(ss);
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
  template void launch_pass1<D, R>(const LaunchSet&, hipStream_t);                       \
  template void launch_pass2<D, R>(const LaunchSet&, hipStream_t);                       \
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
