// Device helpers of the SPH step shared by the pass kernels (sph_pass.hip)
// and the other step kernels (sph_kernels.hip): SPH kernel functions, EOS,
// accumulators, scene sources, the neighbour-membership walk and reductions.
#pragma once

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
// (dW/dr)/r at u = r/H, i.e. ∇W = dx · dWr_f: -20 (1-u)^3 dwnorm / H (u/r = 1/H).
__device__ __forceinline__ float dWr_f(float u, float dwr) {
  const float um1 = 1.0f - u;
  return u < 1.0f ? dwr * (um1 * um1 * um1) : 0.0f;
}

// Tait EOS (src/gpukernels_sph.jl:691-697): 0 if ptype <= 1 and ρ < ρ0.
__device__ __forceinline__ double eos_d(double rho, int32_t pt, const SphConst& C) {
#pragma clang fp contract(off)
  if (pt <= 1 && rho < C.rho0) return 0.0;
  return C.P0 * (powi_d(rho / C.rho0, C.gamma) - 1);
}
// ptype(a)·ptype(b) > 0 of the reference (an Int64 product) for int32 types:
// both positive or both negative, without the 64-bit multiply.
__device__ __forceinline__ bool same_kind(int32_t pa, int32_t pb) { return (pa > 0 && pb > 0) || (pa < 0 && pb < 0); }

// FP32-split 1/ρb and 1/(ρa + ρb) from one hardware reciprocal of
// ρb(ρa + ρb) (densities near ρ0, so the product stays far from the FP32
// range limits).
struct RhoInv {
  float inv_rhob, inv_sum;
};
__device__ __forceinline__ RhoInv rho_inv2(float rhoa, float rhob) {
  const float sum = rhoa + rhob;
  const float R = __builtin_amdgcn_rcpf(rhob * sum);
  return {sum * R, rhob * R};
}

// FP32-split EOS on δρ: P0((1+e)^7 - 1) = P0 e (7 + 21e + 35e² + 35e³ + 21e⁴ + 7e⁵ + e⁶).
__device__ __forceinline__ float eos_f(float drho, int32_t pt, const SphConst& C) {
  if (pt <= 1 && drho < 0.0f) return 0.0f;
  float e = drho * C.f_inv_rho0;
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

// FP32-split ∂v∂tpF! coefficient s·cos(1.5π r/2h)/(r + 0.1h), for r < 2h: the
// hardware cosine (argument in turns, here in [0, 0.75)) and reciprocal.
__device__ __forceinline__ float surf_f(float r, const SphConst& C) {
  return C.f_s * __builtin_amdgcn_cosf(r * C.f_st_turn) * __builtin_amdgcn_rcpf(r + C.f_st_01h);
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
  const int2* __restrict__ pmap512;  // (problem, 512-particle block of the problem)
};
// Blocks of BS threads (128, 256 or 512) over one problem's particles: m.y is
// the block's index in units of BS particles, so particle a = m.y·BS + tid.
template <int BS = kBlock, typename R>
__device__ __forceinline__ const SceneDev<R>& particle_scene(const OneScene<R>& s, int2& m) {
  m = make_int2(0, int(xcd_block()));
  return s.d;
}
// Descriptor i of a batch, read through the constant address space: the
// descriptors are not written while a kernel runs, so their loads are
// invariant (scalar loads the compiler may repeat at will, as for a kernel
// argument), and the buffer pointers read from them are used as global-memory
// pointers instead of falling back to flat addressing.
template <typename R>
__device__ __forceinline__ const SceneDev<R>& batch_scene(const ManyScenes<R>& s, int i) {
  using CS = const __attribute__((address_space(4))) SceneDev<R>;
  CS* c = (CS*)(s.S);
  return *(const SceneDev<R>*)(c + i);
}
template <int BS = kBlock, typename R>
__device__ __forceinline__ const SceneDev<R>& particle_scene(const ManyScenes<R>& s, int2& m) {
  static_assert(BS == 128 || BS == 256 || BS == 512, "block sizes of the particle maps");
  const unsigned L = xcd_block();
  if constexpr (BS == 256) {
    m = s.pmap[L];
  } else if constexpr (BS == 128) {  // halves of the 256-particle blocks
    const int2 e = s.pmap[L >> 1];
    m = make_int2(e.x, e.y * 2 + int(L & 1u));
  } else {
    m = s.pmap512[L];
  }
  return batch_scene(s, m.x);
}
template <typename R>
__device__ __forceinline__ const SceneDev<R>& tile_scene(const OneScene<R>& s, int2& m) {
  m = make_int2(0, int(blockIdx.x));
  return s.d;
}
template <typename R>
__device__ __forceinline__ const SceneDev<R>& tile_scene(const ManyScenes<R>& s, int2& m) {
  m = s.cmap[blockIdx.x];
  return batch_scene(s, m.x);
}
template <typename R>
__device__ __forceinline__ const SceneDev<R>& scene_at(const OneScene<R>& s, unsigned) {
  return s.d;
}
template <typename R>
__device__ __forceinline__ const SceneDev<R>& scene_at(const ManyScenes<R>& s, unsigned i) {
  return batch_scene(s, int(i));
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

// Binds the problem of this block of BS threads: B, C, G, F and the particle
// index a.
#define GCLSPH_SCENE_BS(ss, BS)                                \
  int2 m;                                                      \
  const SceneDev<R>& sd = particle_scene<BS, R>(ss, m);        \
  const SphBufs<R>& B = sd.B;                                  \
  const SphConst& C = sd.C;                                    \
  const CellGeom& G = sd.G;                                    \
  const StepFlags F = step_flags(sd);                          \
  const int64_t a = int64_t(m.y) * (BS) + threadIdx.x;         \
  (void)C;                                                     \
  (void)G;                                                     \
  (void)F
#define GCLSPH_SCENE(ss) GCLSPH_SCENE_BS(ss, kBlock)

// Member bits of one window [lo, hi) (hi - lo <= 32) of a stencil row: bit j is
// set when candidate lo + j is a partner of particle a, i.e. b != a and
// r^2 < dist^2 with the reference's expression (no FMA).  Candidates are read
// four at a time so their loads are in flight together; reads run up to 3 past
// hi, which the position arrays' padding keeps in bounds, and those bits are
// dropped.
template <int D>
__device__ __forceinline__ uint32_t window_bits(const double* x0, const double* x1, const double* x2, double xa0,
                                                double xa1, double xa2, int64_t a, int32_t lo, int32_t hi,
                                                double dist2) {
#pragma clang fp contract(off)
  const double* p0 = x0 + lo;
  const double* p1 = x1 + lo;
  const double* p2 = x2 + lo;
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

// window_bits for the stencil scan inside the pair loops (a particle whose
// windows overflowed the membership store): one candidate at a time, so the
// scan adds few registers to the loop around it (the four-wide form above
// sets the register count, and so the occupancy, of every pair kernel).  Same
// expression, same bits.
template <int D>
__device__ __forceinline__ uint32_t window_bits_lean(const double* x0, const double* x1, const double* x2,
                                                     double xa0, double xa1, double xa2, int64_t a, int32_t lo,
                                                     int32_t hi, double dist2) {
#pragma clang fp contract(off)
  uint32_t mk = 0;
#pragma clang loop unroll(disable)
  for (int32_t b = lo; b < hi; ++b) {
    const double d0 = xa0 - x0[b];
    const double d1 = xa1 - x1[b];
    double r2 = d0 * d0 + d1 * d1;
    if (D == 3) {
      const double d2 = xa2 - x2[b];
      r2 = r2 + d2 * d2;
    }
    mk |= uint32_t(r2 < dist2 && b != a) << (b - lo);
  }
  return mk;
}

// Sorted range [b, e) of stencil row k (dy = k%3 - 1, dz = k/3 - 1) of cell c;
// empty when the row lies outside the grid.
template <int D>
__device__ __forceinline__ void row_range(const CellGeom& g, const int32_t* __restrict__ start, int32_t c, int k,
                                          int32_t& b, int32_t& e) {
  const int64_t gx = g.grid[0], gy = g.grid[1];
  int64_t ix, iy, iz;
  cell_xyz<D>(g, c, ix, iy, iz);
  const int64_t xl = ix > 0 ? ix - 1 : 0, xh = ix + 1 < gx ? ix + 1 : gx - 1;
  const int64_t y = iy + (k % 3) - 1;
  const int64_t z = D == 3 ? iz + (k / 3) - 1 : 0;
  const bool in = y >= 0 && y < gy && (D == 2 || (z >= 0 && z < g.grid[2]));
  const int64_t row = gx * (y + gy * z);
  b = in ? start[row + xl] : 0;
  e = in ? start[row + xh + 1] : 0;
}

// Element of a per-particle array at a 32-bit byte offset: the load then uses
// the scalar base + 32-bit vector offset addressing form (no 64-bit address
// arithmetic per partner).  Arrays stay below 4 GiB (the problem size limit
// checked at construction).
template <typename T>
__device__ __forceinline__ T ldo(const T* base, uint32_t byte_off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Record b of a partner-record array (n particles) as 16-byte loads: word i at
// word i·n + b where RecordWords<T> says the array keeps words apart (2D),
// else at word b·(sizeof(T)/16) + i.  Offsets are 64-bit (the arrays are larger
// than the 8-byte elements the 32-bit offsets above are sized for); b, a
// partner index in [0, n), is taken unsigned (no sign extension per load).
template <typename T>
__device__ __forceinline__ T ld_rec(const void* base, uint32_t b, int64_t n) {
  static_assert(sizeof(T) % 16 == 0, "records are whole 16-byte words");
  constexpr int kW = int(sizeof(T) / 16);
  constexpr bool apart = RecordWords<T>::value;
  const uint4* p = reinterpret_cast<const uint4*>(base) + (apart ? uint64_t(b) : uint64_t(b) * kW);
  uint4 w[kW];
#pragma unroll
  for (int i = 0; i < kW; ++i) w[i] = p[apart ? i * n : i];
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

template <typename T>
__device__ __forceinline__ void st_rec(void* base, int64_t a, const T& r, int64_t n) {
  constexpr int kW = int(sizeof(T) / 16);
  constexpr bool apart = RecordWords<T>::value;
  uint4 w[kW];
  __builtin_memcpy(w, &r, sizeof(T));
  uint4* p = reinterpret_cast<uint4*>(base) + (apart ? a : a * kW);
#pragma unroll
  for (int i = 0; i < kW; ++i) p[apart ? i * n : i] = w[i];
}

// ---- 3D cell frames and per-cell membership windows
// 3D membership entries hold one window of at most kWin3 candidates inside
// one stencil cell: the member bits in the low kWin3 bits of the low word and
// above them the cell's offset o from the particle's own cell, coded
// (ox+1) + 3(oy+1) + 9(oz+1); the high word is the window base.  The
// FP32-split 3D pair kernels turn that code into the per-window separation
// constants of their cell frames (partner records in 3D hold positions as
// offsets from their cell's frame origin).
constexpr uint32_t kWinMask3 = (1u << kWin3) - 1u;
__device__ __forceinline__ uint32_t win_code3(int ox, int oy, int oz) {
  return uint32_t(ox + 1) + 3u * uint32_t(oy + 1) + 9u * uint32_t(oz + 1);
}

// Origin of the cell frame along axis d of grid coordinate i (0-based).
// Every writer and reader of a 3D record evaluates this same expression.
__device__ __forceinline__ double frame_origin(const SphConst& C, const CellGeom& G, int d, int64_t i) {
#pragma clang fp contract(off)
  return G.off64[d] + double(i) * C.fcs[d];
}

// Frame origins of sorted cell c.
__device__ __forceinline__ void cell_frame3(const SphConst& C, const CellGeom& G, int32_t c, double fr[3]) {
  int64_t ix, iy, iz;
  cell_xyz<3>(G, c, ix, iy, iz);
  fr[0] = frame_origin(C, G, 0, ix);
  fr[1] = frame_origin(C, G, 1, iy);
  fr[2] = frame_origin(C, G, 2, iz);
}

// Separation constants of a particle at xa in sorted cell c: km[d] = xa_d -
// frame_d(c_d - 1), rounded once, and the frame cell sizes cf (FP32-exact).
// For a partner b in the stencil cell at offset o, xa - xb = K - xo_b with
// K = km - (o + 1)·cf, one FMA whose product is exact.
struct FrameK3 {
  float km[3];
  float cf[3];
};
__device__ __forceinline__ FrameK3 frame_k3(const SphConst& C, const CellGeom& G, int32_t c, const double* xa) {
  int64_t i[3];
  cell_xyz<3>(G, c, i[0], i[1], i[2]);
  FrameK3 t;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    t.km[d] = float(xa[d] - frame_origin(C, G, d, i[d] - 1));
    t.cf[d] = C.f_fcs[d];
  }
  return t;
}
// Window constants of one partner (3D FP32-split); an empty type elsewhere.
struct Win3 {
  float k[3];
};
struct NoWin {};

// Position fields of the FP32-split partner records: FP64 in 2D; in 3D the
// FP32 offset from the frame origins fr of the particle's sorted cell.
template <int D>
__device__ __forceinline__ void put_x(PartnerN<D>& r, int d, double x, const double*) {
  r.x[d] = x;
}
__device__ __forceinline__ void put_x(PartnerN<3>& r, int d, double x, const double* fr) {
  r.xo[d] = float(x - fr[d]);
}
template <int D>
__device__ __forceinline__ void put_x(PartnerN64<D>& r, int d, double x, const double*) {
  r.x[d] = x;
}
template <int D>
__device__ __forceinline__ void put_x(PartnerH<D>& r, int d, double x, const double*) {
  r.xh[d] = x;
}
__device__ __forceinline__ void put_x(PartnerH<3>& r, int d, double x, const double* fr) {
  r.xo[d] = float(x - fr[d]);
}
// FP32 separation x_a - x_b along d from partner b's record: the FP64
// difference rounded once (2D), or the window constant of b's cell minus b's
// frame offset (3D).
template <int D>
__device__ __forceinline__ float sep_x(const double* xa, const PartnerN<D>& q, int d) {
  return float(xa[d] - q.x[d]);
}
__device__ __forceinline__ float sep_x(const double*, const PartnerN<3>& q, int d, const Win3& w) {
  return w.k[d] - q.xo[d];
}
template <int D>
__device__ __forceinline__ float sep_x(const double* xa, const PartnerH<D>& q, int d) {
  return float(xa[d] - q.xh[d]);
}
__device__ __forceinline__ float sep_x(const double*, const PartnerH<3>& q, int d, const Win3& w) {
  return w.k[d] - q.xo[d];
}
// A partner's type as the pair kernels read it, and its pressure.  2D
// records keep the pressure in the type's place and the type's class in the
// two lowest mantissa bits of the FP64 y, read back as the representative
// type -1, 0, 1 or 2 (the same answer to every test a pair term makes: > 0,
// > 1, < 0, <= 1).  Those bits move y by at most 3 ulp of FP64, far below the
// FP32 rounding of a separation; the membership store reads the position
// arrays, not the records.
__device__ __forceinline__ double with_class(double y, int32_t pt) {
  const uint32_t c = pt < 0 ? 0u : (pt == 0 ? 1u : (pt == 1 ? 2u : 3u));
  const uint64_t u = uint64_t(__double_as_longlong(y));
  return __longlong_as_double((long long)((u & ~uint64_t(3)) | c));
}
__device__ __forceinline__ int32_t class_pt(double y) { return int32_t(uint32_t(__double_as_longlong(y)) & 3u) - 1; }
template <int D>
__device__ __forceinline__ int32_t rec_pt(const PartnerN<D>& r) {
  return class_pt(r.x[1]);
}
__device__ __forceinline__ int32_t rec_pt(const PartnerN<3>& r) { return r.pt; }
template <int D>
__device__ __forceinline__ int32_t rec_pt(const PartnerH<D>& r) {
  return class_pt(r.xh[1]);
}
__device__ __forceinline__ int32_t rec_pt(const PartnerH<3>& r) { return r.pt; }
template <int D>
__device__ __forceinline__ float rec_P(const PartnerN<D>& r, const SphConst&) {
  return r.P;
}
__device__ __forceinline__ float rec_P(const PartnerN<3>& r, const SphConst& C) { return eos_f(r.rho, r.pt, C); }
template <int D>
__device__ __forceinline__ float rec_P(const PartnerH<D>& r, const SphConst&) {
  return r.P;
}
__device__ __forceinline__ float rec_P(const PartnerH<3>& r, const SphConst& C) { return eos_f(r.rh, r.pt, C); }
// Type and pressure into a record whose position is already in place.
template <int D>
__device__ __forceinline__ void set_pt(PartnerN<D>& r, int32_t pt, float P) {
  r.x[1] = with_class(r.x[1], pt);
  r.P = P;
}
__device__ __forceinline__ void set_pt(PartnerN<3>& r, int32_t pt, float) { r.pt = pt; }
template <int D>
__device__ __forceinline__ void set_pt(PartnerH<D>& r, int32_t pt, float P) {
  r.xh[1] = with_class(r.xh[1], pt);
  r.P = P;
}
__device__ __forceinline__ void set_pt(PartnerH<3>& r, int32_t pt, float) { r.pt = pt; }

// Frame origins of particle a (3D; nothing to compute in 2D).
template <int D, typename R>
__device__ __forceinline__ void particle_frame(const SphBufs<R>& B, const SphConst& C, const CellGeom& G, int64_t a,
                                               double fr[3]) {
  if constexpr (D == 3) cell_frame3(C, G, B.skey[a], fr);
  else fr[0] = fr[1] = fr[2] = 0.0;
}
// Separation constants of particle a at position xa (3D FP32-split).
template <int D, typename R>
__device__ __forceinline__ FrameK3 particle_k(const SphBufs<R>& B, const SphConst& C, const CellGeom& G, int64_t a,
                                              const double* xa) {
  if constexpr (D == 3) return frame_k3(C, G, B.skey[a], xa);
  else return FrameK3{};
}

// Neighbour-list partners b of particle a in a fixed order (stencil rows in
// order, ascending sorted index inside a row; in 3D each row visited cell by
// cell, which is the same order).  nc >= 0: the nc stored entries of the
// membership store, read one ahead so the next load overlaps the current
// window's pair bodies.  nc < 0: the windows are computed here from the
// membership positions XM.  Each lane runs its own sequence, so all lanes of a
// wave execute the (long) pair body together.  Fr (3D only): the iterator also
// keeps the current window's separation constants from the particle's FrameK3.
template <int D, typename R, bool Fr = false>
struct Partners {
  static_assert(!Fr || cell_windows<D, R>(), "cell frames come with cell windows");
  const SphBufs<R>& B;
  const CellGeom& G;
  const double* const xm0;  // membership positions, held by value
  const double* const xm1;
  const double* const xm2;
  const double dist2;
  const int64_t a;
  const int nc;
  const uint64_t* pe;
  uint64_t nxt;
  int e = 1, k = -1;
  int32_t w0 = 0, re = 0, cb = 0;
  uint32_t c = 0;
  FrameK3 kt;        // Fr: the particle's separation constants
  float wk[3];       // Fr: constants of the current window
  int j = 3;         // 3D scan (nc < 0): cell of row k, 0..2 for ox = -1..1

  __device__ __forceinline__ Partners(const SphBufs<R>& B_, int64_t a_, int nc_, const CellGeom& G_,
                                      const double* const* XM_, double dist2_, const FrameK3* kt_ = nullptr)
      : B(B_),
        G(G_),
        xm0(XM_[0]),
        xm1(XM_[1]),
        xm2(XM_[2]),
        dist2(dist2_),
        a(a_),
        nc(nc_),
        pe(B_.NE + a_),
        nxt(nc_ > 0 ? *pe : 0ull) {
    if constexpr (Fr) kt = *kt_;
    fill();
  }

  // Another partner is waiting (the iterator always stands on the next one).
  __device__ __forceinline__ bool has() const { return c != 0; }

  // The waiting partner; moves on to the one after it.
  __device__ __forceinline__ int32_t take() {
    const int32_t b = cb + __builtin_ctz(c);
    c &= c - 1;
    if (c == 0) fill();
    return b;
  }
  // As take(), with the separation constants of the partner's window.
  __device__ __forceinline__ int32_t take(Win3& w) {
    w.k[0] = wk[0];
    w.k[1] = wk[1];
    w.k[2] = wk[2];
    return take();
  }
  __device__ __forceinline__ int32_t take(NoWin&) { return take(); }

  __device__ __forceinline__ void set_window(uint32_t code) {
    if constexpr (Fr) {
      // base-3 digits of code < 27: q / 3 = (q · 11) >> 5 for q < 27
      const uint32_t q3 = (code * 11u) >> 5, q9 = (q3 * 11u) >> 5;
      const uint32_t o1[3] = {code - 3u * q3, q3 - 3u * q9, q9};
#pragma unroll
      for (int d = 0; d < 3; ++d) wk[d] = fmaf(float(o1[d]), -kt.cf[d], kt.km[d]);
    }
  }

  // c == 0: loads the next window that has members (c stays 0 at the end).
  __device__ __forceinline__ void fill() {
    if (nc >= 0) {
      if (nxt == 0) return;  // stored entries always have member bits
      if constexpr (cell_windows<D, R>()) {
        c = uint32_t(nxt) & kWinMask3;
        set_window(uint32_t(nxt) >> kWin3);
      } else {
        c = uint32_t(nxt);
      }
      cb = int32_t(nxt >> 32);
      pe += B.n;
      nxt = e < nc ? *pe : 0ull;
      ++e;
      return;
    }
    if constexpr (cell_windows<D, R>()) {
      // stencil cell by cell, windows of kWin3 candidates inside a cell
      while (c == 0) {
        if (w0 >= re) {
          if (++j > 2) {
            j = 0;
            if (++k >= Rows<3>::N) return;
          }
          int64_t cx, cy, cz;
          cell_xyz<3>(G, B.skey[a], cx, cy, cz);
          const int64_t x = cx + j - 1, y = cy + (k % 3) - 1, z = cz + (k / 3) - 1;
          const bool in = x >= 0 && x < G.grid[0] && y >= 0 && y < G.grid[1] && z >= 0 && z < G.grid[2];
          const int64_t cell = x + G.grid[0] * (y + G.grid[1] * z);
          w0 = in ? B.start[cell] : 0;
          re = in ? B.start[cell + 1] : 0;
          set_window(win_code3(j - 1, (k % 3) - 1, (k / 3) - 1));
          continue;
        }
        const int32_t hi = re < w0 + kWin3 ? re : w0 + kWin3;
        c = window_bits_lean<3>(xm0, xm1, xm2, xm0[a], xm1[a], xm2[a], a, w0, hi, dist2);
        cb = w0;
        w0 = hi;
      }
    } else {
      while (c == 0) {
        if (w0 >= re) {
          if (++k >= Rows<D>::N) return;
          row_range<D>(G, B.start, B.skey[a], k, w0, re);
          continue;
        }
        const int32_t hi = re < w0 + 32 ? re : w0 + 32;
        c = window_bits_lean<D>(xm0, xm1, xm2, xm0[a], xm1[a], D == 3 ? xm2[a] : 0.0, a, w0, hi, dist2);
        cb = w0;
        w0 = hi;
      }
    }
  }
};

// The partner loops below leave together: a lane without partners left sits
// out the remaining trips (masked off), and the loop ends when no lane of the
// wave has one.  A loop whose lanes leave one by one makes the compiler carry
// every loop value in a second set of registers with copies on each trip.
template <int D, typename R, typename F>
__device__ __forceinline__ void for_each_partner(const SphBufs<R>& B, int64_t a, int nc, const CellGeom& G,
                                                 const double* const* XM, double dist2, F&& body) {
  Partners<D, R> it(B, a, nc, G, XM, dist2);
  for (;;) {
    const bool m = it.has();
    if (!__any(m)) break;
    if (m) body(it.take());
  }
}

// As for_each_partner, with the partner data split off: fetch(b) loads what
// the body needs of partner b.  Pipe: the fetch of the next partner is issued
// before the body of the current one runs, so the loads overlap the pair
// arithmetic (at the cost of the registers holding both partners' data).
// Frames (3D FP32-split readers of partner records): kt holds the particle's
// separation constants and the body receives the partner's window constants
// as a third argument.
template <bool Pipe, int D, typename R, bool Frames = false, typename Fetch, typename F>
__device__ __forceinline__ void for_each_partner_fetch(const SphBufs<R>& B, int64_t a, int nc, const CellGeom& G,
                                                       const double* const* XM, double dist2, Fetch&& fetch,
                                                       F&& body, const FrameK3* kt = nullptr) {
  constexpr bool Fr = Frames && D == 3 && std::is_same<R, float>::value;
  using W = typename std::conditional<Fr, Win3, NoWin>::type;
  Partners<D, R, Fr> it(B, a, nc, G, XM, dist2, kt);
  auto run = [&](const int32_t b, const auto& f, const W& w) {
    if constexpr (Fr) body(b, f, w);
    else body(b, f);
  };
  if constexpr (!Pipe) {
    for (;;) {
      const bool m = it.has();
      if (!__any(m)) break;
      if (m) {
        W w;
        const int32_t b = it.take(w);
        run(b, fetch(b), w);
      }
    }
    return;
  }
  // two partners in flight, alternating between two sets of registers; a lane
  // without a next partner fetches its previous one (or itself) again, which
  // is a valid address, and skips the body
  bool m0 = it.has();
  if (!__any(m0)) return;
  W w0{}, w1{};
  int32_t b0 = m0 ? it.take(w0) : int32_t(a);
  auto f0 = fetch(b0);
  for (;;) {
    const bool m1 = it.has();
    const int32_t b1 = m1 ? it.take(w1) : b0;
    auto f1 = fetch(b1);
    if (m0) run(b0, f0, w0);
    if (!__any(m1)) break;
    m0 = it.has();
    b0 = m0 ? it.take(w0) : b1;
    f0 = fetch(b0);
    if (m1) run(b1, f1, w1);
    if (!__any(m0)) break;
  }
}

__device__ __forceinline__ double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }

constexpr int kRedSlots = 64;

// Block maximum (NaN wins) folded into one of kRedSlots slots by an atomic max
// on its bit pattern (non-negative doubles order like their bits, NaN above
// +Inf); spreading blocks over many slots avoids one hot address.  Every
// thread of the block must call it.
template <int BS = kBlock>
__device__ __forceinline__ void block_max(double v, bool active, double* slots, int blk) {
  __shared__ double wm[BS / 64];
  double m = wave_max(active ? v : 0.0);
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = wm[0];
#pragma unroll
    for (int w = 1; w < BS / 64; ++w) r = nanmax(r, wm[w]);
    atomic_max_nonneg(reinterpret_cast<unsigned long long*>(slots) + (blk % kRedSlots), r);
  }
  __syncthreads();
}

// Calls f with the scene source of L: the descriptor by value for one problem,
// the device descriptor array and maps for several.
template <typename R, typename Fn>
inline void with_scenes(const LaunchSet& L, Fn&& f) {
  if (L.one) f(OneScene<R>{*static_cast<const SceneDev<R>*>(L.one)});
  else f(ManyScenes<R>{static_cast<const SceneDev<R>*>(L.scenes), L.pmap, L.cmap, L.pmap512});
}

}  // namespace gclsph
