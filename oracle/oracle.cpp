// CPU restatement of GPUCellListSPH.jl's hot path: the cell list
// (`GPUCellList`, `update!`, `partialupdate!`) and one weakly-compressible SPH
// step (`_stepsolve!`).  It follows the reference kernel by kernel, in pair-list
// order, so every per-pair expression is evaluated exactly as the Julia source
// writes it (left-to-right, no FMA contraction: build with -ffp-contract=off).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/,
// by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py.  No
// product code includes, links or calls it.
//
// Where the reference's GPU run is nondeterministic (slot order in
// `fillcells_naive!`, pair order from the atomic cursor, FP64 atomic
// accumulation order) this restatement picks one deterministic order:
// particles in index order, cells in column-major (x fastest) order, pairs in
// emission order, and accumulation in pair order.
//
// Third-party arithmetic that is not in the reference tree:
//  * SPHKernels.jl (compat "2.2", Project.toml:14,24; no Manifest):
//    WendlandC2 W(u) = norm*hinv^D*(1-u)^4*(1+4u), dW(u) = -20u(1-u)^3*norm*
//    hinv^(D+1) for u<1, norm = 7/pi (2D) or 21/(2pi) (3D).  The expression
//    order inside is not visible here; the one below is this file's choice.
//  * Roots.jl find_zero (src/sphproblem.jl:218): replaced by bisection to the
//    last bit, the root is unique on [0,1].
//  * Julia's Float64^Int (pressure `(rho/rho0)^gamma`, fbmolforce `^12`,
//    `^4`): evaluated here by binary powering (x2=x*x, ...).  CUDA.jl's device
//    power differs at the ulp level; that is parity unpinned at the ulp level.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return -1;
}

// ---------------------------------------------------------------------------
// Cell list (src/celllist.jl, src/gpukernels_cls.jl)
// ---------------------------------------------------------------------------
template <typename T>
struct CellList {
  int D = 2;
  int64_t n = 0;
  T dist = 0;
  T cs[3] = {0, 0, 0};
  T off[3] = {0, 0, 0};       // offset stored in T (struct field conversion)
  double off64[3] = {0, 0, 0};  // offset as computed in Float64 (ctor path)
  int64_t grid[3] = {1, 1, 1};
  std::vector<T> x[3];
  std::vector<int32_t> pcell[3];  // 1-based cell coordinates
  std::vector<int32_t> cellpnum;  // column-major over grid
  int64_t mppcell = 0;
  std::vector<int32_t> celllist;  // mppcell x ncells
  std::vector<int32_t> pairs;     // 2 * capacity entries, 1-based, (0,0) pad
  int64_t capacity = 0;
  int64_t pairsn = 0;

  int64_t ncells() const { return grid[0] * grid[1] * (D == 3 ? grid[2] : 1); }
  int64_t lin(int64_t ix, int64_t iy, int64_t iz) const {  // 1-based in
    return (ix - 1) + grid[0] * ((iy - 1) + grid[1] * (iz - 1));
  }
};

// kernel_cellmap! (src/gpukernels_cls.jl:29-60); hinv = 1/h on the host in T
// (:21).  `use_off64` reproduces the constructor call (src/celllist.jl:55),
// which passes the Float64 MIN before it is stored in T (matters for T=Float32).
template <typename T>
void cellmap(CellList<T>& c, bool use_off64) {
  T hinv[3];
  for (int k = 0; k < c.D; ++k) hinv[k] = T(1) / c.cs[k];
  std::fill(c.cellpnum.begin(), c.cellpnum.end(), 0);
  for (int64_t i = 0; i < c.n; ++i) {
    int32_t idx[3] = {1, 1, 1};
    for (int k = 0; k < c.D; ++k) {
      double p;
      if (use_off64) {
        p = (double(c.x[k][i]) - c.off64[k]) * double(hinv[k]);
      } else {
        T pt = (c.x[k][i] - c.off[k]) * hinv[k];
        p = double(pt);
      }
      // min(max(p, 1), grid); a NaN coordinate goes to cell 1, as the GPU
      // library documents (the reference leaves it undefined: ceil(Int32, NaN))
      double cl = p >= 1.0 ? p : 1.0;
      cl = cl <= double(c.grid[k]) ? cl : double(c.grid[k]);
      idx[k] = int32_t(std::ceil(cl));
      c.pcell[k][i] = idx[k];
    }
    c.cellpnum[c.lin(idx[0], idx[1], c.D == 3 ? idx[2] : 1)] += 1;
  }
}

// kernel_fillcells_naive! (src/gpukernels_cls.jl:77-101), slots in index order.
template <typename T>
void fillcells(CellList<T>& c) {
  std::fill(c.cellpnum.begin(), c.cellpnum.end(), 0);
  for (int64_t i = 0; i < c.n; ++i) {
    int64_t cell = c.lin(c.pcell[0][i], c.pcell[1][i], c.D == 3 ? c.pcell[2][i] : 1);
    int32_t slot = ++c.cellpnum[cell];
    if (slot <= c.mppcell) c.celllist[(slot - 1) + c.mppcell * cell] = int32_t(i + 1);
  }
}

// мaxpairs_2d / _3d (src/gpukernels_cls.jl:116-286)
template <typename T>
int64_t maxpairs(const CellList<T>& c) {
  static const int off2[4][2] = {{-1, 1}, {0, 1}, {1, 1}, {1, 0}};
  static const int off3[13][3] = {{-1, 1, 0}, {0, 1, 0},  {1, 1, 0},  {1, 0, 0},  {-1, 1, 1},
                                  {0, 1, 1},  {1, 1, 1},  {1, 0, 1},  {0, 0, 1},  {-1, 1, -1},
                                  {0, 1, -1}, {1, 1, -1}, {1, 0, -1}};
  int64_t total = 0;
  int64_t Nx = c.grid[0], Ny = c.grid[1], Nz = c.D == 3 ? c.grid[2] : 1;
  for (int64_t iz = 1; iz <= Nz; ++iz)
    for (int64_t iy = 1; iy <= Ny; ++iy)
      for (int64_t ix = 1; ix <= Nx; ++ix) {
        int32_t nn = c.cellpnum[c.lin(ix, iy, iz)];
        if (nn <= 0) continue;
        int64_t m = 0;
        int no = c.D == 2 ? 4 : 13;
        for (int o = 0; o < no; ++o) {
          int64_t jx = ix + (c.D == 2 ? off2[o][0] : off3[o][0]);
          int64_t jy = iy + (c.D == 2 ? off2[o][1] : off3[o][1]);
          int64_t jz = iz + (c.D == 2 ? 0 : off3[o][2]);
          if (jx > 0 && jx <= Nx && jy > 0 && jy <= Ny && jz > 0 && jz <= Nz)
            m += c.cellpnum[c.lin(jx, jy, jz)];
        }
        // Int((n * (n - 1)) * 0.5) + m * n  (:162, :280)
        int64_t val = int64_t(double(int32_t(nn * (nn - 1))) * 0.5) + m * int64_t(nn);
        total += val;
      }
  return total;
}

// r^2 exactly as the neighbour kernels write it: (a1-b1)^2 + (a2-b2)^2 [+ (a3-b3)^2]
// (src/gpukernels_cls.jl:326,390,456,527); Julia's x^2 is x*x.
template <typename T>
inline T dist2(const CellList<T>& c, int64_t i, int64_t j) {
  T d0 = c.x[0][i] - c.x[0][j];
  T d1 = c.x[1][i] - c.x[1][j];
  T s = d0 * d0 + d1 * d1;
  if (c.D == 3) {
    T d2 = c.x[2][i] - c.x[2][j];
    s = s + d2 * d2;
  }
  return s;
}

template <typename T>
struct Emitter {
  CellList<T>& c;
  int64_t cnt = 0;
  void push(int32_t a, int32_t b) {
    int32_t lo = std::min(a, b), hi = std::max(a, b);  // minmax(indi, indj)
    if (cnt < c.capacity) {
      c.pairs[2 * cnt] = lo;
      c.pairs[2 * cnt + 1] = hi;
    }
    ++cnt;
  }
};

// neib_search! (src/gpukernels_cls.jl:561-587): internal pass
// (kernel_neib_internal!, :309-354 / :372-416) then the half-stencil external
// passes (kernel_neib_external!, :439-486 / :505-554) in launch order.
template <typename T>
int64_t neib_search(CellList<T>& c) {
  static const int ext2[4][2] = {{1, -1}, {0, 1}, {1, 1}, {1, 0}};
  static const int ext3[13][3] = {{1, -1, 0},  {0, 1, 0},  {1, 1, 0},  {1, 0, 0},  {1, -1, 1},
                                  {0, 1, 1},   {1, 1, 1},  {1, 0, 1},  {0, 0, 1},  {1, -1, -1},
                                  {0, 1, -1},  {1, 1, -1}, {1, 0, -1}};
  Emitter<T> em{c};
  T dist2max = c.dist * c.dist;  // dist^2
  int64_t Nx = c.grid[0], Ny = c.grid[1], Nz = c.D == 3 ? c.grid[2] : 1;
  // internal
  for (int64_t iz = 1; iz <= Nz; ++iz)
    for (int64_t iy = 1; iy <= Ny; ++iy)
      for (int64_t ix = 1; ix <= Nx; ++ix) {
        int64_t cell = c.lin(ix, iy, iz);
        int32_t len = c.cellpnum[cell];
        if (len <= 1) continue;
        for (int32_t a = 1; a <= len - 1; ++a) {
          int32_t ia = c.celllist[(a - 1) + c.mppcell * cell];
          for (int32_t b = a + 1; b <= len; ++b) {
            int32_t ib = c.celllist[(b - 1) + c.mppcell * cell];
            if (dist2(c, ia - 1, ib - 1) < dist2max) em.push(ia, ib);
          }
        }
      }
  int no = c.D == 2 ? 4 : 13;
  for (int o = 0; o < no; ++o) {
    int ox = c.D == 2 ? ext2[o][0] : ext3[o][0];
    int oy = c.D == 2 ? ext2[o][1] : ext3[o][1];
    int oz = c.D == 2 ? 0 : ext3[o][2];
    for (int64_t iz = 1; iz <= Nz; ++iz)
      for (int64_t iy = 1; iy <= Ny; ++iy)
        for (int64_t ix = 1; ix <= Nx; ++ix) {
          int64_t jx = ix + ox, jy = iy + oy, jz = iz + oz;
          if (!(jx > 0 && jx <= Nx && jy > 0 && jy <= Ny && jz > 0 && jz <= Nz)) continue;
          int64_t cell = c.lin(ix, iy, iz), ncell = c.lin(jx, jy, jz);
          int32_t li = c.cellpnum[cell], lj = c.cellpnum[ncell];
          if (li <= 0) continue;
          for (int32_t a = 1; a <= li; ++a) {
            int32_t ia = c.celllist[(a - 1) + c.mppcell * cell];
            for (int32_t b = 1; b <= lj; ++b) {
              int32_t ib = c.celllist[(b - 1) + c.mppcell * ncell];
              if (dist2(c, ia - 1, ib - 1) < dist2max) em.push(ia, ib);
            }
          }
        }
  }
  return em.cnt;
}

template <typename T>
void resize_pairs(CellList<T>& c, int64_t cap) {
  c.capacity = cap;
  c.pairs.assign(size_t(2 * cap), 0);
}

// GPUCellList constructor (src/celllist.jl:22-88)
template <typename T>
int celllist_create(CellList<T>& c, int D, int64_t n, const double* const* pts, const double* cellsize,
                    double dist, const double* MINin, const double* MAXin, int64_t mppcell, int64_t mpairs,
                    int* warn_flags) {
  if (D < 2 || D > 3) return fail("wrong dimention");
  if (n <= 0) return fail("empty point set");
  c.D = D;
  c.n = n;
  c.dist = T(dist);
  for (int k = 0; k < D; ++k) {
    c.x[k].assign(pts[k], pts[k] + n);
    c.pcell[k].assign(size_t(n), 0);
  }
  *warn_flags = 0;
  for (int k = 0; k < D; ++k) {
    c.cs[k] = T(cellsize[k]);
    if (c.cs[k] < c.dist) {  // :32-39
      c.cs[k] = c.dist;
      *warn_flags |= 1;
    }
  }
  const double sq = std::sqrt(std::numeric_limits<double>::epsilon());  // sqrt(eps())
  double range[3];
  for (int k = 0; k < D; ++k) {
    double mn, mx;
    if (MINin) mn = MINin[k];
    else mn = double(*std::min_element(c.x[k].begin(), c.x[k].end()));
    // MIN = @. MIN - abs((MIN + sqrt(eps())) * sqrt(eps()))  (:42)
    mn = mn - std::fabs((mn + sq) * sq);
    if (MAXin) mx = MAXin[k];
    else mx = double(*std::max_element(c.x[k].begin(), c.x[k].end()));
    c.off64[k] = mn;
    c.off[k] = T(mn);
    range[k] = mx - mn;
    double cells = std::ceil(range[k] / double(c.cs[k]));  // :46
    if (!(cells >= 1)) return fail("cell grid has no cells along a dimension");
    c.grid[k] = int64_t(cells);
  }
  c.cellpnum.assign(size_t(c.ncells()), 0);
  cellmap(c, true);
  int32_t mx = *std::max_element(c.cellpnum.begin(), c.cellpnum.end());
  int64_t maxpoint = int64_t(std::ceil(double(mx) * 1.05 + 1));  // :57
  if (mppcell < maxpoint) mppcell = maxpoint;
  c.mppcell = mppcell;
  c.celllist.assign(size_t(c.mppcell * c.ncells()), 0);
  fillcells(c);
  if (mpairs == 0) mpairs = maxpairs(c);
  resize_pairs(c, mpairs);
  int64_t total = neib_search(c);
  if (total > mpairs) *warn_flags |= 2;  // :77-79
  if (double(total) < double(mpairs) * 0.8) {  // :80-86
    int64_t newn = int64_t(std::ceil(double(total) / 0.8));
    std::vector<int32_t> keep(c.pairs.begin(), c.pairs.begin() + 2 * newn);
    c.capacity = newn;
    c.pairs.swap(keep);
  }
  c.pairsn = total;
  return 0;
}

// update! (src/celllist.jl:95-134)
template <typename T>
void celllist_update(CellList<T>& c, bool fillzero) {
  cellmap(c, false);
  int32_t maxpoint = *std::max_element(c.cellpnum.begin(), c.cellpnum.end());
  if (maxpoint > c.mppcell) {
    c.mppcell = int64_t(std::ceil(double(maxpoint) * 1.05 + 1));
    c.celllist.assign(size_t(c.mppcell * c.ncells()), 0);
  } else {
    std::fill(c.celllist.begin(), c.celllist.end(), 0);
  }
  fillcells(c);
  if (c.pairsn > c.capacity || double(c.pairsn) < double(c.capacity) * 0.6) {
    resize_pairs(c, int64_t(std::ceil(double(c.pairsn) / 0.8)));
  } else if (fillzero) {
    std::fill(c.pairs.begin(), c.pairs.end(), 0);
  }
  c.pairsn = neib_search(c);
  if (c.pairsn > c.capacity) {
    resize_pairs(c, int64_t(std::ceil(double(c.pairsn) / 0.8)));
    c.pairsn = neib_search(c);
  }
}

// partialupdate! (src/celllist.jl:141-145)
template <typename T>
void celllist_partialupdate(CellList<T>& c, bool fillzero) {
  if (fillzero) std::fill(c.pairs.begin(), c.pairs.end(), 0);
  c.pairsn = neib_search(c);
}

// ---------------------------------------------------------------------------
// SPH kernel functions (SPHKernels.jl WendlandC2; call sites listed above)
// ---------------------------------------------------------------------------
struct Wendland {
  int D;
  double norm;
  explicit Wendland(int d) : D(d), norm(d == 2 ? 7.0 / M_PI : 21.0 / (2.0 * M_PI)) {}
  static double ipow(double x, int p) {
    double r = x;
    for (int k = 1; k < p; ++k) r = r * x;
    return r;
  }
  double W(double u, double hinv) const {
    double nrm = norm * ipow(hinv, D);
    if (u < 1.0) {
      double um1 = 1.0 - u;
      um1 = um1 * um1;
      um1 = um1 * um1;
      return ((1.0 + 4.0 * u) * um1) * nrm;
    }
    return 0.0;
  }
  double dW(double u, double hinv) const {
    double nrm = norm * ipow(hinv, D + 1);
    if (u < 1.0) {
      double um1 = 1.0 - u;
      double um13 = um1 * um1 * um1;
      return -20.0 * u * um13 * nrm;
    }
    return 0.0;
  }
};

// powfancy7th (src/gpukernels_sph.jl:499-515)
double powfancy7th(double x, double ginv, int gamma) {
  if (gamma == 7) {
    uint64_t bits;
    double ax = std::fabs(x);
    std::memcpy(&bits, &ax, 8);
    uint64_t tb = 0x36cd000000000000ULL + bits / 7;
    double t;
    std::memcpy(&t, &tb, 8);
    t = std::copysign(t, x);
    for (int it = 0; it < 2; ++it) {
      double t2 = t * t;
      double t3 = t2 * t;
      double t4 = t2 * t2;
      double xot4 = x / t4;
      t = t - t * (t3 - xot4) / (4.0 * t3 + 3.0 * xot4);
    }
    return t;
  }
  return std::pow(x, ginv);
}

// Float64^Int by binary powering (see header note).
double powi(double x, int p) {
  if (p == 0) return 1.0;
  if (p < 0) return 1.0 / powi(x, -p);
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

// ---------------------------------------------------------------------------
// SPH problem (src/sphproblem.jl)
// ---------------------------------------------------------------------------
struct Params {
  double dx, h, H, rho0, m0, dt, alpha, c0;
  int64_t gamma;
  double delta, CFL, g, s, nu, dpc_l0, dpc_pmin, dpc_pmax, dpc_lambda, xsph_eps;
  int64_t cspmn;
  double bound_D, bound_l, nui, etime;
};

struct Sph {
  CellList<double>* sys = nullptr;
  int D = 2;
  int64_t N = 0;
  Params p{};
  double hinv = 0, Hinv = 0;
  Wendland kern{2};
  std::vector<double> rho, rhoh, P, buf, drho, sumW, W;
  std::vector<int32_t> ptype;
  std::vector<double> v[3], vh[3], xh[3], acc[3], buf2[3], cdx[3], sumgW[3], gW[3];
  std::vector<double> dvdpc[3];  // Δv_dpc of the last dpcreg! (buf2 is reused by XSPH)
  // pairs view: first pairsn entries (or whole buffer before first update)
  const int32_t* pr = nullptr;
  int64_t npr = 0;
};

using Vec = double[3];

inline void getv(const std::vector<double>* a, int D, int64_t i, double* o) {
  for (int k = 0; k < D; ++k) o[k] = a[k][i];
}
inline double dotv(const double* a, const double* b, int D) {  // StaticArrays dot
  double s = a[0] * b[0] + a[1] * b[1];
  if (D == 3) s = s + a[2] * b[2];
  return s;
}

// sphW! + sph∇W! (src/gpukernels_sph.jl:84-100, :196-213)
void sphW_gradW(Sph& S, const std::vector<double>* x) {
  int D = S.D;
  S.W.assign(size_t(S.npr), 0.0);
  for (int k = 0; k < D; ++k) S.gW[k].assign(size_t(S.npr), 0.0);
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (i == 0) continue;
    --i;
    --j;
    Vec d{};
    for (int k = 0; k < D; ++k) d[k] = x[k][i] - x[k][j];
    double r = std::sqrt(dotv(d, d, D));  // norm(Δx)
    double u = r * S.Hinv;
    S.W[q] = S.kern.W(u, S.Hinv);
    double dwk_r = S.kern.dW(u, S.Hinv) / r;
    for (int k = 0; k < D; ++k) S.gW[k][q] = d[k] * dwk_r;
  }
}

// ∂ρ∂tDDT! (src/gpukernels_sph.jl:542-631)
void ddt(Sph& S, const std::vector<double>& rho, const std::vector<double>* v, const std::vector<double>* x) {
  int D = S.D;
  const Params& p = S.p;
  std::fill(S.drho.begin(), S.drho.end(), 0.0);
  double eta2 = (0.1 * p.h) * (0.1 * p.h);
  double ginv = 1.0 / double(p.gamma);
  double DDTkh = 2 * p.h * p.delta * p.c0;
  double Cb = (p.c0 * p.c0 * p.rho0) * ginv;
  double DDTgz = p.rho0 * p.g / Cb;
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (!(i > 0)) continue;
    --i;
    --j;
    if (!(S.ptype[i] > 0 && S.ptype[j] > 0)) continue;
    Vec dx{}, dv{}, gw{};
    for (int k = 0; k < D; ++k) {
      dx[k] = x[k][i] - x[k][j];
      dv[k] = v[k][i] - v[k][j];
      gw[k] = S.gW[k][q];
    }
    double r2 = dotv(dx, dx, D);
    double ri = rho[i], rj = rho[j];
    double dot3 = -dotv(dx, gw, D);
    double m0dot = p.m0 * dotv(dv, gw, D);
    double si = m0dot, sj = m0dot;
    if (S.ptype[i] > 1) {
      double drhopvp = p.rho0 * powfancy7th(1 + DDTgz * dx[1], ginv, int(p.gamma)) - p.rho0;
      double visc_densi = DDTkh * (rj - ri - drhopvp) / (r2 + eta2);
      double delta_i = visc_densi * dot3 * p.m0 / rj;
      si += delta_i;
    }
    S.drho[i] += si;
    if (S.ptype[j] > 1) {
      double drhopvn = p.rho0 * powfancy7th(1 - DDTgz * dx[1], ginv, int(p.gamma)) - p.rho0;
      double visc_densi = DDTkh * (ri - rj - drhopvn) / (r2 + eta2);
      double delta_j = visc_densi * dot3 * p.m0 / ri;
      sj += delta_j;
    }
    S.drho[j] += sj;
  }
}

// pressure / pressure! (src/gpukernels_sph.jl:691-727)
inline double eos(double rho, int64_t gamma, double rho0, double P0, int32_t ptype) {
  if (ptype <= 1 && rho < rho0) return 0.0;
  return P0 * (powi(rho / rho0, int(gamma)) - 1);
}
void pressure(Sph& S, const std::vector<double>& rho) {
  double P0 = S.p.c0 * S.p.c0 * S.p.rho0 / double(S.p.gamma);  // c₀ ^ 2 * ρ₀ / γ
  for (int64_t i = 0; i < S.N; ++i) S.P[i] = eos(rho[i], S.p.gamma, S.p.rho0, P0, S.ptype[i]);
}

// ∂v∂t! (src/gpukernels_sph.jl:750-776)
void dvdt(Sph& S, const std::vector<double>& rho) {
  int D = S.D;
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (i == 0) continue;
    --i;
    --j;
    if (!(int64_t(S.ptype[i]) * int64_t(S.ptype[j]) > 0)) continue;
    double ri = rho[i], rj = rho[j];
    double Pfac = (S.P[i] + S.P[j]) / (ri * rj);
    for (int k = 0; k < D; ++k) {
      double a = -S.p.m0 * Pfac * S.gW[k][q];
      S.acc[k][i] += a;
      S.acc[k][j] -= a;
    }
  }
}

// ∂v∂t_av! (src/gpukernels_sph.jl:852-901)
void dvdt_av(Sph& S, const std::vector<double>* x, const std::vector<double>& rho, const std::vector<double>* v) {
  int D = S.D;
  const Params& p = S.p;
  double eta2 = (0.1 * p.h) * (0.1 * p.h);
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (i == 0) continue;
    --i;
    --j;
    if (!(S.ptype[i] > 0 && S.ptype[j] > 0)) continue;
    Vec dx{}, dv{};
    for (int k = 0; k < D; ++k) {
      dx[k] = x[k][i] - x[k][j];
      dv[k] = v[k][i] - v[k][j];
    }
    double r2 = dotv(dx, dx, D);
    double ri = rho[i], rj = rho[j];
    double rhom = (ri + rj) * 0.5;
    double cond = dotv(dv, dx, D);
    if (cond < 0) {
      double dmu = p.h * cond / (r2 + eta2);
      double dPi = (-p.alpha * p.c0 * dmu) / rhom;
      for (int k = 0; k < D; ++k) {
        double a = -dPi * p.m0 * S.gW[k][q];
        S.acc[k][i] += a;
        S.acc[k][j] -= a;
      }
    }
  }
}

// ∂v∂t_visc! (src/gpukernels_sph.jl:952-991)
void dvdt_visc(Sph& S, const std::vector<double>* v, const std::vector<double>& rho, const std::vector<double>* x) {
  int D = S.D;
  const Params& p = S.p;
  double eta2 = (0.1 * p.h) * (0.1 * p.h);
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (i == 0) continue;
    --i;
    --j;
    if (!(S.ptype[i] > 0 && S.ptype[j] > 0)) continue;
    double ri = rho[i], rj = rho[j];
    Vec dx{}, dv{}, gw{};
    for (int k = 0; k < D; ++k) {
      dx[k] = x[k][i] - x[k][j];
      dv[k] = v[k][i] - v[k][j];
      gw[k] = S.gW[k][q];
    }
    double r2 = dotv(dx, dx, D);
    double term = 4 * p.nu * p.m0 * dotv(dx, gw, D) / ((ri + rj) * (r2 + eta2));
    for (int k = 0; k < D; ++k) {
      double a = term * dv[k];
      S.acc[k][i] += a;
      S.acc[k][j] -= a;
    }
  }
}

// ∂v∂t_addgrav! + gravvec (src/gpukernels_sph.jl:1041-1051; src/sphproblem.jl:10-17)
void addgrav(Sph& S) {
  double gv[3] = {0.0, S.p.g, 0.0};
  for (int64_t i = 0; i < S.N; ++i)
    for (int k = 0; k < S.D; ++k) S.acc[k][i] -= gv[k];
}

// fbmolforce! (src/gpukernels_sph.jl:1885-1913)
void fbmolforce(Sph& S, const std::vector<double>* x, double d, double r0) {
  int D = S.D;
  double eta2 = 0.01 * (r0 * r0);
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (i == 0) continue;
    --i;
    --j;
    if (!((S.ptype[i] > 0) != (S.ptype[j] > 0))) continue;
    Vec dx{};
    for (int k = 0; k < D; ++k) dx[k] = x[k][i] - x[k][j];
    double r2 = dotv(dx, dx, D);
    double r = std::sqrt(r2);
    if (r < r0) {
      double q1 = r0 / (r + eta2);
      double Fc = d * (powi(q1, 12) - powi(q1, 4)) / (r2 + eta2);
      for (int k = 0; k < D; ++k) {
        double F = dx[k] * Fc;
        S.acc[k][i] += F;
        S.acc[k][j] -= F;
      }
    }
  }
}

// ∂v∂tpF! (src/gpukernels_sph.jl:1382-1409)
void surface_tension(Sph& S, const std::vector<double>* x) {
  int D = S.D;
  const Params& p = S.p;
  const double c15pi = 1.5 * M_PI;
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (i == 0) continue;
    --i;
    --j;
    if (!(S.ptype[i] > 0 && S.ptype[j] > 0)) continue;
    Vec dx{};
    for (int k = 0; k < D; ++k) dx[k] = x[k][i] - x[k][j];
    double r = std::sqrt(dotv(dx, dx, D));
    if (r < 2 * p.h) {
      double scos = p.s * std::cos(c15pi * r / (2 * p.h)) / (r + (0.1 * p.h));
      for (int k = 0; k < D; ++k) {
        double F = scos * dx[k] / p.m0;
        S.acc[k][i] += F;
        S.acc[k][j] -= F;
      }
    }
  }
}

// update_ρp∂ρ∂tΔt! (src/gpukernels_sph.jl:1078-1096)
void update_rho(Sph& S, std::vector<double>& rho, double dt) {
  for (int64_t i = 0; i < S.N; ++i) {
    double rv = rho[i] + S.drho[i] * dt;
    if (rv < S.p.rho0 && S.ptype[i] < 0) rv = S.p.rho0;
    rho[i] = rv;
  }
}
// update_vp∂v∂tΔt! (src/gpukernels_sph.jl:1107-1127)
void update_v(Sph& S, std::vector<double>* v, double dt) {
  for (int64_t i = 0; i < S.N; ++i)
    if (S.ptype[i] > 1)
      for (int k = 0; k < S.D; ++k) v[k][i] += S.acc[k][i] * dt;
}
// update_xpvΔt! (src/gpukernels_sph.jl:1150-1169)
void update_x(Sph& S, std::vector<double>* x, const std::vector<double>* v, double dt) {
  for (int64_t i = 0; i < S.N; ++i)
    for (int k = 0; k < S.D; ++k) x[k][i] += v[k][i] * dt;
}

// symplectic_update! (src/gpukernels_sph.jl:1193-1245)
void symplectic(Sph& S) {
  int D = S.D;
  std::vector<double>* x = S.sys->x;
  double dt = S.p.dt;
  for (int64_t i = 0; i < S.N; ++i) {
    double epsi = -(S.drho[i] / S.rhoh[i]) * dt;
    double rv = S.rho[i] * (2 - epsi) / (2 + epsi);
    if (rv < S.p.rho0 && S.ptype[i] <= 1) rv = S.p.rho0;
    S.rhoh[i] = rv;
    S.rho[i] = rv;
    double ml = S.ptype[i] > 1 ? 1.0 : 0.0;
    for (int k = 0; k < D; ++k) {
      double vv = S.v[k][i];
      double nv = vv + S.acc[k][i] * dt * ml;
      S.vh[k][i] = nv;
      S.v[k][i] = nv;
      double d = (vv + nv) * 0.5 * dt;
      S.cdx[k][i] += d;
      double xv = x[k][i] + d;
      S.xh[k][i] = xv;
      x[k][i] = xv;
    }
  }
}

// Δt_stepping (src/gpukernels_sph.jl:1293-1353); Julia's `maximum` propagates NaN.
double nanmax(double a, double b) {
  if (std::isnan(a) || std::isnan(b)) return std::numeric_limits<double>::quiet_NaN();
  return a > b ? a : b;
}
double dt_stepping(Sph& S, double tl0, double tl1) {
  int D = S.D;
  const Params& p = S.p;
  std::vector<double>* x = S.sys->x;
  double eta2 = (0.01 * p.h) * (0.01 * p.h);
  double amax = -std::numeric_limits<double>::infinity();
  for (int64_t i = 0; i < S.N; ++i) {
    Vec a{};
    getv(S.acc, D, i, a);
    S.buf[i] = std::sqrt(dotv(a, a, D));
    amax = i == 0 ? S.buf[i] : nanmax(amax, S.buf[i]);
  }
  double dt1 = std::sqrt(p.h / (3 * amax));
  double vmax = 0;
  for (int64_t i = 0; i < S.N; ++i) {
    Vec vp{}, pp{};
    getv(S.v, D, i, vp);
    getv(x, D, i, pp);
    S.buf[i] = std::fabs(p.h * dotv(vp, pp, D) / (dotv(pp, pp, D) + eta2));
    vmax = i == 0 ? S.buf[i] : nanmax(vmax, S.buf[i]);
  }
  double dt2 = p.h / (p.c0 + vmax);
  double dt = p.CFL * std::min(dt1, dt2);
  dt = std::min(std::max(dt, tl0), tl1);
  return dt;
}

// dpcreg! + update_dpcreg! (src/gpukernels_sph.jl:1485-1553, :1607-1632)
void dpc(Sph& S, double dpckernlim) {
  int D = S.D;
  const Params& p = S.p;
  std::vector<double>* x = S.sys->x;
  for (int k = 0; k < D; ++k) std::fill(S.buf2[k].begin(), S.buf2[k].end(), 0.0);
  double l0inv = 1 / p.dpc_l0;
  double whinv = 1 / S.kern.W(0.5, l0inv);
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (i == 0) continue;
    --i;
    --j;
    if (!(S.ptype[i] > 0 && S.ptype[j] > 0)) continue;
    double eta2 = (0.1 * p.dpc_l0) * (0.1 * p.dpc_l0);
    Vec dx{}, dv{};
    for (int k = 0; k < D; ++k) {
      dx[k] = x[k][i] - x[k][j];
      dv[k] = S.v[k][i] - S.v[k][j];
    }
    double ri = S.rho[i], rj = S.rho[j];
    double r2 = dotv(dx, dx, D);
    double r = std::sqrt(r2);
    double u = r * l0inv;
    double w = S.kern.W(u, l0inv);
    double chi = std::sqrt(w * whinv);
    double kk = u < dpckernlim ? 1.0 : chi;
    double Pb = chi * std::max(std::min(p.dpc_lambda * std::fabs(S.P[i] + S.P[j]), p.dpc_lambda * p.dpc_pmax),
                               p.dpc_pmin);
    double vr = dotv(dv, dx, D);
    Vec dd{};
    if (vr < 0) {
      double vrdr = vr / (r2 + eta2);
      for (int k = 0; k < D; ++k) dd[k] = kk * (-vrdr * dx[k]);
    } else {
      double tvar = 2 * p.dt * Pb / ((ri + rj) * (r2 + eta2));
      for (int k = 0; k < D; ++k) dd[k] = tvar * dx[k];
    }
    for (int k = 0; k < D; ++k) {
      S.buf2[k][i] += dd[k];
      S.buf2[k][j] -= dd[k];
    }
  }
  for (int k = 0; k < D; ++k) S.dvdpc[k] = S.buf2[k];
  for (int64_t i = 0; i < S.N; ++i)
    if (S.ptype[i] > 1)
      for (int k = 0; k < D; ++k) {
        double dv = S.buf2[k][i];
        S.v[k][i] += dv;
        x[k][i] += dv * p.dt;
      }
}

// xsphcorr! + update_xsphcorr! (src/gpukernels_sph.jl:1745-1779, :1817-1838)
void xsph(Sph& S) {
  int D = S.D;
  const Params& p = S.p;
  std::vector<double>* x = S.sys->x;
  for (int k = 0; k < D; ++k) std::fill(S.buf2[k].begin(), S.buf2[k].end(), 0.0);
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (i == 0) continue;
    --i;
    --j;
    if (!(S.ptype[i] > 1 && S.ptype[j] > 1)) continue;
    double ri = S.rho[i], rj = S.rho[j];
    double xs = 2 * p.m0 * p.xsph_eps * S.W[q] / (ri + rj);
    for (int k = 0; k < D; ++k) {
      double xv = (S.v[k][j] - S.v[k][i]) * xs;
      S.buf2[k][i] += xv;
      S.buf2[k][j] -= xv;
    }
  }
  for (int64_t i = 0; i < S.N; ++i)
    if (S.ptype[i] > 1)
      for (int k = 0; k < D; ++k) x[k][i] += S.buf2[k][i] * p.dt;
}

// cspmcorr! (src/gpukernels_sph.jl:1673-1719)
void cspm(Sph& S) {
  const Params& p = S.p;
  std::vector<double>& s1 = S.buf2[0];
  std::vector<double>& s2 = S.buf2[1];
  for (int k = 0; k < S.D; ++k) std::fill(S.buf2[k].begin(), S.buf2[k].end(), 0.0);
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (i == 0) continue;
    --i;
    --j;
    if (!(S.ptype[i] > 0 && S.ptype[j] > 0)) continue;
    double ri = S.rho[i], rj = S.rho[j];
    double w = S.W[q];
    s1[i] += p.m0 * w;
    s2[i] += w * p.m0 / rj;
    s1[j] += p.m0 * w;
    s2[j] += w * p.m0 / ri;
  }
  for (int64_t i = 0; i < S.N; ++i) {
    double nr = s1[i] / s2[i];
    if (!std::isnan(nr)) S.rho[i] = nr;
  }
}

// sph∑W! + sph∑∇W! (src/gpukernels_sph.jl:139-156, :332-359)
void diagnostics(Sph& S) {
  int D = S.D;
  std::vector<double>* x = S.sys->x;
  std::fill(S.sumW.begin(), S.sumW.end(), 0.0);
  for (int k = 0; k < D; ++k) std::fill(S.sumgW[k].begin(), S.sumgW[k].end(), 0.0);
  for (int64_t q = 0; q < S.npr; ++q) {
    int32_t i = S.pr[2 * q], j = S.pr[2 * q + 1];
    if (i == 0) continue;
    --i;
    --j;
    if ((S.ptype[i] > 0) != (S.ptype[j] > 0)) continue;
    Vec dx{};
    for (int k = 0; k < D; ++k) dx[k] = x[k][i] - x[k][j];
    double r = std::sqrt(dotv(dx, dx, D));
    double u = r * S.Hinv;
    double w = S.kern.W(u, S.Hinv);
    S.sumW[i] += w;
    S.sumW[j] += w;
    double dwk_r = S.kern.dW(u, S.Hinv) / r;
    for (int k = 0; k < D; ++k) {
      double g = dx[k] * dwk_r;
      S.sumgW[k][i] += g;
      S.sumgW[k][j] -= g;
    }
  }
}

double dpc_kernlim(const Wendland& k) {  // root of 1 - W(u, 1) on [0,1]
  double lo = 0.0, hi = 1.0;
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (1.0 - k.W(mid, 1.0) < 0) lo = mid;
    else hi = mid;
  }
  return std::fabs(1.0 - k.W(lo, 1.0)) <= std::fabs(1.0 - k.W(hi, 1.0)) ? lo : hi;
}

// _stepsolve! (src/sphproblem.jl:207-345)
int stepsolve(Sph& S, int64_t n, bool timestepping, double tl0, double tl1, int64_t* updaten_out,
              double* maxcdx_out) {
  CellList<double>& sys = *S.sys;
  int D = S.D;
  const Params& p = S.p;
  bool skipupdate = false;
  int64_t updaten = 0;
  double maxcdxout = 0.0;
  int64_t cspmcorrn = 0;
  double dpckernlim = dpc_kernlim(S.kern);
  S.pr = sys.pairs.data();
  S.npr = sys.capacity;
  for (int64_t iter = 1; iter <= n; ++iter) {
    cspmcorrn += 1;
    if (!skipupdate) {
      celllist_update(sys, true);
      if (sys.pairsn > sys.capacity) return fail("pair list overflow");
      S.pr = sys.pairs.data();
      S.npr = sys.pairsn;
      for (int k = 0; k < D; ++k) std::fill(S.cdx[k].begin(), S.cdx[k].end(), 0.0);
      skipupdate = true;
      updaten += 1;
    }
    for (int k = 0; k < D; ++k) std::fill(S.acc[k].begin(), S.acc[k].end(), 0.0);
    std::vector<double>* x = sys.x;
    sphW_gradW(S, x);
    ddt(S, S.rho, S.v, x);
    pressure(S, S.rho);
    dvdt(S, S.rho);
    dvdt_av(S, x, S.rho, S.v);
    if (p.nu > 0) dvdt_visc(S, S.v, S.rho, x);
    addgrav(S);
    fbmolforce(S, x, 0.4, 2 * p.dx);
    if (p.s > 0) surface_tension(S, x);
    update_rho(S, S.rhoh, p.dt * 0.5);
    update_v(S, S.vh, p.dt * 0.5);
    update_x(S, S.xh, S.vh, p.dt * 0.5);
    for (int k = 0; k < D; ++k) std::fill(S.acc[k].begin(), S.acc[k].end(), 0.0);
    ddt(S, S.rhoh, S.vh, S.xh);
    pressure(S, S.rhoh);
    dvdt(S, S.rhoh);
    dvdt_av(S, S.xh, S.rhoh, S.vh);
    if (p.nu > 0) dvdt_visc(S, S.vh, S.rhoh, S.xh);
    addgrav(S);
    fbmolforce(S, x, p.bound_D, p.bound_l);
    if (p.s > 0) surface_tension(S, S.xh);
    symplectic(S);
    if (p.dpc_l0 > 0) {
      pressure(S, S.rho);
      dpc(S, dpckernlim);
    }
    if (p.xsph_eps > 0) xsph(S);
    if (p.cspmn > 0 && cspmcorrn == p.cspmn) {
      cspm(S);
      cspmcorrn = 0;
    }
    double maxcdx = 0;
    bool first = true;
    for (int k = 0; k < D; ++k)
      for (int64_t i = 0; i < S.N; ++i) {
        double a = std::fabs(S.cdx[k][i]);
        maxcdx = first ? a : nanmax(maxcdx, a);
        first = false;
      }
    if (maxcdx > 0.9 * p.nui) skipupdate = false;
    maxcdxout = nanmax(maxcdxout, maxcdx);
    S.p.etime += p.dt;
    if (timestepping) S.p.dt = dt_stepping(S, tl0, tl1);
  }
  diagnostics(S);
  *updaten_out = updaten;
  *maxcdx_out = maxcdxout;
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI used by the tests (ctypes)
// ---------------------------------------------------------------------------
extern "C" {

struct orc_gcl {
  int prec;  // 64 or 32
  CellList<double> d;
  CellList<float> f;
};

const char* orc_last_error() { return g_err.c_str(); }

orc_gcl* orc_gcl_create(int prec, int dim, int64_t n, const double* const* pts, const double* cellsize,
                        double dist, const double* MIN, const double* MAX, int64_t mppcell, int64_t mpairs,
                        int* warn) {
  orc_gcl* h = new orc_gcl();
  h->prec = prec;
  int rc;
  if (prec == 32) {
    rc = celllist_create(h->f, dim, n, pts, cellsize, dist, MIN, MAX, mppcell, mpairs, warn);
  } else {
    rc = celllist_create(h->d, dim, n, pts, cellsize, dist, MIN, MAX, mppcell, mpairs, warn);
  }
  if (rc) {
    delete h;
    return nullptr;
  }
  return h;
}

void orc_gcl_destroy(orc_gcl* h) { delete h; }

int orc_gcl_update(orc_gcl* h, int fillzero) {
  if (h->prec == 32) celllist_update(h->f, fillzero != 0);
  else celllist_update(h->d, fillzero != 0);
  return 0;
}
int orc_gcl_partialupdate(orc_gcl* h, int fillzero) {
  if (h->prec == 32) celllist_partialupdate(h->f, fillzero != 0);
  else celllist_partialupdate(h->d, fillzero != 0);
  return 0;
}
int orc_gcl_set_points(orc_gcl* h, const double* const* pts) {
  if (h->prec == 32) {
    for (int k = 0; k < h->f.D; ++k)
      for (int64_t i = 0; i < h->f.n; ++i) h->f.x[k][i] = float(pts[k][i]);
  } else {
    for (int k = 0; k < h->d.D; ++k) std::copy(pts[k], pts[k] + h->d.n, h->d.x[k].begin());
  }
  return 0;
}
int orc_gcl_get_points(orc_gcl* h, double* const* out) {
  if (h->prec == 32) {
    for (int k = 0; k < h->f.D; ++k)
      for (int64_t i = 0; i < h->f.n; ++i) out[k][i] = h->f.x[k][i];
  } else {
    for (int k = 0; k < h->d.D; ++k) std::copy(h->d.x[k].begin(), h->d.x[k].end(), out[k]);
  }
  return 0;
}
// info: grid[3], offset[3] (as stored, in T), cs[3], and scalars
void orc_gcl_info(orc_gcl* h, int64_t* grid, double* offset, double* cs, int64_t* scal /*n,mppcell,cap,pairsn*/) {
  if (h->prec == 32) {
    for (int k = 0; k < 3; ++k) {
      grid[k] = h->f.grid[k];
      offset[k] = h->f.off[k];
      cs[k] = h->f.cs[k];
    }
    scal[0] = h->f.n;
    scal[1] = h->f.mppcell;
    scal[2] = h->f.capacity;
    scal[3] = h->f.pairsn;
  } else {
    for (int k = 0; k < 3; ++k) {
      grid[k] = h->d.grid[k];
      offset[k] = h->d.off[k];
      cs[k] = h->d.cs[k];
    }
    scal[0] = h->d.n;
    scal[1] = h->d.mppcell;
    scal[2] = h->d.capacity;
    scal[3] = h->d.pairsn;
  }
}
int64_t orc_gcl_get_pairs(orc_gcl* h, int32_t* out, int64_t cap) {
  const std::vector<int32_t>& pv = h->prec == 32 ? h->f.pairs : h->d.pairs;
  int64_t c = std::min<int64_t>(cap, int64_t(pv.size() / 2));
  std::copy(pv.begin(), pv.begin() + 2 * c, out);
  return c;
}
int orc_gcl_get_cells(orc_gcl* h, int32_t* pcell, int32_t* cellpnum) {
  if (h->prec == 32) {
    for (int k = 0; k < h->f.D; ++k) std::copy(h->f.pcell[k].begin(), h->f.pcell[k].end(), pcell + k * h->f.n);
    std::copy(h->f.cellpnum.begin(), h->f.cellpnum.end(), cellpnum);
  } else {
    for (int k = 0; k < h->d.D; ++k) std::copy(h->d.pcell[k].begin(), h->d.pcell[k].end(), pcell + k * h->d.n);
    std::copy(h->d.cellpnum.begin(), h->d.cellpnum.end(), cellpnum);
  }
  return 0;
}
int64_t orc_gcl_maxpairs(orc_gcl* h) { return h->prec == 32 ? maxpairs(h->f) : maxpairs(h->d); }

// ---- SPH ----
struct orc_sph {
  Sph s;
};

orc_sph* orc_sph_create(orc_gcl* sys, const double* par /* Params order */, const double* rho,
                        const int32_t* ptype, const double* const* v) {
  if (sys->prec != 64) {
    fail("SPHProblem oracle runs in Float64 only");
    return nullptr;
  }
  orc_sph* h = new orc_sph();
  Sph& S = h->s;
  S.sys = &sys->d;
  S.D = sys->d.D;
  S.N = sys->d.n;
  Params& p = S.p;
  p.dx = par[0];
  p.h = par[1];
  p.H = par[2];
  p.rho0 = par[3];
  p.m0 = par[4];
  p.dt = par[5];
  p.alpha = par[6];
  p.c0 = par[7];
  p.gamma = int64_t(par[8]);
  p.delta = par[9];
  p.CFL = par[10];
  p.g = par[11];
  p.s = par[12];
  p.nu = par[13];
  p.dpc_l0 = par[14];
  p.dpc_pmin = par[15];
  p.dpc_pmax = par[16];
  p.dpc_lambda = par[17];
  p.xsph_eps = par[18];
  p.cspmn = int64_t(par[19]);
  p.bound_D = par[20];
  p.bound_l = par[21];
  p.nui = par[22];
  p.etime = par[23];
  S.hinv = 1 / p.h;
  S.Hinv = 1 / p.H;
  S.kern = Wendland(S.D);
  size_t N = size_t(S.N);
  S.rho.assign(rho, rho + N);
  S.rhoh = S.rho;  // deepcopy (src/sphproblem.jl:141)
  S.ptype.assign(ptype, ptype + N);
  S.P.assign(N, 0.0);
  S.buf.assign(N, 0.0);
  S.drho.assign(N, 0.0);
  S.sumW.assign(N, 0.0);
  for (int k = 0; k < S.D; ++k) {
    if (v) S.v[k].assign(v[k], v[k] + N);
    else S.v[k].assign(N, 0.0);
    S.vh[k] = S.v[k];
    S.xh[k] = sys->d.x[k];
    S.acc[k].assign(N, 0.0);
    S.buf2[k].assign(N, 0.0);
    S.dvdpc[k].assign(N, 0.0);
    S.cdx[k].assign(N, 0.0);
    S.sumgW[k].assign(N, 0.0);
  }
  return h;
}

void orc_sph_destroy(orc_sph* h) { delete h; }

int orc_sph_stepsolve(orc_sph* h, int64_t n, int timestepping, double tl0, double tl1, int64_t* updaten,
                      double* maxcdx) {
  return stepsolve(h->s, n, timestepping != 0, tl0, tl1, updaten, maxcdx);
}

// fields: 0 x, 1 v, 2 rho, 3 P, 4 acc, 5 sumW, 6 sumgradW, 7 rhoh, 8 vh, 9 xh, 10 drho, 11 cdx
int orc_sph_get_field(orc_sph* h, int field, double* out) {
  Sph& S = h->s;
  size_t N = size_t(S.N);
  auto vec = [&](const std::vector<double>* a) {
    for (int k = 0; k < S.D; ++k) std::copy(a[k].begin(), a[k].end(), out + k * N);
  };
  switch (field) {
    case 0: vec(S.sys->x); break;
    case 1: vec(S.v); break;
    case 2: std::copy(S.rho.begin(), S.rho.end(), out); break;
    case 3: std::copy(S.P.begin(), S.P.end(), out); break;
    case 4: vec(S.acc); break;
    case 5: std::copy(S.sumW.begin(), S.sumW.end(), out); break;
    case 6: vec(S.sumgW); break;
    case 7: std::copy(S.rhoh.begin(), S.rhoh.end(), out); break;
    case 8: vec(S.vh); break;
    case 9: vec(S.xh); break;
    case 10: std::copy(S.drho.begin(), S.drho.end(), out); break;
    case 11: vec(S.cdx); break;
    case 13: vec(S.dvdpc); break;
    default: return fail("unknown field");
  }
  return 0;
}

double orc_sph_get_scalar(orc_sph* h, int which) {  // 0 dt, 1 etime
  return which == 0 ? h->s.p.dt : h->s.p.etime;
}
void orc_sph_set_scalar(orc_sph* h, int which, double v) {
  if (which == 0) h->s.p.dt = v;
  else h->s.p.etime = v;
}

// Exposed building blocks for closed-form tests.
double orc_wendland_W(int dim, double u, double hinv) { return Wendland(dim).W(u, hinv); }
double orc_wendland_dW(int dim, double u, double hinv) { return Wendland(dim).dW(u, hinv); }
double orc_powfancy7th(double x) { return powfancy7th(x, 1.0 / 7.0, 7); }
double orc_dpckernlim(int dim) { return dpc_kernlim(Wendland(dim)); }
double orc_eos(double rho, int64_t gamma, double rho0, double c0, int32_t ptype) {
  double P0 = c0 * c0 * rho0 / double(gamma);
  return eos(rho, gamma, rho0, P0, ptype);
}

}  // extern "C"
