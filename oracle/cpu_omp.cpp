// Multi-core CPU step of the reference's SPH hot path: the CPU baseline that
// bench.py times beside the GPU.  It is the reference's algorithm
// (src/sphproblem.jl:207-345, per-pair expressions of src/gpukernels_sph.jl)
// reorganised for OpenMP: every particle gathers over its own neighbours (one
// thread writes each particle, so there are no atomics and the sums are
// deterministic for a fixed thread count or none), the neighbour list of a
// rebuild is a per-particle CSR, and pressure is evaluated once per particle
// per pass.  FP64 throughout, FMA contraction off, the per-pair arithmetic of
// oracle.cpp (shared through the include below).
//
// TEST INFRASTRUCTURE ONLY: used by bench.py's cpu_baseline leg and checked
// against the serial oracle in tests/.  No product code links or calls it.
//
// Scope: the physics of the headline workload and of the step's optional
// pair terms (kinematic viscosity, surface tension).  DPC, XSPH and CSPM
// (src/sphproblem.jl:304-322) are not restated here; a problem that enables
// them is refused.
#include "oracle.cpp"

#include <omp.h>

namespace {

struct OmpSph {
  int D = 2;
  int64_t n = 0;
  Params p{};
  double dist = 0, cs[3] = {0, 0, 0}, off[3] = {0, 0, 0};
  int64_t grid[3] = {1, 1, 1};
  Wendland kern{2};
  std::vector<double> x[3], v[3], xh[3], vh[3], acc[3], cdx[3];
  std::vector<double> rho, rhoh, P, drho, red;
  std::vector<int32_t> ptype;
  // neighbour list of the last rebuild: partners of particle i are
  // nb[nbo[i] .. nbo[i+1]), cells in stencil order, ascending index inside a cell
  std::vector<int64_t> nbo;
  std::vector<int32_t> nb;
  std::vector<int32_t> cellof, cstart, corder;
  int threads = 1;

  int64_t ncells() const { return grid[0] * grid[1] * (D == 3 ? grid[2] : 1); }
};

// kernel_cellmap! (src/gpukernels_cls.jl:29-60) as update! calls it: the
// offset stored in Float64, h⁻¹ = 1/cs.
void omp_rebuild(OmpSph& S) {
  const int D = S.D;
  const int64_t n = S.n, nc = S.ncells();
  double hinv[3];
  for (int k = 0; k < D; ++k) hinv[k] = 1.0 / S.cs[k];
  S.cellof.resize(size_t(n));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int64_t c = 0, mul = 1;
    for (int k = 0; k < D; ++k) {
      double q = (S.x[k][i] - S.off[k]) * hinv[k];
      double cl = q >= 1.0 ? q : 1.0;  // min(max(q, 1), grid); NaN -> cell 1 as in the oracle
      cl = cl <= double(S.grid[k]) ? cl : double(S.grid[k]);
      c += (int64_t(std::ceil(cl)) - 1) * mul;
      mul *= S.grid[k];
    }
    S.cellof[size_t(i)] = int32_t(c);
  }
  // counting sort by cell; inside a cell by particle index
  S.cstart.assign(size_t(nc + 1), 0);
  std::vector<int32_t>& cnt = S.cstart;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
#pragma omp atomic
    cnt[size_t(S.cellof[size_t(i)]) + 1] += 1;
  }
  for (int64_t c = 0; c < nc; ++c) cnt[size_t(c + 1)] += cnt[size_t(c)];
  std::vector<int32_t> fill(S.cstart.begin(), S.cstart.end() - 1);
  S.corder.resize(size_t(n));
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int32_t s;
#pragma omp atomic capture
    s = fill[size_t(S.cellof[size_t(i)])]++;
    S.corder[size_t(s)] = int32_t(i);
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t c = 0; c < nc; ++c)
    if (S.cstart[size_t(c + 1)] - S.cstart[size_t(c)] > 1)
      std::sort(S.corder.begin() + S.cstart[size_t(c)], S.corder.begin() + S.cstart[size_t(c + 1)]);
  // neighbour list: r² < dist² with the reference's expression (src/gpukernels_cls.jl:293-554)
  const double d2 = S.dist * S.dist;
  S.nbo.assign(size_t(n + 1), 0);
  auto scan = [&](int64_t i, auto&& emit) {
    const int64_t c = S.cellof[size_t(i)];
    const int64_t gx = S.grid[0], gy = S.grid[1], gz = D == 3 ? S.grid[2] : 1;
    const int64_t ix = c % gx, iy = (c / gx) % gy, iz = c / (gx * gy);
    for (int64_t dz = (D == 3 ? -1 : 0); dz <= (D == 3 ? 1 : 0); ++dz)
      for (int64_t dy = -1; dy <= 1; ++dy)
        for (int64_t dxo = -1; dxo <= 1; ++dxo) {
          const int64_t jx = ix + dxo, jy = iy + dy, jz = iz + dz;
          if (jx < 0 || jx >= gx || jy < 0 || jy >= gy || jz < 0 || jz >= gz) continue;
          const int64_t cj = jx + gx * (jy + gy * jz);
          for (int32_t s = S.cstart[size_t(cj)]; s < S.cstart[size_t(cj + 1)]; ++s) {
            const int32_t j = S.corder[size_t(s)];
            if (j == i) continue;
            double r2 = 0;
            for (int k = 0; k < D; ++k) {
              const double dd = S.x[k][i] - S.x[k][j];
              r2 = k == 0 ? dd * dd : r2 + dd * dd;
            }
            if (r2 < d2) emit(j);
          }
        }
  };
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i) {
    int64_t m = 0;
    scan(i, [&](int32_t) { ++m; });
    S.nbo[size_t(i + 1)] = m;
  }
  for (int64_t i = 0; i < n; ++i) S.nbo[size_t(i + 1)] += S.nbo[size_t(i)];
  S.nb.resize(size_t(S.nbo[size_t(n)]));
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i) {
    int64_t o = S.nbo[size_t(i)];
    scan(i, [&](int32_t j) { S.nb[size_t(o++)] = j; });
  }
}

// One evaluation of the pair physics for particle i over its partners
// (pass 1: state at tⁿ; pass 2: half-step state, ∇W and the wall force at xⁿ).
// For the pair (i, j) it is the reference's expression for i's side with
// Δx = x_i - x_j (src/gpukernels_sph.jl:571-625, 750-772, 866-897, 963-987,
// 1385-1405, 1889-1909).
void omp_pass(OmpSph& S, bool second) {
  const int D = S.D;
  const Params& p = S.p;
  const std::vector<double>* X = second ? S.xh : S.x;
  const std::vector<double>* V = second ? S.vh : S.v;
  const std::vector<double>& RHO = second ? S.rhoh : S.rho;
  const double P0 = p.c0 * p.c0 * p.rho0 / double(p.gamma);
  const double eta2 = (0.1 * p.h) * (0.1 * p.h);
  const double ginv = 1.0 / double(p.gamma);
  const double DDTkh = 2 * p.h * p.delta * p.c0;
  const double Cb = (p.c0 * p.c0 * p.rho0) * ginv;
  const double DDTgz = p.rho0 * p.g / Cb;
  const double fd = second ? p.bound_D : 0.4, fr0 = second ? p.bound_l : 2 * p.dx;
  const double feta2 = 0.01 * (fr0 * fr0);
  const double Hinv = 1.0 / p.H;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < S.n; ++i) S.P[size_t(i)] = eos(RHO[size_t(i)], p.gamma, p.rho0, P0, S.ptype[size_t(i)]);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < S.n; ++i) {
    const int32_t pti = S.ptype[size_t(i)];
    const double ri = RHO[size_t(i)], Pi = S.P[size_t(i)];
    double a[3] = {0, 0, 0}, dr = 0;
    for (int64_t e = S.nbo[size_t(i)]; e < S.nbo[size_t(i + 1)]; ++e) {
      const int32_t j = S.nb[size_t(e)];
      const int32_t ptj = S.ptype[size_t(j)];
      Vec dxn{}, dx{}, dv{}, gw{};
      for (int k = 0; k < D; ++k) {
        dxn[k] = S.x[k][size_t(i)] - S.x[k][size_t(j)];
        dx[k] = X[k][size_t(i)] - X[k][size_t(j)];
        dv[k] = V[k][size_t(i)] - V[k][size_t(j)];
      }
      const double r2n = dotv(dxn, dxn, D);
      const double rn = std::sqrt(r2n);
      const double dwk_r = S.kern.dW(rn * Hinv, Hinv) / rn;
      for (int k = 0; k < D; ++k) gw[k] = dxn[k] * dwk_r;
      const double r2 = dotv(dx, dx, D);
      const double rj = RHO[size_t(j)];
      const bool both = pti > 0 && ptj > 0;
      if (both) {
        double s = p.m0 * dotv(dv, gw, D);
        if (pti > 1) {
          const double drhop = p.rho0 * powfancy7th(1 + DDTgz * dx[1], ginv, int(p.gamma)) - p.rho0;
          const double vd = DDTkh * (rj - ri - drhop) / (r2 + eta2);
          s += vd * -dotv(dx, gw, D) * p.m0 / rj;
        }
        dr += s;
      }
      if (int64_t(pti) * int64_t(ptj) > 0) {
        const double Pfac = (Pi + S.P[size_t(j)]) / (ri * rj);
        for (int k = 0; k < D; ++k) a[k] += -p.m0 * Pfac * gw[k];
      }
      if (both) {
        const double cond = dotv(dv, dx, D);
        if (cond < 0) {
          const double dmu = p.h * cond / (r2 + eta2);
          const double dPi = (-p.alpha * p.c0 * dmu) / ((ri + rj) * 0.5);
          for (int k = 0; k < D; ++k) a[k] += -dPi * p.m0 * gw[k];
        }
        if (p.nu > 0) {
          const double term = 4 * p.nu * p.m0 * dotv(dx, gw, D) / ((ri + rj) * (r2 + eta2));
          for (int k = 0; k < D; ++k) a[k] += term * dv[k];
        }
        if (p.s > 0) {
          const double r = std::sqrt(r2);
          if (r < 2 * p.h) {
            const double scos = p.s * std::cos(1.5 * M_PI * r / (2 * p.h)) / (r + (0.1 * p.h));
            for (int k = 0; k < D; ++k) a[k] += scos * dx[k] / p.m0;
          }
        }
      }
      if ((pti > 0) != (ptj > 0) && rn < fr0) {
        const double q1 = fr0 / (rn + feta2);
        const double Fc = fd * (powi(q1, 12) - powi(q1, 4)) / (r2n + feta2);
        for (int k = 0; k < D; ++k) a[k] += dxn[k] * Fc;
      }
    }
    a[1] = a[1] - p.g;
    for (int k = 0; k < D; ++k) S.acc[k][size_t(i)] = a[k];
    S.drho[size_t(i)] = dr;
  }
}

// _stepsolve! (src/sphproblem.jl:207-345) for the physics above.
void omp_stepsolve(OmpSph& S, int64_t nsteps, bool ts, double tl0, double tl1, int64_t* updaten, double* maxcdx) {
  const int D = S.D;
  Params& p = S.p;
  bool skip = false;
  int64_t upd = 0;
  double mxout = 0;
  for (int64_t it = 1; it <= nsteps; ++it) {
    if (!skip) {
      omp_rebuild(S);
      for (int k = 0; k < D; ++k) std::fill(S.cdx[k].begin(), S.cdx[k].end(), 0.0);
      skip = true;
      ++upd;
    }
    omp_pass(S, false);
    const double hdt = p.dt * 0.5;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < S.n; ++i) {  // update_ρ / update_v / update_x (src/gpukernels_sph.jl:1078-1169)
      double rv = S.rhoh[size_t(i)] + S.drho[size_t(i)] * hdt;
      if (rv < p.rho0 && S.ptype[size_t(i)] < 0) rv = p.rho0;
      S.rhoh[size_t(i)] = rv;
      for (int k = 0; k < D; ++k) {
        if (S.ptype[size_t(i)] > 1) S.vh[k][size_t(i)] += S.acc[k][size_t(i)] * hdt;
        S.xh[k][size_t(i)] += S.vh[k][size_t(i)] * hdt;
      }
    }
    omp_pass(S, true);
    double mx = 0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int64_t i = 0; i < S.n; ++i) {  // symplectic_update! (src/gpukernels_sph.jl:1204-1245)
      const double epsi = -(S.drho[size_t(i)] / S.rhoh[size_t(i)]) * p.dt;
      double rv = S.rho[size_t(i)] * (2 - epsi) / (2 + epsi);
      if (rv < p.rho0 && S.ptype[size_t(i)] <= 1) rv = p.rho0;
      S.rhoh[size_t(i)] = rv;
      S.rho[size_t(i)] = rv;
      const double ml = S.ptype[size_t(i)] > 1 ? 1.0 : 0.0;
      for (int k = 0; k < D; ++k) {
        const double vv = S.v[k][size_t(i)];
        const double nv = vv + S.acc[k][size_t(i)] * p.dt * ml;
        S.vh[k][size_t(i)] = nv;
        S.v[k][size_t(i)] = nv;
        const double d = (vv + nv) * 0.5 * p.dt;
        S.cdx[k][size_t(i)] += d;
        mx = std::max(mx, std::fabs(S.cdx[k][size_t(i)]));
        const double xv = S.x[k][size_t(i)] + d;
        S.xh[k][size_t(i)] = xv;
        S.x[k][size_t(i)] = xv;
      }
    }
    if (mx > 0.9 * p.nui) skip = false;
    mxout = std::max(mxout, mx);
    p.etime += p.dt;
    if (ts) {  // Δt_stepping (src/gpukernels_sph.jl:1293-1353)
      const double eta2 = (0.01 * p.h) * (0.01 * p.h);
      double amax = 0, vmax = 0;
#pragma omp parallel for schedule(static) reduction(max : amax, vmax)
      for (int64_t i = 0; i < S.n; ++i) {
        Vec a{}, vp{}, pp{};
        for (int k = 0; k < D; ++k) {
          a[k] = S.acc[k][size_t(i)];
          vp[k] = S.v[k][size_t(i)];
          pp[k] = S.x[k][size_t(i)];
        }
        amax = std::max(amax, std::sqrt(dotv(a, a, D)));
        vmax = std::max(vmax, std::fabs(p.h * dotv(vp, pp, D) / (dotv(pp, pp, D) + eta2)));
      }
      const double dt1 = std::sqrt(p.h / (3 * amax));
      const double dt2 = p.h / (p.c0 + vmax);
      p.dt = std::min(std::max(p.CFL * std::min(dt1, dt2), tl0), tl1);
    }
  }
  *updaten = upd;
  *maxcdx = mxout;
}

}  // namespace

extern "C" {

// The problem's cell grid follows the reference constructor
// (src/celllist.jl:22-57): cell size raised to dist, offset MIN - |(MIN +
// √eps)·√eps|, grid = ceil(range / cs) over the points' extent.
OmpSph* omp_sph_create(int dim, int64_t n, const double* const* pts, const double* cellsize, double dist,
                       const double* par, const double* rho, const int32_t* ptype, const double* const* v) {
  if (dim < 2 || dim > 3 || n <= 0) {
    fail("dim must be 2 or 3 and n > 0");
    return nullptr;
  }
  if (par[14] > 0 || par[18] > 0) {
    fail("the multi-core step does not restate DPC or XSPH");
    return nullptr;
  }
  OmpSph* S = new OmpSph();
  S->D = dim;
  S->n = n;
  S->dist = dist;
  S->kern = Wendland{dim};
  const double sq = std::sqrt(std::numeric_limits<double>::epsilon());
  for (int k = 0; k < dim; ++k) {
    S->cs[k] = std::max(cellsize[k], dist);
    double mn = *std::min_element(pts[k], pts[k] + n), mx = *std::max_element(pts[k], pts[k] + n);
    mn = mn - std::fabs((mn + sq) * sq);
    S->off[k] = mn;
    S->grid[k] = std::max<int64_t>(1, int64_t(std::ceil((mx - mn) / S->cs[k])));
    S->x[k].assign(pts[k], pts[k] + n);
    S->xh[k] = S->x[k];
    S->v[k].assign(size_t(n), 0.0);
    if (v) S->v[k].assign(v[k], v[k] + n);
    S->vh[k] = S->v[k];
    S->acc[k].assign(size_t(n), 0.0);
    S->cdx[k].assign(size_t(n), 0.0);
  }
  Params& p = S->p;
  p.dx = par[0]; p.h = par[1]; p.H = par[2]; p.rho0 = par[3]; p.m0 = par[4]; p.dt = par[5]; p.alpha = par[6];
  p.c0 = par[7]; p.gamma = int64_t(par[8]); p.delta = par[9]; p.CFL = par[10]; p.g = par[11]; p.s = par[12];
  p.nu = par[13]; p.dpc_l0 = par[14]; p.dpc_pmin = par[15]; p.dpc_pmax = par[16]; p.dpc_lambda = par[17];
  p.xsph_eps = par[18]; p.cspmn = int64_t(par[19]); p.bound_D = par[20]; p.bound_l = par[21]; p.nui = par[22];
  p.etime = par[23];
  S->rho.assign(rho, rho + n);
  S->rhoh = S->rho;
  S->ptype.assign(ptype, ptype + n);
  S->P.assign(size_t(n), 0.0);
  S->drho.assign(size_t(n), 0.0);
  S->threads = omp_get_max_threads();
  return S;
}

void omp_sph_destroy(OmpSph* S) { delete S; }

// CSPM (cspmn > 0) fires only on iterations that are multiples of cspmn
// inside one call: calls shorter than cspmn never reach it.
int omp_sph_stepsolve(OmpSph* S, int64_t n, int ts, double tl0, double tl1, int64_t* updaten, double* maxcdx) {
  if (S->p.cspmn > 0 && n >= S->p.cspmn) return fail("the multi-core step does not restate CSPM");
  omp_stepsolve(*S, n, ts != 0, tl0, tl1, updaten, maxcdx);
  return 0;
}

// fields: 0 x, 1 v, 2 rho, 4 acc, 10 drho (vectors as dim blocks of n)
int omp_sph_get_field(OmpSph* S, int code, double* out) {
  const size_t n = size_t(S->n);
  auto vec = [&](const std::vector<double>* a) {
    for (int k = 0; k < S->D; ++k) std::copy(a[k].begin(), a[k].end(), out + k * n);
  };
  switch (code) {
    case 0: vec(S->x); return 0;
    case 1: vec(S->v); return 0;
    case 2: std::copy(S->rho.begin(), S->rho.end(), out); return 0;
    case 4: vec(S->acc); return 0;
    case 10: std::copy(S->drho.begin(), S->drho.end(), out); return 0;
    default: return fail("unknown field");
  }
}

double omp_sph_get_scalar(OmpSph* S, int which) { return which == 0 ? S->p.dt : S->p.etime; }
int omp_sph_threads(OmpSph* S) { return S->threads; }
int64_t omp_sph_pairs(OmpSph* S) { return S->nbo.empty() ? 0 : S->nbo.back() / 2; }

}  // extern "C"
