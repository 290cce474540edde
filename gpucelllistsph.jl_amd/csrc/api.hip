// Host objects and the C ABI of libgclsph (include/gclsph.h).
//
// gcl_s mirrors GPUCellList (src/celllist.jl:1-15) and its host driver
// (update!, partialupdate!, neighborlist); sph_s mirrors SPHProblem
// (src/sphproblem.jl:62-193) and _stepsolve! (:207-345).  The SPH step keeps
// its whole state on the device, in cell-sorted order, and runs a batch of
// steps with no host synchronisation until the batch ends.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gclsph.h"
#include "sph.hpp"

using namespace gclsph;

namespace {

thread_local std::string g_last_error;

template <typename F>
int guard(F&& f) {
  try {
    f();
    g_last_error.clear();
    return GCLSPH_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return GCLSPH_EHIP;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return GCLSPH_EHIP;
  }
}

void require_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) throw Error(kErrNoDevice, "no HIP device is available; libgclsph has no CPU path");
}

template <typename T>
__global__ void k_gather(const T* __restrict__ src, T* __restrict__ dst, const int32_t* __restrict__ perm, int64_t n) {
  int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s < n) dst[s] = src[perm[s]];
}
template <typename T>
__global__ void k_scatter_by(const T* __restrict__ src, T* __restrict__ dst, const int32_t* __restrict__ id,
                             int64_t n) {
  int64_t s = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (s < n) dst[s] = src[id[s]];
}

template <typename T>
void gather(const T* src, T* dst, const int32_t* perm, int64_t n, hipStream_t s) {
  k_gather<T><<<grid_for(n), kBlock, 0, s>>>(src, dst, perm, n);
  GCL_HIP(hipGetLastError());
}

__global__ void k_maxpairs(const int32_t* __restrict__ cnt, CellGeom g, unsigned long long* out) {
  int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t nc = g.ncells();
  long long val = 0;
  if (c < nc) {
    int32_t nn = cnt[c];
    if (nn > 0) {
      // мaxpairs half stencil (src/gpukernels_cls.jl:130-167, :184-286)
      const int o2[4][3] = {{-1, 1, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}};
      const int o3[13][3] = {{-1, 1, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0},  {-1, 1, 1},  {0, 1, 1}, {1, 1, 1},
                             {1, 0, 1},  {0, 0, 1}, {-1, 1, -1}, {0, 1, -1}, {1, 1, -1}, {1, 0, -1}};
      int64_t gx = g.grid[0], gy = g.grid[1], gz = g.dim == 3 ? g.grid[2] : 1;
      int64_t ix = c % gx, iy = (c / gx) % gy, iz = c / (gx * gy);
      long long m = 0;
      int no = g.dim == 2 ? 4 : 13;
      for (int o = 0; o < no; ++o) {
        const int* off = g.dim == 2 ? o2[o] : o3[o];
        int64_t jx = ix + off[0], jy = iy + off[1], jz = iz + off[2];
        if (jx >= 0 && jx < gx && jy >= 0 && jy < gy && jz >= 0 && jz < gz) m += cnt[jx + gx * (jy + gy * jz)];
      }
      val = (long long)(double(int32_t(nn * (nn - 1))) * 0.5) + m * (long long)nn;
    }
  }
  if (val) atomicAdd(out, (unsigned long long)val);
}

int64_t ceil_div08(int64_t v) { return int64_t(std::ceil(double(v) / 0.8)); }

}  // namespace

void gclsph::set_last_error(const std::string& m) { g_last_error = m; }

struct sph_s;

// --------------------------------------------------------------------- gcl_s
struct gcl_s {
  int dim = 2, prec = 64;
  int64_t n = 0;
  double dist = 0;
  double cs[3] = {0, 0, 0};
  CellGeom G{};
  int64_t mppcell = 0, pairsn = 0;
  void* xo[3] = {nullptr, nullptr, nullptr};  // points in caller order (T)
  void* xs[3] = {nullptr, nullptr, nullptr};  // points in cell order at the last update (T)
  SortBuffers sb;
  PairBuffers pb;
  hipStream_t stream = nullptr;
  sph_s* attached = nullptr;
  bool pairs_valid = true;

  size_t tsize() const { return prec == 32 ? 4 : 8; }

  ~gcl_s() {
    for (int k = 0; k < 3; ++k) {
      if (xo[k]) (void)hipFree(xo[k]);
      if (xs[k]) (void)hipFree(xs[k]);
    }
    sb.release();
    pb.release();
    if (stream) (void)hipStreamDestroy(stream);
  }

  void upload_points(const double* const* pts) {
    for (int k = 0; k < dim; ++k) {
      if (prec == 32) {
        std::vector<float> tmp(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; ++i) tmp[size_t(i)] = float(pts[k][i]);
        GCL_HIP(hipMemcpy(xo[k], tmp.data(), 4 * size_t(n), hipMemcpyHostToDevice));
      } else {
        GCL_HIP(hipMemcpy(xo[k], pts[k], 8 * size_t(n), hipMemcpyHostToDevice));
      }
    }
  }

  int32_t read_maxcell() {
    int32_t m = 0;
    GCL_HIP(hipMemcpyAsync(&m, sb.maxcell, 4, hipMemcpyDeviceToHost, stream));
    GCL_HIP(hipStreamSynchronize(stream));
    return m;
  }

  int64_t count_pairs(const void* const* pos) {
    pair_count(G, prec, pos, sb.start, sb.skey, n, dist, pb, stream);
    int64_t total = 0;
    GCL_HIP(hipMemcpyAsync(&total, pb.off + n, 8, hipMemcpyDeviceToHost, stream));
    GCL_HIP(hipStreamSynchronize(stream));
    return total;
  }

  // update! (src/celllist.jl:95-134) for a standalone cell list.
  void update(bool fillzero) {
    bin_sort(G, prec, xo, n, false, sb, nullptr, stream, nullptr, nullptr, xs);
    int32_t maxpoint = read_maxcell();
    if (maxpoint > mppcell) mppcell = int64_t(std::ceil(double(maxpoint) * 1.05 + 1));
    bool fresh = false;  // a reallocated buffer is zero-filled whole
    if (pairsn > pb.capacity || double(pairsn) < double(pb.capacity) * 0.6) {
      pb.set_capacity(ceil_div08(pairsn), stream);
      fresh = true;
    }
    int64_t total = count_pairs(xs);
    if (total > pb.capacity) {
      pb.set_capacity(ceil_div08(total), stream);
      fresh = true;
    }
    if (fillzero && !fresh) pb.zero_from(total, stream);
    pair_emit(G, prec, xs, sb.start, sb.skey, sb.perm, n, dist, total, pb, stream);
    pairsn = total;
    GCL_HIP(hipStreamSynchronize(stream));
  }

  // partialupdate! (src/celllist.jl:141-145): current points, cells of the last update!.
  void partialupdate(bool fillzero) {
    gather_positions(dim, prec, xo, xs, sb.perm, n, stream);
    int64_t total = count_pairs(xs);
    if (fillzero) pb.zero_from(total, stream);
    pair_emit(G, prec, xs, sb.start, sb.skey, sb.perm, n, dist, total, pb, stream);
    pairsn = total;
    GCL_HIP(hipStreamSynchronize(stream));
  }
};

// --------------------------------------------------------------------- sph_s
struct sph_s {
  gcl_s* sys = nullptr;
  int dim = 2, prec = GCLSPH_F64;
  int64_t n = 0;
  double par[SPH_NPARAMS] = {};
  CellGeom G{};
  hipStream_t stream = nullptr;
  SortBuffers sb;
  std::vector<void*> allocs;
  SphBufs<double> Bd{};
  SphBufs<float> Bf{};
  StepCtl* ctl = nullptr;
  StepCtl hctl{};
  bool synced_prev = true;
  bool p_from_dpc = false;  // get_pressure reads the DPC-stage pressure
  int xrb_mode = 0;  // membership positions for diagnostics / pair list: 0 XRB, 1 other, 2 current
  uint64_t* ne_buf = nullptr;  // neighbour-mask entries, necap per particle
  int necap = 0;
  bool masks_valid = false;  // NE/NC hold the membership of the last rebuild
  // sumW / sum∇W of the last stepsolve call are computed on first use (or
  // before anything changes the state they describe)
  bool diag_pending = false;
  int diag_xm = 0, diag_masks = 0;
  LaunchSet own;     // this problem as a batch of one
  std::vector<unsigned char> own_desc;  // its descriptor (passed by value to the kernels)
  int batches = 0;   // batches that step this problem
  bool timing = false;
  double stage_ms[5] = {0, 0, 0, 0, 0};
  int64_t stage_cnt[5] = {0, 0, 0, 0, 0};
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_marks;

  bool split() const { return prec == GCLSPH_F32SPLIT; }

  ~sph_s() {
    for (void* p : allocs)
      if (p) (void)hipFree(p);
    if (ne_buf) (void)hipFree(ne_buf);
    own.release();
    sb.release();
    for (hipEvent_t e : ev_pool) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <typename T>
  T* dalloc(int64_t cnt) {
    void* p = nullptr;
    GCL_HIP(hipMalloc(&p, sizeof(T) * size_t(cnt < 1 ? 1 : cnt)));
    GCL_HIP(hipMemsetAsync(p, 0, sizeof(T) * size_t(cnt < 1 ? 1 : cnt), stream));
    allocs.push_back(p);
    return static_cast<T*>(p);
  }

  template <typename R>
  void alloc_bufs(SphBufs<R>& B) {
    B.n = n;
    for (int w = 0; w < 2; ++w) {
      for (int k = 0; k < dim; ++k) {
        B.X[w][k] = dalloc<double>(n + kPosPad);
        B.V[w][k] = dalloc<R>(n);
      }
      B.RHO[w] = dalloc<R>(n);
      B.PT[w] = dalloc<int32_t>(n);
      B.ID[w] = dalloc<int32_t>(n);
    }
    for (int k = 0; k < dim; ++k) {
      B.XH[k] = dalloc<double>(n);
      B.VH[k] = dalloc<R>(n);
      B.XH2[k] = dalloc<double>(n);
      B.VH2[k] = dalloc<R>(n);
      B.XRB[k] = dalloc<double>(n + kPosPad);
      B.CDX[k] = dalloc<double>(n);
      B.ACC[k] = dalloc<R>(n);
      B.DV[k] = dalloc<R>(n);
      B.SUMGW[k] = dalloc<double>(n);
    }
    for (int k = dim; k < 3; ++k) {  // unused third component aliases the second
      for (int w = 0; w < 2; ++w) {
        B.X[w][k] = B.X[w][1];
        B.V[w][k] = B.V[w][1];
      }
      B.XH[k] = B.XH[1];
      B.VH[k] = B.VH[1];
      B.XH2[k] = B.XH2[1];
      B.VH2[k] = B.VH2[1];
      B.XRB[k] = B.XRB[1];
      B.CDX[k] = B.CDX[1];
      B.ACC[k] = B.ACC[1];
      B.DV[k] = B.DV[1];
      B.SUMGW[k] = B.SUMGW[1];
    }
    B.RH = dalloc<R>(n);
    B.RH2 = dalloc<R>(n);
    B.DRHO = dalloc<R>(n);
    B.PBUF = dalloc<R>(n);
    B.PN = dalloc<R>(n);
    B.PH = dalloc<R>(n);
    B.TMP = dalloc<double>(n);
    B.nblk = int(grid_for(n));
    B.RED = dalloc<double>(3 * 64);  // zeroed here; k_finalize clears it after every step
    B.SUMW = dalloc<double>(n);
    B.start = sb.start;
    B.skey = sb.skey;
    B.perm = sb.perm;
    B.maxcell = sb.maxcell;
    B.ctl = ctl;
    B.NC = dalloc<int32_t>(n);
    B.NE = nullptr;
    B.necap = 0;
    const size_t rn = dim == 2 ? sizeof(typename Records<2, R>::N) : sizeof(typename Records<3, R>::N);
    const size_t rh = dim == 2 ? sizeof(typename Records<2, R>::H) : sizeof(typename Records<3, R>::H);
    B.PRN = dalloc<uint8_t>(int64_t(rn) * n);
    B.PRH = dalloc<uint8_t>(int64_t(rh) * n);
  }

  // Entry store of the membership: a particle needs at most one entry per
  // window, 32-candidate windows of its 3^(D-1) stencil rows (each holding at
  // most 3 * mppcell candidates), or in 3D FP32-split kWin3-candidate windows
  // of its 27 stencil cells.  GCLSPH_NMASK_CAP lowers the capacity (particles
  // beyond it scan their stencil instead).  Reallocation drops the stored
  // membership.
  template <typename R>
  void ensure_nmask(SphBufs<R>& B) {
    const int64_t mpp = std::max<int64_t>({hctl.mppcell, sys->mppcell, 1});
    const bool cells = dim == 3 && std::is_same<R, float>::value;  // cell_windows<3, R>()
    int64_t need = cells ? 27 * ((mpp + kWin3 - 1) / kWin3) : (dim == 2 ? 3 : 9) * ((3 * mpp + 31) / 32);
    if (const char* env = std::getenv("GCLSPH_NMASK_CAP")) {
      const long v = std::strtol(env, nullptr, 10);
      if (v >= 0 && v < need) need = v;
    }
    need = std::min<int64_t>(need, 4096);
    if (need == necap && ne_buf) return;
    GCL_HIP(hipStreamSynchronize(stream));
    if (ne_buf) GCL_HIP(hipFree(ne_buf));
    ne_buf = nullptr;
    GCL_HIP(hipMalloc(&ne_buf, sizeof(uint64_t) * size_t(std::max<int64_t>(need, 1)) * size_t(n)));
    necap = int(need);
    B.NE = ne_buf;
    B.necap = necap;
    masks_valid = false;
  }

  // Constants derived from the parameters, as the reference derives them
  // before each launch.
  SphConst make_const(double tl0, double tl1) const {
    SphConst C{};
    const double* p = par;
    C.dim = dim;
    C.gamma = int(p[SPH_GAMMA]);
    C.dx = p[SPH_DX];
    C.h = p[SPH_H];
    C.H = p[SPH_HH];
    C.Hinv = 1 / C.H;
    C.rho0 = p[SPH_RHO0];
    C.m0 = p[SPH_M0];
    C.alpha = p[SPH_ALPHA];
    C.c0 = p[SPH_C0];
    C.g = p[SPH_G];
    C.s = p[SPH_S];
    C.nu = p[SPH_NU];
    C.eta2 = (0.1 * C.h) * (0.1 * C.h);
    C.ginv = 1.0 / double(C.gamma);
    C.DDTkh = 2 * C.h * p[SPH_DELTA] * C.c0;
    C.Cb = (C.c0 * C.c0 * C.rho0) * C.ginv;
    C.DDTgz = C.rho0 * C.g / C.Cb;
    C.P0 = C.c0 * C.c0 * C.rho0 / double(C.gamma);
    const double norm = dim == 2 ? 7.0 / M_PI : 21.0 / (2.0 * M_PI);
    auto ipow = [](double x, int q) {
      double r = x;
      for (int k = 1; k < q; ++k) r = r * x;
      return r;
    };
    C.wnorm = norm * ipow(C.Hinv, dim);
    C.dwnorm = norm * ipow(C.Hinv, dim + 1);
    C.dist = sys->dist;
    C.dist2 = sys->dist * sys->dist;
    C.fb_d[0] = 0.4;
    C.fb_r0[0] = 2 * C.dx;
    C.fb_d[1] = p[SPH_BOUND_D];
    C.fb_r0[1] = p[SPH_BOUND_L];
    for (int k = 0; k < 2; ++k) C.fb_eta2[k] = 0.01 * (C.fb_r0[k] * C.fb_r0[k]);
    C.st_2h = 2 * C.h;
    C.st_c = 1.5 * M_PI;
    C.st_01h = 0.1 * C.h;
    C.dpc_l0 = p[SPH_DPC_L0];
    if (C.dpc_l0 > 0) {
      C.dpc_l0inv = 1 / C.dpc_l0;
      C.dpc_eta2 = (0.1 * C.dpc_l0) * (0.1 * C.dpc_l0);
      C.dpc_wnorm = norm * ipow(C.dpc_l0inv, dim);
      double um1 = 1.0 - 0.5;
      um1 = um1 * um1;
      um1 = um1 * um1;
      C.dpc_whinv = 1 / (((1.0 + 4.0 * 0.5) * um1) * C.dpc_wnorm);
      // root of 1 - W(u, 1) on [0, 1] (Roots.find_zero at src/sphproblem.jl:218)
      double w1 = norm;
      auto W1 = [&](double u) {
        if (!(u < 1.0)) return 0.0;
        double m = 1.0 - u;
        m = m * m;
        m = m * m;
        return ((1.0 + 4.0 * u) * m) * w1;
      };
      double lo = 0.0, hi = 1.0;
      for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (1.0 - W1(mid) < 0) lo = mid;
        else hi = mid;
      }
      C.dpc_lim = std::fabs(1.0 - W1(lo)) <= std::fabs(1.0 - W1(hi)) ? lo : hi;
    }
    C.dpc_lambda = p[SPH_DPC_LAMBDA];
    C.dpc_pmin = p[SPH_DPC_PMIN];
    C.dpc_pmax = p[SPH_DPC_PMAX];
    C.xsph_c = 2 * C.m0 * p[SPH_XSPH_EPS];
    C.dt_eta2 = (0.01 * C.h) * (0.01 * C.h);
    C.CFL = p[SPH_CFL];
    C.nui = p[SPH_NUI];
    C.tl0 = tl0;
    C.tl1 = tl1;
    C.f_h = float(C.h);
    C.f_Hinv = float(C.Hinv);
    C.f_rho0 = float(C.rho0);
    C.f_m0 = float(C.m0);
    C.f_alpha = float(C.alpha);
    C.f_c0 = float(C.c0);
    C.f_g = float(C.g);
    C.f_s = float(C.s);
    C.f_nu = float(C.nu);
    C.f_eta2 = float(C.eta2);
    C.f_DDTkh = float(C.DDTkh);
    C.f_DDTgz = float(C.DDTgz);
    C.f_P0 = float(C.P0);
    C.f_wnorm = float(C.wnorm);
    C.f_dwnorm = float(C.dwnorm);
    C.f_inv_rho0 = float(1.0 / C.rho0);
    C.f_st_turn = float(C.st_c / (2.0 * M_PI * C.st_2h));
    C.f_dwr = float(-20.0 * C.dwnorm * C.Hinv);
    for (int k = 0; k < 2; ++k) {
      C.f_fb_d[k] = float(C.fb_d[k]);
      C.f_fb_r0[k] = float(C.fb_r0[k]);
      C.f_fb_eta2[k] = float(C.fb_eta2[k]);
    }
    C.f_st_2h = float(C.st_2h);
    C.f_st_c = float(C.st_c);
    C.f_st_01h = float(C.st_01h);
    C.f_xsph_c = float(C.xsph_c);
    C.f_dpc_l0inv = float(C.dpc_l0inv);
    C.f_dpc_eta2 = float(C.dpc_eta2);
    C.f_dpc_whinv = float(C.dpc_whinv);
    C.f_dpc_lim = float(C.dpc_lim);
    C.f_dpc_wnorm = float(C.dpc_wnorm);
    C.f_dpc_lambda = float(C.dpc_lambda);
    C.f_dpc_pmin = float(C.dpc_pmin);
    C.f_dpc_pmax = float(C.dpc_pmax);
    for (int k = 0; k < 3; ++k) {
      C.f_fcs[k] = k < dim ? float(sys->cs[k]) : 0.0f;
      C.fcs[k] = double(C.f_fcs[k]);
    }
    return C;
  }

  // ---- timing
  hipEvent_t take_event() {
    if (ev_used == ev_pool.size()) {
      hipEvent_t e;
      GCL_HIP(hipEventCreate(&e));
      ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
  }
  hipEvent_t mark_begin(hipStream_t st) {
    if (!timing) return nullptr;
    hipEvent_t e = take_event();
    GCL_HIP(hipEventRecord(e, st));
    return e;
  }
  void mark_end(int stage, hipEvent_t b, hipStream_t st) {
    if (!timing) return;
    hipEvent_t e = take_event();
    GCL_HIP(hipEventRecord(e, st));
    ev_marks.push_back({stage, {b, e}});
  }
  void collect_timing() {
    for (auto& m : ev_marks) {
      float ms = 0;
      GCL_HIP(hipEventElapsedTime(&ms, m.second.first, m.second.second));
      stage_ms[m.first] += ms;
      stage_cnt[m.first] += 1;
    }
    ev_marks.clear();
    ev_used = 0;
  }

  void write_ctl() { write_ctl(stream); }
  void write_ctl(hipStream_t st) { GCL_HIP(hipMemcpyAsync(ctl, &hctl, sizeof(StepCtl), hipMemcpyHostToDevice, st)); }
  void read_ctl() {
    GCL_HIP(hipMemcpyAsync(&hctl, ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, stream));
    GCL_HIP(hipStreamSynchronize(stream));
  }

  template <typename R>
  SphBufs<R>& bufs();

  // Per-call switches of this problem (the per-step ones follow on the device).
  StepFlags call_flags(bool ts) const {
    const bool dpc_on = par[SPH_DPC_L0] > 0, xsph_on = par[SPH_XSPH_EPS] > 0;
    StepFlags F{};
    F.need_xrb = (par[SPH_NUI] > 0 || dpc_on || xsph_on) ? 1 : 0;
    F.need_cdx = par[SPH_NUI] > 0 ? 1 : 0;
    F.ts = ts ? 1 : 0;
    F.visc_in_pass2 = (ts && !dpc_on && !xsph_on) ? 1 : 0;
    F.nu_on = par[SPH_NU] > 0 ? 1 : 0;
    F.s_on = par[SPH_S] > 0 ? 1 : 0;
    F.write_xrb = F.need_xrb;
    F.dpc_on = dpc_on ? 1 : 0;
    F.xsph_on = xsph_on ? 1 : 0;
    const int64_t cspmn = int64_t(par[SPH_CSPMN]);  // src/sphproblem.jl:319
    F.cspmn = cspmn > 0 ? int(std::min<int64_t>(cspmn, std::numeric_limits<int>::max())) : 0;
    return F;
  }

  template <typename R>
  SceneDev<R> scene_desc(const SphConst& C, const StepFlags& F) {
    SceneDev<R> d{};
    d.B = bufs<R>();
    d.C = C;
    d.G = G;
    d.F = F;
    d.key = sb.key;
    d.slot = sb.slot;
    d.cnt = sb.cnt;
    d.start = sb.start;
    d.tmp = sb.tmp;
    d.perm = sb.perm;
    d.skey = sb.skey;
    d.maxcell = sb.maxcell;
    d.scan_tmp = sb.scan_tmp;
    d.ncells = G.ncells();
    return d;
  }

  // Descriptor of this problem for its own one-problem launch set.
  template <typename R>
  void upload_own(const SphConst& C, const StepFlags& F) {
    const SceneDev<R> d = scene_desc<R>(C, F);
    own_desc.resize(sizeof(d));
    std::memcpy(own_desc.data(), &d, sizeof(d));
    own.one = own_desc.data();
  }

  void stepsolve(int64_t nsteps, bool ts, double tl0, double tl1, int64_t* updaten, double* maxcdx);

  // Membership positions of the current pair list (host pointer choice).
  const double* membership(int k) {
    SphBufs<double>& Bd_ = Bd;
    double* const* X0 = split() ? Bf.X[0] : Bd_.X[0];
    double* const* X1 = split() ? Bf.X[1] : Bd_.X[1];
    double* const* XRB = split() ? Bf.XRB : Bd_.XRB;
    int p = hctl.parity;
    if (xrb_mode == 0) return XRB[k];
    if (xrb_mode == 1) return (p ? X0 : X1)[k];
    return (p ? X1 : X0)[k];
  }
  double* xcur(int k) {
    double* const* X = split() ? Bf.X[hctl.parity] : Bd.X[hctl.parity];
    return X[k];
  }
  const int32_t* idcur() { return split() ? Bf.ID[hctl.pparity] : Bd.ID[hctl.pparity]; }
  const int32_t* ptcur() { return split() ? Bf.PT[hctl.pparity] : Bd.PT[hctl.pparity]; }

  // update!(prob.system) called by the user between steps.
  template <int D, typename R>
  void user_rebuild_t() {
    SphBufs<R>& B = bufs<R>();
    read_ctl();
    ensure_nmask(B);
    hctl.rebuild = 1;
    hctl.synced = synced_prev ? 1 : 0;
    write_ctl();
    StepFlags F = call_flags(false);
    F.write_xrb = 1;
    upload_own<R>(make_const(0.0, 0.0), F);
    launch_rebuild<D, R>(own, stream);
    GCL_HIP(hipGetLastError());
    masks_valid = true;
    // per-particle arrays outside the reorder follow the new order
    auto permute_d = [&](double* a) {
      double* t = B.TMP;
      GCL_HIP(hipMemcpyAsync(t, a, 8 * size_t(n), hipMemcpyDeviceToDevice, stream));
      gather<double>(t, a, sb.perm, n, stream);
    };
    auto permute_r = [&](R* a) {
      R* t = reinterpret_cast<R*>(B.TMP);  // n doubles hold n elements of either precision
      GCL_HIP(hipMemcpyAsync(t, a, sizeof(R) * size_t(n), hipMemcpyDeviceToDevice, stream));
      gather<R>(t, a, sb.perm, n, stream);
    };
    for (int k = 0; k < D; ++k) {
      permute_d(B.CDX[k]);
      permute_d(B.SUMGW[k]);
      permute_r(B.ACC[k]);
      permute_r(B.DV[k]);
    }
    permute_d(B.SUMW);
    permute_r(B.DRHO);
    permute_r(B.PBUF);
    permute_r(B.PH);
    if (!synced_prev) {  // the reorder moved the half-step state into XH/VH/RH
      for (int k = 0; k < D; ++k) {
        GCL_HIP(hipMemcpyAsync(B.XH2[k], B.XH[k], 8 * size_t(n), hipMemcpyDeviceToDevice, stream));
        GCL_HIP(hipMemcpyAsync(B.VH2[k], B.VH[k], sizeof(R) * size_t(n), hipMemcpyDeviceToDevice, stream));
      }
      GCL_HIP(hipMemcpyAsync(B.RH2, B.RH, sizeof(R) * size_t(n), hipMemcpyDeviceToDevice, stream));
    }
    read_ctl();
    xrb_mode = 0;
    sys->mppcell = hctl.mppcell;
    sys->pairs_valid = false;
  }
  template <int D, typename R>
  void flush_diag_t() {
    upload_own<R>(make_const(0.0, 0.0), call_flags(false));
    launch_diag<D, R>(own, diag_xm, diag_masks, stream);
    GCL_HIP(hipGetLastError());
  }
  void flush_diag() {
    if (!diag_pending) return;
    diag_pending = false;
    if (split()) {
      if (dim == 2) flush_diag_t<2, float>();
      else flush_diag_t<3, float>();
    } else {
      if (dim == 2) flush_diag_t<2, double>();
      else flush_diag_t<3, double>();
    }
  }

  void user_rebuild() {
    flush_diag();
    if (split()) {
      if (dim == 2) user_rebuild_t<2, float>();
      else user_rebuild_t<3, float>();
    } else {
      if (dim == 2) user_rebuild_t<2, double>();
      else user_rebuild_t<3, double>();
    }
  }

  // Writing the points of an attached cell list leaves the half-step copy
  // untouched (the reference's xΔt½ is a separate array).
  void materialize_halfstep() {
    if (!synced_prev) return;
    int p = hctl.parity;
    for (int k = 0; k < dim; ++k) {
      if (split()) {
        GCL_HIP(hipMemcpyAsync(Bf.XH2[k], Bf.X[p][k], 8 * size_t(n), hipMemcpyDeviceToDevice, stream));
        GCL_HIP(hipMemcpyAsync(Bf.VH2[k], Bf.V[p][k], 4 * size_t(n), hipMemcpyDeviceToDevice, stream));
      } else {
        GCL_HIP(hipMemcpyAsync(Bd.XH2[k], Bd.X[p][k], 8 * size_t(n), hipMemcpyDeviceToDevice, stream));
        GCL_HIP(hipMemcpyAsync(Bd.VH2[k], Bd.V[p][k], 8 * size_t(n), hipMemcpyDeviceToDevice, stream));
      }
    }
    if (split()) GCL_HIP(hipMemcpyAsync(Bf.RH2, Bf.RHO[p], 4 * size_t(n), hipMemcpyDeviceToDevice, stream));
    else GCL_HIP(hipMemcpyAsync(Bd.RH2, Bd.RHO[p], 8 * size_t(n), hipMemcpyDeviceToDevice, stream));
    synced_prev = false;
  }
};

template <>
SphBufs<double>& sph_s::bufs<double>() {
  return Bd;
}
template <>
SphBufs<float>& sph_s::bufs<float>() {
  return Bf;
}

void LaunchSet::build(const int64_t* n, const int64_t* ncells, unsigned cnt, size_t bytes) {
  release();
  count = cnt;
  scene_bytes = bytes;
  std::vector<int2> pm, cm, pm512;
  for (unsigned i = 0; i < cnt; ++i) {
    const int64_t nb = (n[i] + kBlock - 1) / kBlock;
    for (int64_t b = 0; b < nb; ++b) pm.push_back(make_int2(int(i), int(b)));
    const int64_t nb512 = (n[i] + 511) / 512;  // the pass kernels' 512-thread blocks
    for (int64_t b = 0; b < nb512; ++b) pm512.push_back(make_int2(int(i), int(b)));
    const int64_t nt = (ncells[i] + 1 + 1023) / 1024;  // cnt[0..ncells] in 1024-cell tiles
    for (int64_t t = 0; t < nt; ++t) cm.push_back(make_int2(int(i), int(t)));
  }
  if (pm.empty()) pm.push_back(make_int2(0, 0));  // a no-op block (a >= n everywhere)
  if (pm512.empty()) pm512.push_back(make_int2(0, 0));
  pblocks = unsigned(pm.size());
  pblocks512 = unsigned(pm512.size());
  ctiles = unsigned(cm.size());
  GCL_HIP(hipMalloc(&scenes, bytes * cnt));
  GCL_HIP(hipMalloc(&pmap, sizeof(int2) * pm.size()));
  GCL_HIP(hipMalloc(&pmap512, sizeof(int2) * pm512.size()));
  GCL_HIP(hipMalloc(&cmap, sizeof(int2) * cm.size()));
  GCL_HIP(hipMemcpy(pmap, pm.data(), sizeof(int2) * pm.size(), hipMemcpyHostToDevice));
  GCL_HIP(hipMemcpy(pmap512, pm512.data(), sizeof(int2) * pm512.size(), hipMemcpyHostToDevice));
  GCL_HIP(hipMemcpy(cmap, cm.data(), sizeof(int2) * cm.size(), hipMemcpyHostToDevice));
}

void LaunchSet::release() {
  if (step_graph) (void)hipGraphExecDestroy(step_graph);
  step_graph = nullptr;
  graph_key.clear();
  if (scenes) (void)hipFree(scenes);
  if (pmap) (void)hipFree(pmap);
  if (pmap512) (void)hipFree(pmap512);
  if (cmap) (void)hipFree(cmap);
  scenes = nullptr;
  pmap = cmap = pmap512 = nullptr;
  pblocks = ctiles = count = pblocks512 = 0;
}

namespace {

// `stepsolve!` (src/sphproblem.jl:207-345) for count problems stepped
// together: every stage is one launch over all of them, each problem keeping
// its own control block (Δt, rebuilds, time, corrections).  The problems'
// results are those of stepping each on its own.
template <int D, typename R>
void run_steps(sph_s* const* P, unsigned count, LaunchSet& L, hipStream_t st, int64_t nsteps, bool ts,
               const double* tl0, const double* tl1, int64_t* updaten, double* maxcdx) {
  GCL_CHECK(nsteps <= std::numeric_limits<int>::max(), kErrArgument, "too many steps in one call");
  std::vector<SceneDev<R>> descs(count);
  std::vector<StepFlags> flags(count);
  bool any_dpc = false, any_xsph = false, any_dtv = false;
  for (unsigned i = 0; i < count; ++i) {
    sph_s* p = P[i];
    SphBufs<R>& B = p->bufs<R>();
    p->diag_pending = false;  // replaced by this call's
    p->ensure_nmask(B);
    const StepFlags F = p->call_flags(ts);
    flags[i] = F;
    descs[i] = p->scene_desc<R>(p->make_const(tl0[i], tl1[i]), F);
    any_dpc |= F.dpc_on != 0;
    any_xsph |= F.xsph_on != 0;
    any_dtv |= F.ts && !F.visc_in_pass2;
    StepCtl& h = p->hctl;
    h.rebuild = 1;  // the first iteration of every call rebuilds (src/sphproblem.jl:212,222)
    h.updaten = 0;
    h.maxcdx_out = 0;
    h.dt = p->par[SPH_DT];
    h.etime = p->par[SPH_ETIME];
    h.red_cdx = h.red_acc = h.red_visc = 0;
    h.iter = 1;
    h.nsteps = int(nsteps);
    h.synced = p->synced_prev ? 1 : 0;
    p->write_ctl(st);
  }
  // one problem: the descriptor goes by value with every launch
  if (count == 1) L.one = descs.data();
  else GCL_HIP(hipMemcpyAsync(L.scenes, descs.data(), sizeof(SceneDev<R>) * count, hipMemcpyHostToDevice, st));
  sph_s* timer = (count == 1 && P[0]->timing) ? P[0] : nullptr;
  bool any_cspmn = false;
  for (unsigned i = 0; i < count; ++i) any_cspmn |= flags[i].cspmn > 0;
  // Pass variants (same results, different code generation): forced by
  // GCLSPH_PASS_VARIANT=<pass1>,<pass2>, else the ones this launch set chose by
  // timing, else the fastest measured on the dam-break scenes per dimension
  // and precision until it has been tuned.
  int pv[2];
  if (std::is_same<R, float>::value) {
    pv[0] = 2;  // prefetch, >= 3 waves per SIMD
    pv[1] = 0;  // prefetch, compiler's occupancy
  } else if (D == 2) {
    pv[0] = pv[1] = 2;
  } else {
    pv[0] = 0;
    pv[1] = 1;
  }
  bool forced = false;
  if (const char* env = std::getenv("GCLSPH_PASS_VARIANT")) {
    int v1 = -1, v2 = -1;
    if (std::sscanf(env, "%d,%d", &v1, &v2) == 2 && v1 >= 0 && v1 < kPassVariants && v2 >= 0 &&
        v2 < kPassVariants) {
      pv[0] = v1;
      pv[1] = v2;
      forced = true;
    }
  }
  if (!forced && L.pass_variant[0] >= 0) {
    pv[0] = L.pass_variant[0];
    pv[1] = L.pass_variant[1];
  }
  // one step's launches; with the corrections of every problem that may need
  // them (each kernel returns at once for a problem that does not)
  auto launch_step = [&](bool cspm, int v1, int v2, hipEvent_t* ev) {
    launch_rebuild<D, R>(L, st);
    if (ev) GCL_HIP(hipEventRecord(ev[0], st));
    launch_pass1<D, R>(L, st, v1);
    if (ev) GCL_HIP(hipEventRecord(ev[1], st));
    launch_pass2<D, R>(L, st, v2);
    if (ev) GCL_HIP(hipEventRecord(ev[2], st));
    if (any_dpc) launch_dpc<D, R>(L, st);
    if (any_xsph) launch_xsph<D, R>(L, st);
    if (cspm) launch_cspm<D, R>(L, st);
    if (any_dtv) launch_dt_visc<D, R>(L, st);
    launch_finalize<D, R>(L, st);
    GCL_HIP(hipGetLastError());
  };
  auto cspm_at = [&](int64_t it) {
    bool any = false;
    for (unsigned i = 0; i < count; ++i) any |= flags[i].cspmn > 0 && it % flags[i].cspmn == 0;
    return any;
  };
  int64_t done = 0;
  const char* no_tune = std::getenv("GCLSPH_AUTOTUNE");
  if (!forced && L.pass_variant[0] < 0 && !timer && nsteps > kTuneTimed && !(no_tune && no_tune[0] == '0')) {
    // Autotune on the first steps of the call (live steps: every variant
    // computes the same values).  Up to kTuneWarm untimed steps first, as
    // many as the call has beyond the kTuneTimed timed ones: the first dozen
    // steps of a run on a fresh device run up to 10 % slower while it warms
    // up, which would favour the variants timed late.  Then both passes are
    // timed with events: the kPassCodegen code-generation variants at each
    // block size, then the fastest of them per pass with the occupancy cap
    // and the two fastest once more (timings of variants within a few per
    // cent of each other are noisy; a variant's time is its shortest).  The
    // steps of each stage are queued back to back, as in production, and
    // read after the stage.  The variant of each pass with the shortest time
    // is kept for this launch set.
    const int warm = int(std::min<int64_t>(kTuneWarm, nsteps - kTuneTimed));
    std::vector<hipEvent_t> ev(3 * (kTuneWarm + kTuneTimed), nullptr);
    for (auto& e : ev) GCL_HIP(hipEventCreate(&e));
    std::vector<float> tv[2] = {std::vector<float>(kPassVariants, -1.0f), std::vector<float>(kPassVariants, -1.0f)};
    std::vector<std::pair<int, int>> queued;  // (step, variant pass 1 * kPassVariants + variant pass 2)
    auto time_step = [&](int& k, int v1, int v2) {
      queued.emplace_back(k, v1 * kPassVariants + v2);
      launch_step(cspm_at(k + 1), v1, v2, &ev[3 * size_t(k)]);
      ++k;
    };
    auto read_stage = [&] {
      GCL_HIP(hipStreamSynchronize(st));
      for (const auto& q : queued) {
        const size_t k = size_t(q.first);
        float t[2] = {0, 0};
        GCL_HIP(hipEventElapsedTime(&t[0], ev[3 * k], ev[3 * k + 1]));
        GCL_HIP(hipEventElapsedTime(&t[1], ev[3 * k + 1], ev[3 * k + 2]));
        const int v[2] = {q.second / kPassVariants, q.second % kPassVariants};
        for (int p = 0; p < 2; ++p)
          if (tv[p][v[p]] < 0 || t[p] < tv[p][v[p]]) tv[p][v[p]] = t[p];
      }
      queued.clear();
    };
    // the fastest (rank 0) or second fastest (rank 1) timed variant of pass p
    auto ranked = [&](int p, int rank) {
      int b0 = -1, b1 = -1;
      for (int v = 0; v < kPassVariants; ++v) {
        if (tv[p][v] < 0) continue;
        if (b0 < 0 || tv[p][v] < tv[p][b0]) {
          b1 = b0;
          b0 = v;
        } else if (b1 < 0 || tv[p][v] < tv[p][b1]) {
          b1 = v;
        }
      }
      return rank == 0 || b1 < 0 ? b0 : b1;
    };
    const int nocap = kPassCodegen * kPassBlockSizes;
    int k = 0;
    for (; k < warm; ++k) launch_step(cspm_at(k + 1), pv[0], pv[1], nullptr);
    for (int v = 0; v < nocap; ++v) time_step(k, v, v);
    read_stage();
    const int r0[2] = {ranked(0, 0), ranked(1, 0)}, r1[2] = {ranked(0, 1), ranked(1, 1)};
    time_step(k, r0[0] % nocap + nocap, r0[1] % nocap + nocap);
    time_step(k, r0[0], r0[1]);
    time_step(k, r1[0], r1[1]);
    read_stage();
    for (auto& e : ev) (void)hipEventDestroy(e);
    pv[0] = ranked(0, 0);
    pv[1] = ranked(1, 0);
    L.pass_variant[0] = pv[0];
    L.pass_variant[1] = pv[1];
    done = k;
  }
  const char* no_graph = std::getenv("GCLSPH_NO_GRAPH");
  const bool graph = !timer && nsteps - done >= 2 && !(no_graph && no_graph[0] == '1');
  if (graph) {
    // The step as a captured graph, replayed for the remaining steps: every
    // per-step decision lives in the control blocks, so one capture serves all
    // steps.  Kept for the next call while the launches, pass variants and
    // one-problem descriptor (baked into the graph) are unchanged.
    std::vector<unsigned char> key = {uint8_t(any_dpc), uint8_t(any_xsph), uint8_t(any_cspmn), uint8_t(any_dtv),
                                      uint8_t(pv[0]), uint8_t(pv[1])};
    if (L.one) {
      const unsigned char* d = static_cast<const unsigned char*>(L.one);
      key.insert(key.end(), d, d + sizeof(SceneDev<R>));
    }
    if (!L.step_graph || key != L.graph_key) {
      if (L.step_graph) GCL_HIP(hipGraphExecDestroy(L.step_graph));
      L.step_graph = nullptr;
      hipGraph_t g = nullptr;
      GCL_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      launch_step(any_cspmn, pv[0], pv[1], nullptr);
      GCL_HIP(hipStreamEndCapture(st, &g));
      const hipError_t e = hipGraphInstantiate(&L.step_graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      GCL_HIP(e);
      L.graph_key = std::move(key);
    }
    for (int64_t it = done + 1; it <= nsteps; ++it) GCL_HIP(hipGraphLaunch(L.step_graph, st));
  } else {
    auto mark_begin = [&]() { return timer ? timer->mark_begin(st) : hipEvent_t(nullptr); };
    auto mark_end = [&](int stage, hipEvent_t e) {
      if (timer) timer->mark_end(stage, e, st);
    };
    for (int64_t it = done + 1; it <= nsteps; ++it) {
      const bool any_cspm = cspm_at(it);
      if (!timer) {
        launch_step(any_cspm, pv[0], pv[1], nullptr);
        continue;
      }
      hipEvent_t e0 = mark_begin();
      launch_rebuild<D, R>(L, st);
      GCL_HIP(hipGetLastError());
      mark_end(0, e0);
      e0 = mark_begin();
      launch_pass1<D, R>(L, st, pv[0]);
      GCL_HIP(hipGetLastError());
      mark_end(1, e0);
      e0 = mark_begin();
      launch_pass2<D, R>(L, st, pv[1]);
      GCL_HIP(hipGetLastError());
      mark_end(2, e0);
      if (any_dpc || any_xsph || any_cspm || any_dtv) {
        e0 = mark_begin();
        if (any_dpc) launch_dpc<D, R>(L, st);
        if (any_xsph) launch_xsph<D, R>(L, st);
        if (any_cspm) launch_cspm<D, R>(L, st);
        if (any_dtv) launch_dt_visc<D, R>(L, st);
        GCL_HIP(hipGetLastError());
        mark_end(3, e0);
      }
      e0 = mark_begin();
      launch_finalize<D, R>(L, st);
      GCL_HIP(hipGetLastError());
      mark_end(4, e0);
    }
  }
  // host mirror of the device's half-step sync flag after the last step
  for (unsigned i = 0; i < count && nsteps > 0; ++i) {
    const StepFlags& F = flags[i];
    const bool cspm_fire = F.cspmn > 0 && nsteps % F.cspmn == 0;
    P[i]->synced_prev = !(F.dpc_on || F.xsph_on || cspm_fire);
  }
  for (unsigned i = 0; i < count; ++i) {
    sph_s* p = P[i];
    if (nsteps > 0) {
      p->xrb_mode = flags[i].need_xrb ? 0 : 1;
      p->p_from_dpc = flags[i].dpc_on != 0;
      p->masks_valid = true;
    }
    p->diag_pending = true;
    p->diag_xm = p->xrb_mode;
    p->diag_masks = p->masks_valid ? 1 : 0;
    GCL_HIP(hipMemcpyAsync(&p->hctl, p->ctl, sizeof(StepCtl), hipMemcpyDeviceToHost, st));
  }
  GCL_HIP(hipStreamSynchronize(st));
  if (count == 1) L.one = nullptr;
  if (timer) timer->collect_timing();
  for (unsigned i = 0; i < count; ++i) {
    sph_s* p = P[i];
    p->par[SPH_DT] = p->hctl.dt;
    p->par[SPH_ETIME] = p->hctl.etime;
    updaten[i] = p->hctl.updaten;
    maxcdx[i] = p->hctl.maxcdx_out;
    p->sys->mppcell = p->hctl.mppcell;
    p->sys->pairs_valid = false;
  }
}

template <int D, typename R>
void run_steps_one(sph_s* p, int64_t nsteps, bool ts, double tl0, double tl1, int64_t* updaten, double* maxcdx) {
  run_steps<D, R>(&p, 1, p->own, p->stream, nsteps, ts, &tl0, &tl1, updaten, maxcdx);
}

}  // namespace

void sph_s::stepsolve(int64_t nsteps, bool ts, double tl0, double tl1, int64_t* updaten, double* maxcdx) {
  if (split()) {
    if (dim == 2) run_steps_one<2, float>(this, nsteps, ts, tl0, tl1, updaten, maxcdx);
    else run_steps_one<3, float>(this, nsteps, ts, tl0, tl1, updaten, maxcdx);
  } else {
    if (dim == 2) run_steps_one<2, double>(this, nsteps, ts, tl0, tl1, updaten, maxcdx);
    else run_steps_one<3, double>(this, nsteps, ts, tl0, tl1, updaten, maxcdx);
  }
}

// A set of problems stepped together (one launch per stage for all of them).
struct sph_batch_s {
  std::vector<sph_s*> probs;
  LaunchSet L;
  hipStream_t stream = nullptr;
  ~sph_batch_s() {
    for (sph_s* p : probs) p->batches -= 1;
    L.release();
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Host copy of a sorted device array, returned in caller order.
template <typename T>
std::vector<double> unsort(const T* dev, const int32_t* id_dev, int64_t n, hipStream_t s, double add = 0.0) {
  std::vector<T> v(static_cast<size_t>(n));
  std::vector<int32_t> id(static_cast<size_t>(n));
  GCL_HIP(hipMemcpyAsync(v.data(), dev, sizeof(T) * size_t(n), hipMemcpyDeviceToHost, s));
  GCL_HIP(hipMemcpyAsync(id.data(), id_dev, 4 * size_t(n), hipMemcpyDeviceToHost, s));
  GCL_HIP(hipStreamSynchronize(s));
  std::vector<double> out(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) out[size_t(id[size_t(i)])] = double(v[size_t(i)]) + add;
  return out;
}

// Lazily (re)builds the pair list of a cell list owned by an SPH problem:
// pairs at the positions of the last rebuild, with the reference's capacity rule.
void ensure_pairs(gcl_s* g) {
  if (!g->attached || g->pairs_valid) return;
  sph_s* P = g->attached;
  P->read_ctl();
  const void* xm[3] = {P->membership(0), P->membership(1), P->membership(g->dim == 3 ? 2 : 1)};
  int64_t total = 0;
  pair_count(P->G, 64, xm, P->sb.start, P->sb.skey, g->n, g->dist, g->pb, P->stream);
  GCL_HIP(hipMemcpyAsync(&total, g->pb.off + g->n, 8, hipMemcpyDeviceToHost, P->stream));
  GCL_HIP(hipStreamSynchronize(P->stream));
  int64_t cap = g->pb.capacity;
  if (g->pairsn > cap || double(g->pairsn) < double(cap) * 0.6) cap = ceil_div08(g->pairsn);
  if (total > cap) cap = ceil_div08(total);
  if (cap != g->pb.capacity) g->pb.set_capacity(cap, P->stream);
  else g->pb.zero_from(total, P->stream);
  pair_emit(P->G, 64, xm, P->sb.start, P->sb.skey, P->idcur(), g->n, g->dist, total, g->pb, P->stream);
  GCL_HIP(hipStreamSynchronize(P->stream));
  g->pairsn = total;
  g->pairs_valid = true;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

const char* gclsph_last_error(void) { return g_last_error.c_str(); }
const char* gclsph_version(void) { return "gclsph 0.1 (gfx950)"; }

int gcl_create(int dim, int prec, int64_t n, const double* const* pts, const double* cellsize, double dist,
               const double* min_or_null, const double* max_or_null, int64_t mppcell, int64_t mpairs, int* warn_flags,
               gcl_t* out) {
  return guard([&] {
    GCL_CHECK(out != nullptr, kErrArgument, "out handle is NULL");
    *out = nullptr;
    GCL_CHECK(dim == 2 || dim == 3, kErrArgument, "wrong dimention");  // src/celllist.jl:26
    GCL_CHECK(prec == GCLSPH_F64 || prec == GCLSPH_F32, kErrArgument, "cell list precision must be F64 or F32");
    GCL_CHECK(n > 0 && n < (int64_t(1) << 31) - 1, kErrArgument, "number of points must be in [1, 2^31-2]");
    GCL_CHECK(pts != nullptr && cellsize != nullptr, kErrArgument, "points / cellsize is NULL");
    for (int k = 0; k < dim; ++k) GCL_CHECK(pts[k] != nullptr, kErrArgument, "points component is NULL");
    require_device();
    std::unique_ptr<gcl_s> g(new gcl_s());
    g->dim = dim;
    g->prec = prec;
    g->n = n;
    int warn = 0;
    // T(dist), T.(cellsize); a cell smaller than dist is raised to dist (:27-39)
    double distT = prec == 32 ? double(float(dist)) : dist;
    g->dist = distT;
    for (int k = 0; k < dim; ++k) {
      double c = prec == 32 ? double(float(cellsize[k])) : cellsize[k];
      if (c < distT) {
        c = distT;
        warn |= 1;
      }
      g->cs[k] = c;
    }
    const double sq = std::sqrt(std::numeric_limits<double>::epsilon());
    g->G.dim = dim;
    for (int k = 0; k < 3; ++k) g->G.grid[k] = 1;
    for (int k = 0; k < dim; ++k) {
      double mn, mx;
      if (min_or_null) mn = min_or_null[k];
      else {
        mn = std::numeric_limits<double>::infinity();
        for (int64_t i = 0; i < n; ++i) {
          double v = prec == 32 ? double(float(pts[k][i])) : pts[k][i];
          mn = v < mn ? v : mn;
        }
      }
      if (max_or_null) mx = max_or_null[k];
      else {
        mx = -std::numeric_limits<double>::infinity();
        for (int64_t i = 0; i < n; ++i) {
          double v = prec == 32 ? double(float(pts[k][i])) : pts[k][i];
          mx = v > mx ? v : mx;
        }
      }
      mn = mn - std::fabs((mn + sq) * sq);  // :42
      double cells = std::ceil((mx - mn) / g->cs[k]);  // :44-46
      GCL_CHECK(cells >= 1 && cells < 2.0e9, kErrArgument, "cell grid has no cells (or too many) along a dimension");
      g->G.grid[k] = int64_t(cells);
      g->G.off64[k] = mn;
      g->G.off32[k] = float(mn);
      if (prec == 32) {
        g->G.hinv32[k] = 1.0f / float(g->cs[k]);
        g->G.hinv64[k] = double(g->G.hinv32[k]);
      } else {
        g->G.hinv64[k] = 1.0 / g->cs[k];
        g->G.hinv32[k] = float(g->G.hinv64[k]);
      }
    }
    GCL_CHECK(g->G.ncells() < (int64_t(1) << 31) - 2, kErrArgument, "cell grid exceeds 2^31 cells");
    GCL_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    for (int k = 0; k < dim; ++k) {
      GCL_HIP(hipMalloc(&g->xo[k], g->tsize() * size_t(n)));
      GCL_HIP(hipMalloc(&g->xs[k], g->tsize() * size_t(n + kPosPad)));  // pair windows read past the end
    }
    g->upload_points(pts);
    g->sb.alloc(n, g->G.ncells());
    g->pb.alloc(n, dim == 2 && n <= kPairEntriesMaxPoints ? kPairEntries2D : 0);
    // cellmap! with the Float64 MIN (:55), then fillcells_naive! (:66)
    bin_sort(g->G, prec, g->xo, n, true, g->sb, nullptr, g->stream, nullptr, nullptr, g->xs);
    int32_t maxc = g->read_maxcell();
    int64_t maxpoint = int64_t(std::ceil(double(maxc) * 1.05 + 1));  // :57-58
    g->mppcell = mppcell < maxpoint ? maxpoint : mppcell;
    if (mpairs == 0) {  // мaxpairs (:69-71)
      unsigned long long* d = nullptr;
      GCL_HIP(hipMalloc(&d, 8));
      GCL_HIP(hipMemsetAsync(d, 0, 8, g->stream));
      k_maxpairs<<<grid_for(g->G.ncells()), kBlock, 0, g->stream>>>(g->sb.cnt, g->G, d);
      GCL_HIP(hipGetLastError());
      unsigned long long v = 0;
      GCL_HIP(hipMemcpyAsync(&v, d, 8, hipMemcpyDeviceToHost, g->stream));
      GCL_HIP(hipStreamSynchronize(g->stream));
      (void)hipFree(d);
      mpairs = int64_t(v);
    }
    int64_t total = g->count_pairs(g->xs);
    if (total > mpairs) warn |= 2;  // :77-79
    int64_t cap = mpairs;
    if (double(total) < double(mpairs) * 0.8) cap = ceil_div08(total);  // :80-86
    g->pb.set_capacity(cap, g->stream);
    pair_emit(g->G, prec, g->xs, g->sb.start, g->sb.skey, g->sb.perm, n, g->dist, total, g->pb, g->stream);
    GCL_HIP(hipStreamSynchronize(g->stream));
    g->pairsn = total;
    if (warn_flags) *warn_flags = warn;
    *out = g.release();
  });
}

int gcl_destroy(gcl_t h) {
  return guard([&] {
    GCL_CHECK(h != nullptr, kErrArgument, "NULL handle");
    GCL_CHECK(h->attached == nullptr, kErrState, "cell list is owned by an SPH problem; destroy the problem first");
    delete h;
  });
}

int gcl_update(gcl_t h, int fillzero) {
  return guard([&] {
    GCL_CHECK(h != nullptr, kErrArgument, "NULL handle");
    if (h->attached) {
      h->attached->user_rebuild();
      ensure_pairs(h);
    } else {
      h->update(fillzero != 0);
    }
  });
}

int gcl_partialupdate(gcl_t h, int fillzero) {
  return guard([&] {
    GCL_CHECK(h != nullptr, kErrArgument, "NULL handle");
    if (h->attached) {
      sph_s* P = h->attached;
      P->read_ctl();
      const void* xc[3] = {P->xcur(0), P->xcur(1), P->xcur(h->dim == 3 ? 2 : 1)};
      pair_count(P->G, 64, xc, P->sb.start, P->sb.skey, h->n, h->dist, h->pb, P->stream);
      int64_t total = 0;
      GCL_HIP(hipMemcpyAsync(&total, h->pb.off + h->n, 8, hipMemcpyDeviceToHost, P->stream));
      GCL_HIP(hipStreamSynchronize(P->stream));
      if (fillzero) h->pb.zero_from(total, P->stream);
      pair_emit(P->G, 64, xc, P->sb.start, P->sb.skey, P->idcur(), h->n, h->dist, total, h->pb, P->stream);
      GCL_HIP(hipStreamSynchronize(P->stream));
      h->pairsn = total;
      h->pairs_valid = true;
    } else {
      h->partialupdate(fillzero != 0);
    }
  });
}

int gcl_set_points(gcl_t h, const double* const* pts) {
  return guard([&] {
    GCL_CHECK(h != nullptr && pts != nullptr, kErrArgument, "NULL argument");
    if (h->attached) {
      sph_s* P = h->attached;
      P->read_ctl();
      P->flush_diag();
      P->materialize_halfstep();
      double* tmp = nullptr;
      GCL_HIP(hipMalloc(&tmp, 8 * size_t(h->n)));
      for (int k = 0; k < h->dim; ++k) {
        GCL_HIP(hipMemcpyAsync(tmp, pts[k], 8 * size_t(h->n), hipMemcpyHostToDevice, P->stream));
        k_scatter_by<double><<<grid_for(h->n), kBlock, 0, P->stream>>>(tmp, P->xcur(k), P->idcur(), h->n);
        GCL_HIP(hipGetLastError());
      }
      GCL_HIP(hipStreamSynchronize(P->stream));
      (void)hipFree(tmp);
    } else {
      h->upload_points(pts);
    }
  });
}

int gcl_get_points(gcl_t h, double* const* out) {
  return guard([&] {
    GCL_CHECK(h != nullptr && out != nullptr, kErrArgument, "NULL argument");
    if (h->attached) {
      sph_s* P = h->attached;
      P->read_ctl();
      for (int k = 0; k < h->dim; ++k) {
        std::vector<double> v = unsort<double>(P->xcur(k), P->idcur(), h->n, P->stream);
        std::memcpy(out[k], v.data(), 8 * size_t(h->n));
      }
    } else {
      for (int k = 0; k < h->dim; ++k) {
        if (h->prec == 32) {
          std::vector<float> v(static_cast<size_t>(h->n));
          GCL_HIP(hipMemcpy(v.data(), h->xo[k], 4 * size_t(h->n), hipMemcpyDeviceToHost));
          for (int64_t i = 0; i < h->n; ++i) out[k][i] = v[size_t(i)];
        } else {
          GCL_HIP(hipMemcpy(out[k], h->xo[k], 8 * size_t(h->n), hipMemcpyDeviceToHost));
        }
      }
    }
  });
}

int gcl_pairsn(gcl_t h, int64_t* out) {
  return guard([&] {
    GCL_CHECK(h != nullptr && out != nullptr, kErrArgument, "NULL argument");
    ensure_pairs(h);
    *out = h->pairsn;
  });
}

int gcl_capacity(gcl_t h, int64_t* out) {
  return guard([&] {
    GCL_CHECK(h != nullptr && out != nullptr, kErrArgument, "NULL argument");
    ensure_pairs(h);
    *out = h->pb.capacity;
  });
}

int gcl_get_pairs(gcl_t h, int32_t* ij, int64_t cap, int64_t* written) {
  return guard([&] {
    GCL_CHECK(h != nullptr && ij != nullptr, kErrArgument, "NULL argument");
    ensure_pairs(h);
    int64_t c = std::min<int64_t>(cap, h->pb.capacity);
    hipStream_t s = h->attached ? h->attached->stream : h->stream;
    if (c > 0) GCL_HIP(hipMemcpyAsync(ij, h->pb.pairs, 8 * size_t(c), hipMemcpyDeviceToHost, s));
    GCL_HIP(hipStreamSynchronize(s));
    if (written) *written = c;
  });
}

int gcl_info(gcl_t h, gcl_info_t* out) {
  return guard([&] {
    GCL_CHECK(h != nullptr && out != nullptr, kErrArgument, "NULL argument");
    if (h->attached) {
      h->attached->read_ctl();
      h->mppcell = h->attached->hctl.mppcell;
    }
    std::memset(out, 0, sizeof(*out));
    out->dim = h->dim;
    out->prec = h->prec;
    out->n = h->n;
    for (int k = 0; k < 3; ++k) {
      out->grid[k] = k < h->dim ? h->G.grid[k] : 1;
      out->offset[k] = k < h->dim ? (h->prec == 32 ? double(h->G.off32[k]) : h->G.off64[k]) : 0.0;
      out->cellsize[k] = k < h->dim ? h->cs[k] : 0.0;
    }
    out->dist = h->dist;
    out->mppcell = h->mppcell;
  });
}

int gcl_get_cellpnum(gcl_t h, int32_t* out) {
  return guard([&] {
    GCL_CHECK(h != nullptr && out != nullptr, kErrArgument, "NULL argument");
    SortBuffers& sb = h->attached ? h->attached->sb : h->sb;
    hipStream_t s = h->attached ? h->attached->stream : h->stream;
    GCL_HIP(hipMemcpyAsync(out, sb.cnt, 4 * size_t(h->G.ncells()), hipMemcpyDeviceToHost, s));
    GCL_HIP(hipStreamSynchronize(s));
  });
}

int gcl_get_pcell(gcl_t h, int32_t* out) {
  return guard([&] {
    GCL_CHECK(h != nullptr && out != nullptr, kErrArgument, "NULL argument");
    SortBuffers& sb = h->attached ? h->attached->sb : h->sb;
    hipStream_t s = h->attached ? h->attached->stream : h->stream;
    if (h->attached) h->attached->read_ctl();
    const int32_t* id = h->attached ? h->attached->idcur() : sb.perm;
    std::vector<int32_t> skey(size_t(h->n)), idh(size_t(h->n));
    GCL_HIP(hipMemcpyAsync(skey.data(), sb.skey, 4 * size_t(h->n), hipMemcpyDeviceToHost, s));
    GCL_HIP(hipMemcpyAsync(idh.data(), id, 4 * size_t(h->n), hipMemcpyDeviceToHost, s));
    GCL_HIP(hipStreamSynchronize(s));
    int64_t gx = h->G.grid[0], gy = h->G.grid[1];
    for (int64_t s_ = 0; s_ < h->n; ++s_) {
      int64_t c = skey[size_t(s_)], i = idh[size_t(s_)];
      out[i] = int32_t(c % gx + 1);
      out[h->n + i] = int32_t((c / gx) % gy + 1);
      if (h->dim == 3) out[2 * h->n + i] = int32_t(c / (gx * gy) + 1);
    }
  });
}

int gcl_maxpairs(gcl_t h, int64_t* out) {
  return guard([&] {
    GCL_CHECK(h != nullptr && out != nullptr, kErrArgument, "NULL argument");
    SortBuffers& sb = h->attached ? h->attached->sb : h->sb;
    hipStream_t s = h->attached ? h->attached->stream : h->stream;
    unsigned long long* d = nullptr;
    GCL_HIP(hipMalloc(&d, 8));
    GCL_HIP(hipMemsetAsync(d, 0, 8, s));
    k_maxpairs<<<grid_for(h->G.ncells()), kBlock, 0, s>>>(sb.cnt, h->G, d);
    GCL_HIP(hipGetLastError());
    unsigned long long v = 0;
    GCL_HIP(hipMemcpyAsync(&v, d, 8, hipMemcpyDeviceToHost, s));
    GCL_HIP(hipStreamSynchronize(s));
    (void)hipFree(d);
    *out = int64_t(v);
  });
}

// ---------------------------------------------------------------- SPHProblem
int sph_create(gcl_t system, int prec, const double* params, const double* rho, const int32_t* ptype,
               const double* const* v_or_null, sph_t* out) {
  return guard([&] {
    GCL_CHECK(out != nullptr, kErrArgument, "out handle is NULL");
    *out = nullptr;
    GCL_CHECK(system != nullptr && params != nullptr && rho != nullptr && ptype != nullptr, kErrArgument,
              "NULL argument");
    GCL_CHECK(system->prec == GCLSPH_F64, kErrUnsupported,
              "SPHProblem needs a Float64 cell list (use GCLSPH_F32SPLIT for single-precision pair sums)");
    GCL_CHECK(system->attached == nullptr, kErrState, "cell list already belongs to an SPH problem");
    GCL_CHECK(prec == GCLSPH_F64 || prec == GCLSPH_F32SPLIT, kErrArgument, "SPH precision must be F64 or F32SPLIT");
    double g = params[SPH_GAMMA];
    GCL_CHECK(g >= 1 && g <= 64 && std::floor(g) == g, kErrArgument, "gamma must be an integer in [1, 64]");
    GCL_CHECK(prec != GCLSPH_F32SPLIT || g == 7, kErrUnsupported, "the FP32-split path implements gamma = 7 only");
    GCL_CHECK(params[SPH_H] > 0 && params[SPH_HH] > 0, kErrArgument, "h and H must be positive");
    // pair kernels address per-particle arrays with 32-bit byte offsets
    GCL_CHECK(system->n < (int64_t(1) << 29) - kPosPad, kErrUnsupported,
              "an SPH problem holds fewer than 2^29 - 64 particles");
    require_device();
    std::unique_ptr<sph_s> P(new sph_s());
    P->sys = system;
    P->dim = system->dim;
    P->prec = prec;
    P->n = system->n;
    std::memcpy(P->par, params, sizeof(P->par));
    P->G = system->G;
    const int64_t n = P->n;
    GCL_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
    P->sb.alloc(n, P->G.ncells());
    {
      void* c = nullptr;
      GCL_HIP(hipMalloc(&c, sizeof(StepCtl)));
      P->allocs.push_back(c);
      P->ctl = static_cast<StepCtl*>(c);
    }
    // cell structure of the system's last update!
    GCL_HIP(hipMemcpyAsync(P->sb.start, system->sb.start, 4 * size_t(P->G.ncells() + 1), hipMemcpyDeviceToDevice,
                           P->stream));
    GCL_HIP(hipMemcpyAsync(P->sb.cnt, system->sb.cnt, 4 * size_t(P->G.ncells() + 1), hipMemcpyDeviceToDevice,
                           P->stream));
    GCL_HIP(hipMemcpyAsync(P->sb.skey, system->sb.skey, 4 * size_t(n), hipMemcpyDeviceToDevice, P->stream));
    GCL_HIP(hipMemcpyAsync(P->sb.perm, system->sb.perm, 4 * size_t(n), hipMemcpyDeviceToDevice, P->stream));
    GCL_HIP(hipStreamSynchronize(system->stream));
    auto init = [&](auto& B, auto tag) {
      using R = decltype(tag);
      P->alloc_bufs(B);
      const int32_t* perm = system->sb.perm;
      for (int k = 0; k < P->dim; ++k) {
        gather<double>(static_cast<const double*>(system->xo[k]), B.X[0][k], perm, n, P->stream);
        GCL_HIP(hipMemcpyAsync(B.XRB[k], system->xs[k], 8 * size_t(n), hipMemcpyDeviceToDevice, P->stream));
      }
      GCL_HIP(hipMemcpyAsync(B.ID[0], perm, 4 * size_t(n), hipMemcpyDeviceToDevice, P->stream));
      GCL_HIP(hipMemcpyAsync(B.PT[1], ptype, 4 * size_t(n), hipMemcpyHostToDevice, P->stream));
      gather<int32_t>(B.PT[1], B.PT[0], perm, n, P->stream);
      std::vector<R> tmp(static_cast<size_t>(n));
      const double rho0 = P->par[SPH_RHO0];
      for (int64_t i = 0; i < n; ++i) tmp[size_t(i)] = std::is_same<R, float>::value ? R(rho[i] - rho0) : R(rho[i]);
      GCL_HIP(hipMemcpyAsync(B.RHO[1], tmp.data(), sizeof(R) * size_t(n), hipMemcpyHostToDevice, P->stream));
      gather<R>(B.RHO[1], B.RHO[0], perm, n, P->stream);
      GCL_HIP(hipStreamSynchronize(P->stream));
      for (int k = 0; k < P->dim; ++k) {
        for (int64_t i = 0; i < n; ++i) tmp[size_t(i)] = v_or_null ? R(v_or_null[k][i]) : R(0);
        GCL_HIP(hipMemcpyAsync(B.V[1][k], tmp.data(), sizeof(R) * size_t(n), hipMemcpyHostToDevice, P->stream));
        gather<R>(B.V[1][k], B.V[0][k], perm, n, P->stream);
        GCL_HIP(hipStreamSynchronize(P->stream));
      }
    };
    if (prec == GCLSPH_F32SPLIT) init(P->Bf, float());
    else init(P->Bd, double());
    std::memset(&P->hctl, 0, sizeof(StepCtl));
    P->hctl.synced = 1;
    P->hctl.mppcell = system->mppcell;
    P->hctl.dt = P->par[SPH_DT];
    P->hctl.etime = P->par[SPH_ETIME];
    P->write_ctl();
    {
      const int64_t nc = P->G.ncells();
      const SphConst C0 = P->make_const(0.0, 0.0);
      const StepFlags F0 = P->call_flags(false);
      if (prec == GCLSPH_F32SPLIT) {
        P->own.build(&n, &nc, 1, sizeof(SceneDev<float>));
        P->upload_own<float>(C0, F0);
        if (P->dim == 2) launch_init_pn<2, float>(P->own, P->stream);
        else launch_init_pn<3, float>(P->own, P->stream);
      } else {
        P->own.build(&n, &nc, 1, sizeof(SceneDev<double>));
        P->upload_own<double>(C0, F0);
        if (P->dim == 2) launch_init_pn<2, double>(P->own, P->stream);
        else launch_init_pn<3, double>(P->own, P->stream);
      }
    }
    GCL_HIP(hipGetLastError());
    GCL_HIP(hipStreamSynchronize(P->stream));
    P->xrb_mode = 0;
    system->attached = P.get();
    system->pairs_valid = true;  // the system's own pair list is still current
    *out = P.release();
  });
}

int sph_destroy(sph_t h) {
  return guard([&] {
    GCL_CHECK(h != nullptr, kErrArgument, "NULL handle");
    GCL_CHECK(h->batches == 0, kErrState, "problem belongs to a batch (destroy the batch first)");
    // hand the current points back to the cell list (caller order)
    gcl_s* g = h->sys;
    h->read_ctl();
    for (int k = 0; k < g->dim; ++k) {
      std::vector<double> v = unsort<double>(h->xcur(k), h->idcur(), g->n, h->stream);
      GCL_HIP(hipMemcpy(g->xo[k], v.data(), 8 * size_t(g->n), hipMemcpyHostToDevice));
    }
    g->attached = nullptr;
    delete h;
    g->update(true);
  });
}

int sph_set_param(sph_t h, int which, double value) {
  return guard([&] {
    GCL_CHECK(h != nullptr, kErrArgument, "NULL handle");
    GCL_CHECK(which >= 0 && which < SPH_NPARAMS, kErrArgument, "unknown parameter");
    GCL_CHECK(which != SPH_GAMMA || (value >= 1 && std::floor(value) == value), kErrArgument,
              "gamma must be an integer");
    GCL_CHECK(which != SPH_GAMMA || !h->split() || value == 7, kErrUnsupported,
              "the FP32-split path implements gamma = 7 only");
    h->flush_diag();
    h->par[which] = value;
  });
}

int sph_get_param(sph_t h, int which, double* value) {
  return guard([&] {
    GCL_CHECK(h != nullptr && value != nullptr, kErrArgument, "NULL argument");
    GCL_CHECK(which >= 0 && which < SPH_NPARAMS, kErrArgument, "unknown parameter");
    *value = h->par[which];
  });
}

int sph_stepsolve(sph_t h, int64_t n, int timestepping, double tl0, double tl1, int64_t* updaten, double* maxcdx) {
  return guard([&] {
    GCL_CHECK(h != nullptr && updaten != nullptr && maxcdx != nullptr, kErrArgument, "NULL argument");
    GCL_CHECK(n >= 0, kErrArgument, "number of steps must be >= 0");
    h->stepsolve(n, timestepping != 0, tl0, tl1, updaten, maxcdx);
  });
}

int sph_get_field(sph_t h, int field, double* out) {
  return guard([&] {
    GCL_CHECK(h != nullptr && out != nullptr, kErrArgument, "NULL argument");
    h->read_ctl();
    const int64_t n = h->n;
    const int32_t* id = h->idcur();
    const int p = h->hctl.parity;
    const double rho0 = h->par[SPH_RHO0];
    auto put = [&](const std::vector<double>& v, int k) { std::memcpy(out + size_t(k) * size_t(n), v.data(), 8 * size_t(n)); };
    auto vec_r = [&](auto* const* a, double add) {
      for (int k = 0; k < h->dim; ++k) put(unsort(a[k], id, n, h->stream, add), k);
    };
    bool sp = h->split();
    switch (field) {
      case SPH_FIELD_X:
        for (int k = 0; k < h->dim; ++k) put(unsort<double>(h->xcur(k), id, n, h->stream), 0 + k);
        break;
      case SPH_FIELD_V:
        if (sp) vec_r(h->Bf.V[p], 0.0);
        else vec_r(h->Bd.V[p], 0.0);
        break;
      case SPH_FIELD_RHO:
        if (sp) put(unsort<float>(h->Bf.RHO[p], id, n, h->stream, rho0), 0);
        else put(unsort<double>(h->Bd.RHO[p], id, n, h->stream), 0);
        break;
      case SPH_FIELD_P:  // the last pressure! call: pass 2's half-step EOS, or EOS(ρ) before DPC
        if (sp) put(unsort<float>(h->p_from_dpc ? h->Bf.PBUF : h->Bf.PH, id, n, h->stream), 0);
        else put(unsort<double>(h->p_from_dpc ? h->Bd.PBUF : h->Bd.PH, id, n, h->stream), 0);
        break;
      case SPH_FIELD_ACC:
        if (sp) vec_r(h->Bf.ACC, 0.0);
        else vec_r(h->Bd.ACC, 0.0);
        break;
      case SPH_FIELD_SUMW:
        h->flush_diag();
        put(unsort<double>(sp ? h->Bf.SUMW : h->Bd.SUMW, id, n, h->stream), 0);
        break;
      case SPH_FIELD_SUMGRADW:
        h->flush_diag();
        vec_r(sp ? h->Bf.SUMGW : h->Bd.SUMGW, 0.0);
        break;
      case SPH_FIELD_DRHO:
        if (sp) put(unsort<float>(h->Bf.DRHO, id, n, h->stream), 0);
        else put(unsort<double>(h->Bd.DRHO, id, n, h->stream), 0);
        break;
      case SPH_FIELD_PTYPE:
        put(unsort<int32_t>(h->ptcur(), id, n, h->stream), 0);
        break;
      case SPH_FIELD_DPC:  // Δv_dpc of the last step's dpcreg!; zero when the last call ran without DPC
        if (!h->p_from_dpc) {
          std::memset(out, 0, sizeof(double) * size_t(n) * size_t(h->dim));
        } else if (sp) {
          vec_r(h->Bf.DV, 0.0);
        } else {
          vec_r(h->Bd.DV, 0.0);
        }
        break;
      default:
        throw Error(kErrArgument, "unknown field");
    }
  });
}

int sph_set_timing(sph_t h, int on) {
  return guard([&] {
    GCL_CHECK(h != nullptr, kErrArgument, "NULL handle");
    h->timing = on != 0;
    for (int k = 0; k < 5; ++k) {
      h->stage_ms[k] = 0;
      h->stage_cnt[k] = 0;
    }
  });
}

int sph_get_timing(sph_t h, double* ms, int64_t* launches, int cap, int* nstages) {
  return guard([&] {
    GCL_CHECK(h != nullptr, kErrArgument, "NULL handle");
    for (int k = 0; k < 5 && k < cap; ++k) {
      if (ms) ms[k] = h->stage_ms[k];
      if (launches) launches[k] = h->stage_cnt[k];
    }
    if (nstages) *nstages = 5;
  });
}

int sph_get_pass_variants(sph_t h, int32_t* out4) {
  return guard([&] {
    GCL_CHECK(h != nullptr && out4 != nullptr, kErrArgument, "NULL argument");
    for (int k = 0; k < 2; ++k) {
      const int v = h->own.pass_variant[k];
      out4[k] = v;
      out4[2 + k] = v >= 0 ? pass_block_size(v) : -1;
    }
  });
}

int sph_batch_create(const sph_t* problems, int32_t count, sph_batch_t* out) {
  return guard([&] {
    GCL_CHECK(problems != nullptr && out != nullptr, kErrArgument, "NULL argument");
    GCL_CHECK(count > 0, kErrArgument, "a batch needs at least one problem");
    *out = nullptr;
    for (int32_t i = 0; i < count; ++i) {
      GCL_CHECK(problems[i] != nullptr, kErrArgument, "NULL problem in batch");
      GCL_CHECK(problems[i]->dim == problems[0]->dim && problems[i]->prec == problems[0]->prec, kErrArgument,
                "batched problems must share dimension and precision");
      for (int32_t j = 0; j < i; ++j)
        GCL_CHECK(problems[j] != problems[i], kErrArgument, "a problem appears twice in the batch");
    }
    require_device();
    std::unique_ptr<sph_batch_s> b(new sph_batch_s());
    b->probs.assign(problems, problems + count);
    for (sph_s* p : b->probs) p->batches += 1;
    GCL_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    std::vector<int64_t> n(static_cast<size_t>(count)), nc(static_cast<size_t>(count));
    for (int32_t i = 0; i < count; ++i) {
      n[size_t(i)] = b->probs[size_t(i)]->n;
      nc[size_t(i)] = b->probs[size_t(i)]->G.ncells();
    }
    const size_t bytes = problems[0]->split() ? sizeof(SceneDev<float>) : sizeof(SceneDev<double>);
    b->L.build(n.data(), nc.data(), unsigned(count), bytes);
    *out = b.release();
  });
}

int sph_batch_destroy(sph_batch_t b) {
  return guard([&] {
    GCL_CHECK(b != nullptr, kErrArgument, "NULL handle");
    delete b;
  });
}

int sph_batch_stepsolve(sph_batch_t b, int64_t n, int timestepping, const double* tl0, const double* tl1,
                        int64_t* updaten, double* maxcdx) {
  return guard([&] {
    GCL_CHECK(b != nullptr && updaten != nullptr && maxcdx != nullptr, kErrArgument, "NULL argument");
    GCL_CHECK(n >= 0, kErrArgument, "number of steps must be >= 0");
    const size_t cnt = b->probs.size();
    std::vector<double> t0(cnt), t1(cnt);
    for (size_t i = 0; i < cnt; ++i) {
      const double* par = b->probs[i]->par;
      // default timelims (sqrt(eps), CFL·H/(3c0)) (src/sphproblem.jl:207,426)
      t0[i] = tl0 ? tl0[i] : std::sqrt(std::numeric_limits<double>::epsilon());
      t1[i] = tl1 ? tl1[i] : par[SPH_CFL] * par[SPH_HH] / (3 * par[SPH_C0]);
    }
    sph_s* const* P = b->probs.data();
    const bool ts = timestepping != 0;
    if (P[0]->split()) {
      if (P[0]->dim == 2) run_steps<2, float>(P, unsigned(cnt), b->L, b->stream, n, ts, t0.data(), t1.data(), updaten, maxcdx);
      else run_steps<3, float>(P, unsigned(cnt), b->L, b->stream, n, ts, t0.data(), t1.data(), updaten, maxcdx);
    } else {
      if (P[0]->dim == 2) run_steps<2, double>(P, unsigned(cnt), b->L, b->stream, n, ts, t0.data(), t1.data(), updaten, maxcdx);
      else run_steps<3, double>(P, unsigned(cnt), b->L, b->stream, n, ts, t0.data(), t1.data(), updaten, maxcdx);
    }
  });
}

}  // extern "C"

int sph_batch_get_pass_variants(sph_batch_t b, int32_t* out4) {
  return guard([&] {
    GCL_CHECK(b != nullptr && out4 != nullptr, kErrArgument, "NULL argument");
    for (int k = 0; k < 2; ++k) {
      const int v = b->L.pass_variant[k];
      out4[k] = v;
      out4[2 + k] = v >= 0 ? pass_block_size(v) : -1;
    }
  });
}
