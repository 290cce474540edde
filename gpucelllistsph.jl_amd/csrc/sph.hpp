// SPH step: device state, constants and the host-side driver.
#pragma once

#include "celllist.hpp"

namespace gclsph {

// Per-problem constants, derived on the host in Float64 exactly as the
// reference derives them before each launch (citations at the use sites).
struct SphConst {
  int dim;
  int gamma;
  double dx, h, H, Hinv, rho0, m0, alpha, c0, g, s, nu;
  double eta2;          // (0.1h)^2               src/gpukernels_sph.jl:556,854,953
  double ginv;          // 1/γ                     :557
  double DDTkh;         // 2 h δ c0                :558
  double Cb;            // (c0 c0 ρ0) γ⁻¹          :559
  double DDTgz;         // ρ0 g / Cb               :560
  double P0;            // c0^2 ρ0 / γ             :712
  double wnorm, dwnorm; // σ_D H⁻ᴰ, σ_D H⁻⁽ᴰ⁺¹⁾ (WendlandC2)
  double dist2;         // neighbour-list cutoff dist^2 (src/gpukernels_cls.jl:294)
  double fb_d[2], fb_r0[2], fb_eta2[2];  // fbmolforce (0.4, 2dx) / (bound_D, bound_l); η² = 0.01 r0^2
  double st_2h, st_c, st_01h;            // 2h, 1.5π, 0.1h       :1393-1394
  double dpc_l0, dpc_l0inv, dpc_eta2, dpc_whinv, dpc_lim, dpc_wnorm, dpc_lambda, dpc_pmin, dpc_pmax;
  double xsph_c;        // 2m0 ε                   :1766
  double dt_eta2;       // (0.01h)^2               :1298
  double CFL, nui, tl0, tl1;
  // single-precision copies for the FP32-split kernels
  float f_h, f_Hinv, f_rho0, f_m0, f_alpha, f_c0, f_g, f_s, f_nu, f_eta2, f_DDTkh, f_DDTgz, f_P0, f_wnorm,
      f_dwnorm, f_fb_d[2], f_fb_r0[2], f_fb_eta2[2], f_st_2h, f_st_c, f_st_01h, f_xsph_c;
};

// Per-launch switches decided on the host.
struct StepFlags {
  int need_xrb;     // membership uses the rebuild positions when a step did not rebuild
  int synced_prev;  // half-step state of the previous step equals the main state
  int want_out;     // write acceleration / dρ/dt / pressure of this step
  int need_cdx;     // keep cumulative displacement across steps (nui > 0)
  int corr_on;      // a correction (DPC, XSPH, CSPM) runs after this step's update
  int ts;           // Δt stepping: reduce |a| and the viscous rate
  int visc_in_pass2;
  int nu_on, s_on;
  int write_xrb;    // the rebuild stores its positions for later membership tests
};

// Position arrays (both parities and the rebuild copy) carry this many zero
// elements past n: the membership test reads candidates in groups of four and
// may run past the end of the last stencil row.
constexpr int64_t kPosPad = 64;

// Device buffers of one problem.  Main state is double-buffered ([parity]);
// ptype / id are double-buffered by the reorder only ([pparity]).
template <typename R>
struct SphBufs {
  int64_t n;
  double* X[2][3];
  R* V[2][3];
  R* RHO[2];  // ρ (F64) or ρ-ρ0 (F32-split)
  int32_t* PT[2];
  int32_t* ID[2];
  double* XH[3];
  R* VH[3];
  R* RH;
  double* XH2[3];
  R* VH2[3];
  R* RH2;
  double* XRB[3];
  double* CDX[3];
  R* ACC[3];
  R* DRHO;
  R* PBUF;
  R* PN;  // EOS of the current density (written with every density update)
  R* PH;  // EOS of the half-step density (written by pass 1)
  R* DV[3];
  double* TMP;
  double* SUMW;
  double* SUMGW[3];
  const int32_t* start;
  const int32_t* skey;
  const int32_t* perm;
  const int32_t* maxcell;
  StepCtl* ctl;
  double* RED;  // per-block partial maxima: [|cdx|, |a|, visc] x nblk
  int nblk;
  // neighbour-list membership of the last rebuild (k_nmask): entry e of
  // particle a at NE[e * n + a] = (window base << 32) | member bits; NC[a]
  // entries, or -1 when more than necap were needed
  uint64_t* NE;
  int32_t* NC;
  int necap;
};

template <int D, typename R>
void launch_reorder(const SphBufs<R>& B, const SphConst& C, const StepFlags& F, hipStream_t s);
template <typename R>
void launch_init_pn(const SphBufs<R>& B, const SphConst& C, hipStream_t s);
template <int D, typename R>
void launch_pass1(const SphBufs<R>& B, const SphConst& C, const CellGeom& G, const StepFlags& F, hipStream_t s);
template <int D, typename R>
void launch_pass2(const SphBufs<R>& B, const SphConst& C, const CellGeom& G, const StepFlags& F, hipStream_t s);
template <int D, typename R>
void launch_dpc(const SphBufs<R>& B, const SphConst& C, const CellGeom& G, const StepFlags& F, hipStream_t s);
template <int D, typename R>
void launch_xsph(const SphBufs<R>& B, const SphConst& C, const CellGeom& G, const StepFlags& F, hipStream_t s);
template <int D, typename R>
void launch_cspm(const SphBufs<R>& B, const SphConst& C, const CellGeom& G, const StepFlags& F, hipStream_t s);
template <int D, typename R>
void launch_dt_visc(const SphBufs<R>& B, const SphConst& C, hipStream_t s);
template <int D, typename R>
void launch_diag(const SphBufs<R>& B, const SphConst& C, const CellGeom& G, int xm_mode, int use_masks,
                 hipStream_t s);
template <int D, typename R>
void launch_nmask(const SphBufs<R>& B, const CellGeom& G, double dist2, hipStream_t s);
void launch_post_rebuild(StepCtl* ctl, const int32_t* maxcell, int write_xrb, hipStream_t s);
void launch_finalize(StepCtl* ctl, const double* red, int nblk, const SphConst& C, int ts, int corr_on, hipStream_t s);

}  // namespace gclsph
