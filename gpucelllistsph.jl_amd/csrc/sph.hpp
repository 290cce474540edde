// SPH step: device state, constants and the host-side driver.
#pragma once

#include <type_traits>
#include <vector>

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
  double dist, dist2;   // neighbour-list cutoff and dist^2 (src/gpukernels_cls.jl:294)
  double fb_d[2], fb_r0[2], fb_eta2[2];  // fbmolforce (0.4, 2dx) / (bound_D, bound_l); η² = 0.01 r0^2
  double st_2h, st_c, st_01h;            // 2h, 1.5π, 0.1h       :1393-1394
  double dpc_l0, dpc_l0inv, dpc_eta2, dpc_whinv, dpc_lim, dpc_wnorm, dpc_lambda, dpc_pmin, dpc_pmax;
  double xsph_c;        // 2m0 ε                   :1766
  double dt_eta2;       // (0.01h)^2               :1298
  double CFL, nui, tl0, tl1;
  // single-precision copies for the FP32-split kernels
  float f_h, f_Hinv, f_rho0, f_m0, f_alpha, f_c0, f_g, f_s, f_nu, f_eta2, f_DDTkh, f_DDTgz, f_P0, f_wnorm,
      f_dwnorm, f_fb_d[2], f_fb_r0[2], f_fb_eta2[2], f_st_2h, f_st_c, f_st_01h, f_xsph_c;
  float f_dpc_l0inv, f_dpc_eta2, f_dpc_whinv, f_dpc_lim, f_dpc_wnorm, f_dpc_lambda, f_dpc_pmin, f_dpc_pmax;
  float f_inv_rho0;  // 1/ρ0
  float f_st_turn;   // 1.5π/(2h) in turns: cos(1.5π r/2h) = cos(2π · r·f_st_turn)
  float f_dwr;       // dW/dr / r = f_dwr (1-u)^3: -20 dwnorm H⁻¹
  double fcs[3];     // cell-frame spacing of the 3D FP32-split partner records: the
  float f_fcs[3];    // cell sizes rounded to FP32 (exact in both copies)
};

// Switches of one problem.  The host sets the per-call ones; the per-step ones
// (synced_prev, want_out, corr_on, cspm_fire) are derived on the device from the
// control block at every launch (step_flags), so a step's launch sequence is the
// same for every step and every problem of a batch.
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
  int dpc_on, xsph_on;
  int cspmn;        // CSPM every cspmn iterations of a call (0: never)
  int cspm_fire;    // CSPM runs after this step
};

// Candidates per 3D FP32-split membership window (one stencil cell; the
// cell's offset code, 0..26, takes the five bits above the member bits).  The
// other layouts (2D, 3D FP64) use 32-candidate windows of a stencil row.
constexpr int kWin3 = 27;
template <int D, typename R>
constexpr bool cell_windows() {
  return D == 3 && std::is_same<R, float>::value;
}

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
  // partner records of the pair kernels, one per sorted particle: PartnerN /
  // PartnerH(X) below in FP32-split, PartnerN64 / PartnerH64 in FP64.  PRH
  // holds the half-step records from pass 1 to pass 2, then the XSPH records
  // (N layout, xⁿ in place of x).
  void* PRN;
  void* PRH;
};

// What a pair pass reads of a partner, packed into one record of 16-byte words
// so that a partner costs a few 16-byte loads instead of one load per field
// (in 2D the arrays keep word i of every particle's record together:
// RecordWords below).
// PartnerN: the state at tⁿ read by pass 1 — written by the reorder, by pass 2
// for the next step and by the corrections that change x, v or ρ.
// PartnerH: the half-step state read by pass 2, written by pass 1.  Pass 2
// also needs the partner's tⁿ position (∇W at tⁿ): when the half step started
// from xⁿ (no correction in the previous step) x½ = xⁿ + v½·Δt/2, so the tⁿ
// separation is the x½ separation minus (v½_a - v½_b)·Δt/2, exact to FP64
// rounding; otherwise the records are PartnerHX, which append x½ - xⁿ in FP32
// (a displacement of order v·Δt, so its FP32 rounding lies far below the FP32
// rounding of the separation itself; the step flag synced_prev tells both
// passes which layout the step uses).  In 3D the
// pressure is not stored: it is the EOS of (rho, pt), recomputed with the same
// function that fills PN / PH; the 2D records store it (the EOS is ~10 VALU
// of a pair's ~110, 2-3 % of a pass).
//
// In 3D the position is held as FP32 offsets from the origin of the
// particle's cell in the current sort (its cell frame, cell_frame below): a
// partner in the stencil cell at offset o from a's cell is at separation
// (x_a - frame(c_a + o)) - xo_b, where the first term is a per-window constant
// of a (the stored windows name their cell, so it is computed once per
// window), and the record shrinks to 32 bytes: two 16-byte loads per partner
// instead of three.  The offsets of a particle inside its cell are below one
// cell size, so their FP32 rounding (2^-24 of a cell) is of the same order as
// the rounding of the FP32 separation itself.  A particle outside the grid
// (binned into an edge cell) sits further from its frame and loses
// precision in proportion.  2D keeps FP64 positions: its record is two loads
// either way.
// The 2D records hold the partner's pressure (the EOS of (rho, pt)) in the
// place of its type; the type's class (< 0, 0, 1, > 1: all that a pair term
// asks of a partner's type) rides in the two lowest mantissa bits of the FP64
// y (set_pt / rec_pt).
template <int D>
struct alignas(16) PartnerN {
  double x[D];
  float v[D];
  float rho;  // ρ - ρ0
  float P;
};
template <>
struct alignas(16) PartnerN<3> {
  float xo[3];  // x - cell frame
  float v[3];
  float rho;
  int32_t pt;
};
template <int D>
struct alignas(16) PartnerH {
  double xh[D];
  float vh[D];
  float rh;  // ρ½ - ρ0
  float P;   // P½
};
template <>
struct alignas(16) PartnerH<3> {
  float xo[3];  // x½ - cell frame
  float vh[3];
  float rh;
  int32_t pt;
};
static_assert(sizeof(PartnerN<2>) == 32 && sizeof(PartnerN<3>) == 32, "partner record sizes");
template <int D>
struct alignas(16) PartnerHX {
  PartnerH<D> h;
  float dn[D];  // x½ - xⁿ
};
static_assert(sizeof(PartnerH<2>) == 32 && sizeof(PartnerH<3>) == 32, "partner record sizes");
static_assert(sizeof(PartnerHX<2>) == 48 && sizeof(PartnerHX<3>) == 48, "partner record sizes");

// FP64 path: the pressure is stored (its EOS costs an FP64 division and a
// power per partner), and pass 2's record always carries xⁿ (an FP64-exact
// tⁿ separation cannot be derived from the half-step state).
template <int D>
struct alignas(16) PartnerN64 {
  double x[D];
  double v[D];
  double rho;
  double P;
  int32_t pt;
};
template <int D>
struct alignas(16) PartnerH64 {
  double xn[D];
  double xh[D];
  double vh[D];
  double rh;
  double P;
  int32_t pt;
};
static_assert(sizeof(PartnerN64<2>) == 64 && sizeof(PartnerN64<3>) == 80, "partner record sizes");
static_assert(sizeof(PartnerH64<2>) == 80 && sizeof(PartnerH64<3>) == 96, "partner record sizes");

// How an array keeps its records (ld_rec / st_rec): in 2D as 16-byte words,
// word i of every particle together (the lanes of a wave, whose partners lie
// near each other in the x-ordered rows, then touch half the cache lines per
// load), in 3D whole records in a row (a particle's ~160 partners in 27
// cells are spread out, and a record split over two lines costs two).
template <typename T>
struct RecordWords : std::false_type {};
template <>
struct RecordWords<PartnerN<2>> : std::true_type {};
template <>
struct RecordWords<PartnerH<2>> : std::true_type {};
template <>
struct RecordWords<PartnerHX<2>> : std::true_type {};
template <>
struct RecordWords<PartnerN64<2>> : std::true_type {};
template <>
struct RecordWords<PartnerH64<2>> : std::true_type {};

// Record types of a precision: N (state at tⁿ, pass 1 and the corrections) and
// H (pass 2; FP32-split uses PartnerH or PartnerHX by step, see above).
template <int D, typename R>
struct Records;
template <int D>
struct Records<D, float> {
  using N = PartnerN<D>;
  using H = PartnerHX<D>;
};
template <int D>
struct Records<D, double> {
  using N = PartnerN64<D>;
  using H = PartnerH64<D>;
};
// The XSPH pass writes N-layout records into the half-step record buffer.
static_assert(sizeof(Records<2, float>::H) >= sizeof(Records<2, float>::N) &&
                  sizeof(Records<3, float>::H) >= sizeof(Records<3, float>::N) &&
                  sizeof(Records<2, double>::H) >= sizeof(Records<2, double>::N) &&
                  sizeof(Records<3, double>::H) >= sizeof(Records<3, double>::N),
              "half-step records hold an N record");

// Device descriptor of one problem: everything a kernel needs, read by each
// block from the scene it works on.
template <typename R>
struct SceneDev {
  SphBufs<R> B;
  SphConst C;
  CellGeom G;
  StepFlags F;
  // counting-sort buffers of the rebuild
  int32_t *key, *slot, *cnt, *start, *tmp, *perm, *skey, *maxcell;
  int64_t* scan_tmp;
  int64_t ncells;
};

// Launch geometry of a set of problems stepped together (a batch; a single
// problem is a batch of one).  Particle kernels run one block per 256 particles
// of one problem; pmap maps a logical block to (problem, block of the problem);
// pmap512 does the same for the 512-thread blocks of the pass kernels (their
// 128-thread blocks halve the pmap entries).
// Cell kernels run one block per 1024-cell tile; cmap maps a block to (problem,
// tile).  Per-problem kernels run one block per problem.
struct LaunchSet {
  void* scenes = nullptr;  // device SceneDev<R>[count]
  const void* one = nullptr;  // host SceneDev<R> passed by value when set (count == 1)
  int2* pmap = nullptr;
  int2* cmap = nullptr;
  int2* pmap512 = nullptr;
  unsigned pblocks = 0, ctiles = 0, count = 0, pblocks512 = 0;
  size_t scene_bytes = 0;
  // pass variants chosen by timing (-1: not yet)
  int pass_variant[2] = {-1, -1};
  // the captured step of the last call and what it was captured for
  hipGraphExec_t step_graph = nullptr;
  std::vector<unsigned char> graph_key;
  // n[i] particles and ncells[i] cells of problem i
  void build(const int64_t* n, const int64_t* ncells, unsigned count, size_t scene_bytes);
  void release();
};

// Rebuild stage of one step for every problem whose control block asks for it:
// binning, scan, scatter, rank, reorder of the state, commit, membership store.
template <int D, typename R>
void launch_rebuild(const LaunchSet& L, hipStream_t s);
// Pair passes in one of kPassVariants variants that differ only in code
// generation and launch shape (same results): variant = cg + 4·bs + 12·oc,
// with cg in [0, 4) the prefetch / occupancy-target choice, bs in [0, 3) the
// block size (256, 128, 512 threads) and oc in [0, 2) an occupancy cap (1: at
// most kPassCapWaves waves per SIMD, by launching with unused LDS).
constexpr int kPassCodegen = 4;
constexpr int kPassBlockSizes = 3;
constexpr int kPassCaps = 2;
constexpr int kPassCapWaves = 4;
constexpr int kPassVariants = kPassCodegen * kPassBlockSizes * kPassCaps;
// Steps the autotuner times: every code-generation variant at every block
// size (the best shape at one block size need not be the best at another),
// then the best of them with the cap and the two best once more; before them
// up to kTuneWarm untimed steps (as many as the call has to spare).
constexpr int kTuneTimed = kPassCodegen * kPassBlockSizes + kPassCaps - 1 + 2;
constexpr int kTuneWarm = 16;
constexpr int pass_block_size(int variant) {
  return (variant / kPassCodegen) % kPassBlockSizes == 1 ? 128
         : (variant / kPassCodegen) % kPassBlockSizes == 2 ? 512
                                                           : 256;
}
constexpr int pass_cap(int variant) { return variant / (kPassCodegen * kPassBlockSizes); }
template <int D, typename R>
void launch_pass1(const LaunchSet& L, hipStream_t s, int variant);
template <int D, typename R>
void launch_pass2(const LaunchSet& L, hipStream_t s, int variant);
template <int D, typename R>
void launch_dpc(const LaunchSet& L, hipStream_t s);
template <int D, typename R>
void launch_xsph(const LaunchSet& L, hipStream_t s);
template <int D, typename R>
void launch_cspm(const LaunchSet& L, hipStream_t s);
template <int D, typename R>
void launch_dt_visc(const LaunchSet& L, hipStream_t s);
template <int D, typename R>
void launch_finalize(const LaunchSet& L, hipStream_t s);
template <int D, typename R>
void launch_diag(const LaunchSet& L, int xm_mode, int use_masks, hipStream_t s);
template <int D, typename R>
void launch_init_pn(const LaunchSet& L, hipStream_t s);

}  // namespace gclsph
