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
#include <atomic>

#include "sph_device.hpp"

namespace gclsph {

// Pass kernels come in variants that differ only in code generation and
// launch shape (so their results are bit-identical: one thread computes one
// particle whatever the block): Pipe fetches the next partner's data before the
// current pair body (two register sets), W asks the compiler for at least W
// waves per SIMD (fewer registers), BS is the block size.  The step driver
// picks one per pass by timing them on live steps.
//
// ---- pass 1 ---------------------------------------------------------------
template <int D, typename R, typename SS, bool Pipe, int W, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(W))) void k_pass1(SS ss) {
#pragma clang fp contract(off)
  GCLSPH_SCENE_BS(ss, BS);
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
    using Nb = PartnerN64<D>;
    auto fetch = [&](const int32_t b) { return ld_rec<Nb>(B.PRN, b, B.n); };
    for_each_partner_fetch<Pipe, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& q) {
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
      if (same_kind(pta, ptb)) {  // ∂v∂t! (:755-772)
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
    const double Ph = eos_d(rv, pta, C);
    B.RH[a] = rv;
    B.PH[a] = Ph;
    PartnerH64<D> rec{};
#pragma unroll
    for (int d = 0; d < D; ++d) {
      double vv = vh0[d];
      if (pta > 1) vv += av[d] * hdt;
      const double xh = xh0[d] + vv * hdt;
      B.VH[d][a] = vv;
      B.XH[d][a] = xh;
      rec.xn[d] = xa[d];
      rec.xh[d] = xh;
      rec.vh[d] = vv;
    }
    rec.rh = rv;
    rec.P = Ph;
    rec.pt = pta;
    st_rec(B.PRH, a, rec, B.n);
  } else {
    // ---------------- FP32-split
#pragma clang fp contract(fast)
    const float dra = sa;
    const float rhoa = C.f_rho0 + dra;
    const float inv_rhoa = __builtin_amdgcn_rcpf(rhoa);
    const float mPa = -C.f_m0 * inv_rhoa;  // -m0/ρ_a of the pressure term
    const float Pa = B.PN[a];
    using Nb = PartnerN<D>;
    auto fetch = [&](const int32_t b) { return ld_rec<Nb>(B.PRN, b, B.n); };
    const FrameK3 kt = particle_k<D, R>(B, C, G, a, xa);
    // ∇W = dx·dwr, so every term along ∇W is gathered into one coefficient fg
    // of dx, and dx·∇W = r²·dwr, dv·∇W = (dv·dx)·dwr.
    for_each_partner_fetch<Pipe, D, R, true>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& q, const auto&... win) {
      const int32_t ptb = rec_pt(q);
      float dx[3], dv[3];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        dx[d] = sep_x(xa, q, d, win...);
        dv[d] = va[d] - q.v[d];
      }
      const float r2 = dotf<D>(dx, dx);
      const float r = __builtin_amdgcn_sqrtf(r2);
      const float dwr = dWr_f(r * C.f_Hinv, C.f_dwr);
      const float inv_den = __builtin_amdgcn_rcpf(r2 + C.f_eta2);
      const float dvdx = dotf<D>(dv, dx);
      const float drb = q.rho;
      const float rhob = C.f_rho0 + drb;
      const RhoInv ri = rho_inv2(rhoa, rhob);
      const float inv_rhob = ri.inv_rhob;
      // The common terms are evaluated for every pair and selected (no
      // divergent branches in the hot loop): continuity + δ-SPH, pressure, AV.
      const bool both = pta > 0 && ptb > 0;
      float ax[3] = {0.0f, 0.0f, 0.0f};
      {
        const float s = C.f_m0 * dwr * dvdx;
        const float drhop = hydro_drho_f(C.f_DDTgz * dx[1], C.f_rho0);
        const float vd = C.f_DDTkh * ((drb - dra) - drhop) * inv_den;
        const float sd = s - vd * (r2 * dwr) * C.f_m0 * inv_rhob;
        drho.add(both ? (pta > 1 ? sd : s) : 0.0f);
      }
      const float Pt = (Pa + rec_P(q, C)) * inv_rhob * mPa;
      float fg = same_kind(pta, ptb) ? Pt : 0.0f;
      {
        const float dmu = C.f_h * fminf(dvdx, 0.0f) * inv_den;  // AV acts on approaching pairs only
        const float dPi = (-C.f_alpha * C.f_c0 * dmu) * (2.0f * ri.inv_sum);
        fg -= both ? dPi * C.f_m0 : 0.0f;
      }
      if (both) {
        if (F.nu_on) {
          float term = 4.0f * C.f_nu * C.f_m0 * (r2 * dwr) * (ri.inv_sum * inv_den);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += term * dv[d];
        }
      }
      fg *= dwr;
      if (both && F.s_on && r < C.f_st_2h) fg += surf_f(r, C) * (1.0f / C.f_m0);
#pragma unroll
      for (int d = 0; d < D; ++d) ax[d] += fg * dx[d];
      if ((pta > 0) != (ptb > 0) && r < C.f_fb_r0[0]) {
        // wall force: FP32 while q = r0/(r + η²) < 32 (q^12 < 1.2e18, far inside
        // the FP32 range), FP64 for the (rare) closer pairs
        const float q1 = C.f_fb_r0[0] * __builtin_amdgcn_rcpf(r + C.f_fb_eta2[0]);
        if (q1 < 32.0f) {
          const float q2 = q1 * q1, q4 = q2 * q2, q8 = q4 * q4;
          const float Fc = C.f_fb_d[0] * (q4 * q8 - q4) * __builtin_amdgcn_rcpf(r2 + C.f_fb_eta2[0]);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += dx[d] * Fc;
        } else {
          double p1 = C.fb_r0[0] / (double(r) + C.fb_eta2[0]);
          double p2 = p1 * p1, p4 = p2 * p2, p8 = p4 * p4;
          double Fc = C.fb_d[0] * (p4 * p8 - p4) / (double(r2) + C.fb_eta2[0]);
#pragma unroll
          for (int d = 0; d < D; ++d) ax[d] += float(double(dx[d]) * Fc);
        }
      }
#pragma unroll
      for (int d = 0; d < D; ++d) acc[d].add(ax[d]);
    }, &kt);
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
    const float Ph = eos_f(rv, pta, C);
    B.PH[a] = Ph;
    PartnerHX<D> rec{};
    double fr[3];
    particle_frame<D, R>(B, C, G, a, fr);
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float vv = vh0[d];
      if (pta > 1) vv += av[d] * hdt;
      const double xh = xh0[d] + double(vv) * (dt * 0.5);
      B.VH[d][a] = vv;
      B.XH[d][a] = xh;
      put_x(rec.h, d, xh, fr);
      rec.h.vh[d] = vv;
      rec.dn[d] = float(xh - xa[d]);
    }
    rec.h.rh = rv;
    set_pt(rec.h, pta, Ph);
    if (F.synced_prev) st_rec(B.PRH, a, rec.h, B.n);  // pass 2 derives xⁿ
    else st_rec(B.PRH, a, rec, B.n);
  }
}

// ---- pass 2 + symplectic update ------------------------------------------
template <int D, typename R, typename SS, bool Pipe, int W, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(W))) void k_pass2(SS ss) {
#pragma clang fp contract(off)
  GCLSPH_SCENE_BS(ss, BS);
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
      using Nb = PartnerH64<D>;
      auto fetch = [&](const int32_t b) { return ld_rec<Nb>(B.PRH, b, B.n); };
      for_each_partner_fetch<Pipe, D>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& q) {
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
        if (same_kind(pta, ptb)) {
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
      const double Pn = eos_d(rv, pta, C);
      B.RHO[q][a] = rv;
      B.PN[a] = Pn;
      if (F.corr_on) B.RH2[a] = rv;
      const double ml = pta > 1 ? 1.0 : 0.0;
      double xn[3], vn[3];
      PartnerN64<D> rec{};
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
        rec.x[d] = xv;
        rec.v[d] = nv;
        if (F.corr_on) {
          B.VH2[d][a] = nv;
          B.XH2[d][a] = xv;
        }
      }
      rec.rho = rv;
      rec.P = Pn;
      rec.pt = pta;
      st_rec(B.PRN, a, rec, B.n);  // partner record of the next step's pass 1
      if (F.ts) {
        red_acc = sqrt(dotd<D>(av, av));
        if (F.visc_in_pass2) red_visc = fabs(C.h * dotd<D>(vn, xn) / (dotd<D>(xn, xn) + C.dt_eta2));
      }
    } else {
#pragma clang fp contract(fast)
      const float dra = sha;
      const float rhoa = C.f_rho0 + dra;
      const float inv_rhoa = __builtin_amdgcn_rcpf(rhoa);
      const float mPa = -C.f_m0 * inv_rhoa;  // -m0/ρ_a of the pressure term
      Pha = B.PH[a];
      // tⁿ separation from the half-step state when x½ = xⁿ + v½·Δt/2, else
      // from the displacement x½ - xⁿ that pass 1 appended to the records
      // xⁿa - xⁿb = (x½a - x½b) - (da - db) with d = x½ - xⁿ, the displacement
      // part in FP32 (its size is of order |v|·Δt): derived, d = v½·Δt/2
      // (PartnerH records); else d from the records (PartnerHX).  The choice
      // holds for the whole step, so the loop is compiled once for each.
      const float fhdt = float(dt * 0.5);
      float dna[3];
#pragma unroll
      for (int d = 0; d < D; ++d) dna[d] = float(xha[d] - xa[d]);
      const FrameK3 kt = particle_k<D, R>(B, C, G, a, xha);
      auto gather = [&](auto derive_c) {
      constexpr bool derive = decltype(derive_c)::value;
      using Nb = std::conditional_t<derive, PartnerH<D>, PartnerHX<D>>;
      auto fetch = [&](const int32_t b) { return ld_rec<Nb>(B.PRH, b, B.n); };
      for_each_partner_fetch<Pipe, D, R, true>(B, a, nc, G, XM, C.dist2, fetch, [&](const int32_t, const Nb& r,
                                                                            const auto&... win) {
        const PartnerH<D>* qp;
        if constexpr (derive) qp = &r;
        else qp = &r.h;
        const PartnerH<D>& q = *qp;
        const int32_t ptb = rec_pt(q);
        float dxn[3], dx[3], dv[3];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          dx[d] = sep_x(xha, q, d, win...);
          dv[d] = vha[d] - q.vh[d];
          if constexpr (derive) dxn[d] = dx[d] - dv[d] * fhdt;
          else dxn[d] = dx[d] - (dna[d] - r.dn[d]);
        }
        // ∇W from pass 1 (positions xⁿ): ∇W = dxn·dwr; terms along ∇W are
        // gathered into one coefficient fg of dxn
        const float r2n = dotf<D>(dxn, dxn);
        const float rn = __builtin_amdgcn_sqrtf(r2n);
        const float dwr = dWr_f(rn * C.f_Hinv, C.f_dwr);
        const float r2 = dotf<D>(dx, dx);
        const float inv_den = __builtin_amdgcn_rcpf(r2 + C.f_eta2);
        const float dvdx = dotf<D>(dv, dx);
        const float drb = q.rh;
        const float rhob = C.f_rho0 + drb;
        const RhoInv ri = rho_inv2(rhoa, rhob);
        const float inv_rhob = ri.inv_rhob;
        // common terms evaluated for every pair and selected, as in pass 1
        const bool both = pta > 0 && ptb > 0;
        float ax[3] = {0.0f, 0.0f, 0.0f};
        const float dxg = dotf<D>(dx, dxn) * dwr;  // dx·∇W
        {
          const float s = C.f_m0 * dwr * dotf<D>(dv, dxn);
          const float drhop = hydro_drho_f(C.f_DDTgz * dx[1], C.f_rho0);
          const float vd = C.f_DDTkh * ((drb - dra) - drhop) * inv_den;
          const float sd = s - vd * dxg * C.f_m0 * inv_rhob;
          drho.add(both ? (pta > 1 ? sd : s) : 0.0f);
        }
        if (both) {
          if (F.nu_on) {
            float term = 4.0f * C.f_nu * C.f_m0 * dxg * (ri.inv_sum * inv_den);
#pragma unroll
            for (int d = 0; d < D; ++d) ax[d] += term * dv[d];
          }
          if (F.s_on) {
            const float r = __builtin_amdgcn_sqrtf(r2);
            if (r < C.f_st_2h) {
              const float scos = surf_f(r, C);
#pragma unroll
              for (int d = 0; d < D; ++d) ax[d] += scos * dx[d] * (1.0f / C.f_m0);
            }
          }
        }
        const float Pt = (Pha + rec_P(q, C)) * inv_rhob * mPa;
        float fg = same_kind(pta, ptb) ? Pt : 0.0f;
        {
          const float dmu = C.f_h * fminf(dvdx, 0.0f) * inv_den;
          const float dPi = (-C.f_alpha * C.f_c0 * dmu) * (2.0f * ri.inv_sum);
          fg -= both ? dPi * C.f_m0 : 0.0f;
        }
        fg *= dwr;
#pragma unroll
        for (int d = 0; d < D; ++d) ax[d] += fg * dxn[d];
        if ((pta > 0) != (ptb > 0) && rn < C.f_fb_r0[1]) {
          // wall force: FP32 while q < 32, FP64 for closer pairs (as in pass 1)
          const float q1 = C.f_fb_r0[1] * __builtin_amdgcn_rcpf(rn + C.f_fb_eta2[1]);
          if (q1 < 32.0f) {
            const float q2 = q1 * q1, q4 = q2 * q2, q8 = q4 * q4;
            const float Fc = C.f_fb_d[1] * (q4 * q8 - q4) * __builtin_amdgcn_rcpf(r2n + C.f_fb_eta2[1]);
#pragma unroll
            for (int d = 0; d < D; ++d) ax[d] += dxn[d] * Fc;
          } else {
            double p1 = C.fb_r0[1] / (double(rn) + C.fb_eta2[1]);
            double p2 = p1 * p1, p4 = p2 * p2, p8 = p4 * p4;
            double Fc = C.fb_d[1] * (p4 * p8 - p4) / (double(r2n) + C.fb_eta2[1]);
#pragma unroll
            for (int d = 0; d < D; ++d) ax[d] += float(double(dxn[d]) * Fc);
          }
        }
#pragma unroll
        for (int d = 0; d < D; ++d) acc[d].add(ax[d]);
      }, &kt);
      };
      if (F.synced_prev) gather(std::true_type{});
      else gather(std::false_type{});
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
      const float Pn = eos_f(rv, pta, C);
      B.PN[a] = Pn;
      if (F.corr_on) B.RH2[a] = rv;
      const float ml = pta > 1 ? 1.0f : 0.0f;
      double xn[3], vn[3];
      PartnerN<D> rec{};
      double fr[3];
      particle_frame<D, R>(B, C, G, a, fr);
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
        put_x(rec, d, xv, fr);
        rec.v[d] = nv;
        if (F.corr_on) {
          B.VH2[d][a] = nv;
          B.XH2[d][a] = xv;
        }
      }
      rec.rho = rv;
      set_pt(rec, pta, Pn);
      st_rec(B.PRN, a, rec, B.n);  // partner record of the next step's pass 1
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
  block_max<BS>(red_cdx, active, B.RED, m.y);
  if (F.ts) {
    block_max<BS>(red_acc, active, B.RED + kRedSlots, m.y);
    if (F.visc_in_pass2) block_max<BS>(red_visc, active, B.RED + 2 * kRedSlots, m.y);
  }
}

// ------------------------------------------------------------ host launchers
namespace {
// Blocks of BS threads covering every problem of L: the 256-particle blocks of
// the particle map, twice as many 128-thread blocks, or the 512-particle map.
template <int BS>
unsigned pass_grid(const LaunchSet& L) {
  return BS == 256 ? L.pblocks : BS == 128 ? 2 * L.pblocks : L.pblocks512;
}
template <int D, typename R, bool Pipe, int W, int BS>
void pass1_as(const LaunchSet& L, hipStream_t s, size_t lds) {
  with_scenes<R>(L, [&](auto ss) { k_pass1<D, R, decltype(ss), Pipe, W, BS><<<pass_grid<BS>(L), BS, lds, s>>>(ss); });
}
template <int D, typename R, bool Pipe, int W, int BS>
void pass2_as(const LaunchSet& L, hipStream_t s, size_t lds) {
  with_scenes<R>(L, [&](auto ss) { k_pass2<D, R, decltype(ss), Pipe, W, BS><<<pass_grid<BS>(L), BS, lds, s>>>(ss); });
}
// Unused dynamic LDS per block of BS threads that leaves room for at most
// kPassCapWaves waves per SIMD (B = 4·kPassCapWaves·64/BS blocks per CU: B
// blocks fit the CU's LDS, B + 1 do not); 0 when the cap would need more LDS
// than a block may have.  Queried once per device.
size_t cap_lds(int bs) {
  static std::atomic<int64_t> per_cu[64], per_block[64];  // 0: not queried yet
  int dev = 0;
  GCL_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return 0;
  if (per_cu[dev].load() == 0) {
    int cu = 0, blk = 0;
    GCL_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    GCL_HIP(hipDeviceGetAttribute(&blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    per_block[dev].store(blk);
    per_cu[dev].store(cu);
  }
  const size_t cu = size_t(per_cu[dev].load()), blk = size_t(per_block[dev].load());
  const size_t blocks = size_t(4 * kPassCapWaves * 64 / bs);
  const size_t lds = cu / (blocks + 1) + 1024;  // the static LDS of a pass block is below 1 KiB
  return blocks * lds <= cu && lds <= blk ? lds : 0;
}
// code generation cg: 0 pipelined; 1 not pipelined; 2 pipelined, >= 3
// waves/SIMD; 3 not pipelined, >= 5 waves/SIMD
template <int D, typename R, int BS>
void pass1_cg(const LaunchSet& L, hipStream_t s, int cg, size_t lds) {
  switch (cg) {
    case 1: pass1_as<D, R, false, 1, BS>(L, s, lds); break;
    case 2: pass1_as<D, R, true, 3, BS>(L, s, lds); break;
    case 3: pass1_as<D, R, false, 5, BS>(L, s, lds); break;
    default: pass1_as<D, R, true, 1, BS>(L, s, lds); break;
  }
}
template <int D, typename R, int BS>
void pass2_cg(const LaunchSet& L, hipStream_t s, int cg, size_t lds) {
  switch (cg) {
    case 1: pass2_as<D, R, false, 1, BS>(L, s, lds); break;
    case 2: pass2_as<D, R, true, 3, BS>(L, s, lds); break;
    case 3: pass2_as<D, R, false, 5, BS>(L, s, lds); break;
    default: pass2_as<D, R, true, 1, BS>(L, s, lds); break;
  }
}
}  // namespace

// variant = cg + kPassCodegen·bs + kPassCodegen·kPassBlockSizes·oc (bs 0: 256,
// 1: 128, 2: 512 threads per block; oc 1: occupancy cap)
template <int D, typename R>
void launch_pass1(const LaunchSet& L, hipStream_t s, int variant) {
  const int cg = variant % kPassCodegen, bs = pass_block_size(variant);
  const size_t lds = pass_cap(variant) ? cap_lds(bs) : 0;
  switch (bs) {
    case 128: pass1_cg<D, R, 128>(L, s, cg, lds); break;
    case 512: pass1_cg<D, R, 512>(L, s, cg, lds); break;
    default: pass1_cg<D, R, 256>(L, s, cg, lds); break;
  }
}
template <int D, typename R>
void launch_pass2(const LaunchSet& L, hipStream_t s, int variant) {
  const int cg = variant % kPassCodegen, bs = pass_block_size(variant);
  const size_t lds = pass_cap(variant) ? cap_lds(bs) : 0;
  switch (bs) {
    case 128: pass2_cg<D, R, 128>(L, s, cg, lds); break;
    case 512: pass2_cg<D, R, 512>(L, s, cg, lds); break;
    default: pass2_cg<D, R, 256>(L, s, cg, lds); break;
  }
}

#define GCLSPH_INSTANTIATE_PASS(D, R)                                     \
  template void launch_pass1<D, R>(const LaunchSet&, hipStream_t, int);  \
  template void launch_pass2<D, R>(const LaunchSet&, hipStream_t, int);

GCLSPH_INSTANTIATE_PASS(2, double)
GCLSPH_INSTANTIATE_PASS(3, double)
GCLSPH_INSTANTIATE_PASS(2, float)
GCLSPH_INSTANTIATE_PASS(3, float)

}  // namespace gclsph
