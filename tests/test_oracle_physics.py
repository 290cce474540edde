"""Pins for the oracle's pair physics, each fixed by mathematics rather than by
the oracle's own code: continuum limits of the SPH sums on a regular lattice
(the Laplacian and viscosity operators, the kernel's partition of unity and
second moment), exact cancellations and ratios that follow from the
definitions (hydrostatic δ-SPH correction, Shepard renormalisation, elastic
reflection), zero crossings of the pairwise forces, finite differences and
brute-force counts.  Each test names the mistake it catches.

Lattice facts used below (regular square lattice, spacing dx, kernel support
H = 2h, h = 1.2·√2·dx, WendlandC2 with ∫W = 1, W' = dW/dr ≤ 0):
* Σ_j V_j (f_i - f_j) (x_ij·∇W_ij) / r² → ½ ∇²f   (Brookshaw; ∫ r W'(r) dV = -d)
* Σ_j V_j W_ij Δy² → ∫ W y² dV = ½ ∫ W r² dV
* Σ_j V_j (v_ij·x_ij) x_ij W'/(r r²) → -[∇²v + 2∇(∇·v)] / (2(d+2))   (isotropic 4th moment)
"""
import math
import os

import numpy as np
import pytest
from conftest import GOLDEN
from scipy import integrate


def _run(O, S, x, y, rho, ptype, v=None, dx=0.02, dt=1e-12, nsteps=1, **over):
    p, dist = S.standard_params(dx, 2, dt=dt)
    for k, val in over.items():
        setattr(p, k, val)
    sc = S.Scene(2, (np.array(x, float), np.array(y, float)), np.array(rho, float), np.array(ptype, np.int32), p,
                 dist, (p.H, p.H), v=v)
    cl = O.CellList(sc.points, sc.cellsize, sc.dist)
    sph = O.SPHProblem(cl, sc.params, sc.rho, sc.ptype, v=sc.v)
    sph.stepsolve(nsteps)
    return p, sph


def _lattice(S, n=30, dx=0.02):
    g = 1.0 + np.arange(n) * dx
    X, Y = np.meshgrid(g, g, indexing="ij")
    x, y = X.ravel().copy(), Y.ravel().copy()
    xc, yc = x.mean(), y.mean()
    # particles whose whole kernel support lies inside the lattice
    inner = (np.abs(x - xc) < 0.08) & (np.abs(y - yc) < 0.08)
    return x, y, xc, yc, inner


QUIET = dict(g=0.0, alpha=0.0, cspmn=0)  # no gravity, no artificial viscosity, no CSPM


# --------------------------------------------------------------------------
# ∂ρ∂tDDT! (src/gpukernels_sph.jl:542-631): continuity + δ-SPH diffusion with
# the hydrostatic correction ρ0·powfancy7th(1 ± DDTgz·Δx₂) - ρ0
# --------------------------------------------------------------------------


def test_ddt_is_hdelta_c0_laplacian_of_density(O, S):
    """At rest (no continuity term) and without gravity (no correction) the
    diffusion term of particle i is 2hδc0 Σ_j V_j (ρ_j - ρ_i)(-x_ij·∇W)/(r²+η²)
    = hδc0 ∇²ρ in the continuum limit.  A quadratic ρ = ρ0 + A(x-xc)² has
    ∇²ρ = 2A everywhere.  Catches a wrong DDTkh (h↔H, a dropped 2, δ or c0),
    a wrong sign of ρ_j - ρ_i and a wrong volume factor m0/ρ."""
    x, y, xc, yc, inner = _lattice(S)
    A = 50.0
    p, sph = _run(O, S, x, y, 1000.0 + A * (x - xc) ** 2, np.full(len(x), 2), **QUIET)
    got = sph.get("drho")[inner]
    expect = p.h * p.delta * p.c0 * 2 * A
    assert np.all(np.abs(got / expect - 1) < 0.04)  # measured 0.978-0.980 (η² and the lattice)


def test_ddt_hydrostatic_correction_cancels_the_hydrostatic_pair(O, S):
    """Two fluid particles at rest, j a distance d above i, with the Tait-
    hydrostatic densities ρ_i = ρ0, ρ_j = ρ0(1 - kd)^(1/7), k = ρ0 g / B.
    i's term is K_i (ρ_j - ρ_i - ρ0((1 + k(y_i - y_j))^(1/7) - 1)) = 0 exactly;
    j's mirrored term, with (1 - k(y_i - y_j))^(1/7), is the closed-form
    fraction R of its uncorrected value (same run with g = 0, which removes
    the correction only).  Catches a flipped sign of either correction, the
    correction on the wrong axis, and ρ_i/ρ_j swapped in the difference."""
    p, _ = S.standard_params(0.02, 2)
    B = p.c0 ** 2 * p.rho0 / p.gamma
    k = p.rho0 * p.g / B
    d = 0.5 * p.H
    rho = [p.rho0, p.rho0 * (1 - k * d) ** (1 / 7)]
    args = ([0.5, 0.5], [0.5, 0.5 + d], rho, [2, 2])
    _, corr = _run(O, S, *args, dt=1e-14, alpha=0.0, cspmn=0)
    _, raw = _run(O, S, *args, dt=1e-14, alpha=0.0, cspmn=0, g=0.0)
    dc, dr = corr.get("drho"), raw.get("drho")
    assert dr[0] < 0 < dr[1]  # uncorrected: diffuses from the denser (lower) particle to the lighter one
    # powfancy7th carries a relative error below 1e-12 (pinned above), i.e. an
    # absolute error of 1e-12·ρ0 in each bracket whose exact value is known
    raw_bracket = p.rho0 * (1 - (1 - k * d) ** (1 / 7))
    slack = 2 * 1e-12 * p.rho0 / raw_bracket
    assert abs(dc[0] / dr[0]) < slack  # measured 1e-8, against an uncorrected size of 1
    R = (2 - (1 - k * d) ** (1 / 7) - (1 + k * d) ** (1 / 7)) / (1 - (1 - k * d) ** (1 / 7))
    assert R > 1e-4 and abs(dc[1] / dr[1] - R) < slack


def test_ddt_density_step_diffuses_horizontally_without_correction(O, S):
    """Equal height: the correction is ρ0(pf7(1) - 1) ≈ 0, so gravity changes
    nothing, and a density step diffuses from high to low ρ."""
    args = ([0.5, 0.5 + 0.4 * 0.0679], [0.5, 0.5], [1003.0, 1000.0], [2, 2])
    _, a = _run(O, S, *args, alpha=0.0, cspmn=0)
    _, b = _run(O, S, *args, alpha=0.0, cspmn=0, g=0.0)
    da, db = a.get("drho"), b.get("drho")
    assert da[0] < 0 < da[1]
    assert np.allclose(da, db, rtol=1e-9)


def test_ddt_fixture_at_tait_hydrostatic_density(O, S):
    """The reference's Dp0.02 fixture (test/input) at rest with the exact
    Tait-hydrostatic density everywhere: near the floor and the free surface
    the neighbourhoods are one-sided, so the uncorrected diffusion term (g = 0
    run) is large there; the correction cancels it to second order in kΔy."""
    pts, _, ptype = S.load_dualsphysics_csv(os.path.join(GOLDEN, "FluidPoints_Dp0.02.csv"),
                                            os.path.join(GOLDEN, "BoundaryPoints_Dp0.02.csv"))
    p, _ = S.standard_params(0.02, 2)
    fl = ptype == 2
    exact = S.hydrostatic_rho(pts[1], pts[1][fl].max(), p.rho0, p.g, p.c0)
    _, corr = _run(O, S, pts[0], pts[1], exact, ptype, alpha=0.0, cspmn=0)
    _, raw = _run(O, S, pts[0], pts[1], exact, ptype, alpha=0.0, cspmn=0, g=0.0)
    dc, dr = np.abs(corr.get("drho")[fl]), np.abs(raw.get("drho")[fl])
    assert dr.max() > 5.0
    assert dc.max() < 0.02 * dr.max()  # measured 0.004


# --------------------------------------------------------------------------
# ∂v∂t_av! (src/gpukernels_sph.jl:852-901) and ∂v∂t_visc! (:952-991)
# --------------------------------------------------------------------------


def test_artificial_viscosity_is_monaghans_effective_viscosity(O, S):
    """v_x = -β(x-xc) - γ(x-xc)² compresses every pair near the centre, so
    the Monaghan term is active for all of them and its continuum limit is
    αc0h/(2(d+2)) [∇²v + 2∇(∇·v)] = -3αc0hγ/4 (d = 2) along x.  Catches h↔H,
    (ρi+ρj) for their mean, a dropped c0 or α, a wrong η² and a wrong sign."""
    x, y, xc, yc, inner = _lattice(S)
    beta, gam, alpha = 1.0, 0.5, 0.1
    vx = -beta * (x - xc) - gam * (x - xc) ** 2
    p, sph = _run(O, S, x, y, np.full(len(x), 1000.0), np.full(len(x), 2), v=(vx, np.zeros_like(x)), dt=1e-10,
                  g=0.0, alpha=alpha, cspmn=0)
    a = sph.get("acc")
    expect = -3 * alpha * p.c0 * p.h * gam / 4
    assert np.all(np.abs(a[0][inner] / expect - 1) < 0.03)  # measured 1.0014
    assert np.abs(a[1][inner]).max() < 1e-9 * abs(expect)


def test_artificial_viscosity_ignores_receding_pairs(O, S):
    """The same field reversed (every pair recedes): the switch cond < 0 is
    never true, so the term vanishes."""
    x, y, xc, yc, inner = _lattice(S)
    vx = (x - xc) + 0.5 * (x - xc) ** 2
    _, sph = _run(O, S, x, y, np.full(len(x), 1000.0), np.full(len(x), 2), v=(vx, np.zeros_like(x)), dt=1e-10,
                  g=0.0, alpha=0.1, cspmn=0)
    assert np.abs(sph.get("acc")[:, inner]).max() < 1e-6


def test_laminar_viscosity_is_nu_laplacian(O, S):
    """Morris' operator 4ν m0 (x_ij·∇W)/((ρi+ρj)(r²+η²)) v_ij → ν∇²v: for the
    shear field v_x = a(y-yc)², a_x = 2νa.  Catches (ρi+ρj)/2, a dropped 4 or
    m0, the sign of v_ij."""
    x, y, xc, yc, inner = _lattice(S)
    aa, nu = 2.0, 0.1
    _, sph = _run(O, S, x, y, np.full(len(x), 1000.0), np.full(len(x), 2), v=(aa * (y - yc) ** 2, np.zeros_like(x)),
                  dt=1e-10, nu=nu, **QUIET)
    a = sph.get("acc")
    assert np.all(np.abs(a[0][inner] / (2 * nu * aa) - 1) < 0.04)  # measured 0.980
    assert np.abs(a[1][inner]).max() < 1e-9


# --------------------------------------------------------------------------
# xsphcorr! + update_xsphcorr! (src/gpukernels_sph.jl:1745-1838)
# --------------------------------------------------------------------------


def test_xsph_is_epsilon_times_kernel_second_moment(O, S):
    """x += Δt Σ_j 2m0ε W_ij (v_j - v_i)/(ρi+ρj).  For v_x = a(y-yc)² the
    linear part cancels and the sum is εa Σ_j V_j W_ij Δy² → εa ∫W y² dV
    (quadrature of the pinned kernel).  Catches v_i - v_j, (ρi+ρj)/2, ∇W for W,
    a dropped ε, and correcting v instead of x."""
    x, y, xc, yc, inner = _lattice(S)
    aa, eps, dt = 2.0, 0.5, 1e-3
    v = (aa * (y - yc) ** 2, np.zeros_like(x))
    rho, pt = np.full(len(x), 1000.0), np.full(len(x), 2)
    p, base = _run(O, S, x, y, rho, pt, v=v, dt=dt, **QUIET)
    _, corr = _run(O, S, x, y, rho, pt, v=v, dt=dt, xsph_eps=eps, **QUIET)
    H = p.H
    m2 = 0.5 * integrate.quad(lambda r: O.wendland_W(2, r / H, 1 / H) * 2 * math.pi * r ** 3, 0, H, epsabs=0,
                              epsrel=1e-13)[0]
    shift = (corr.get("x")[0] - base.get("x")[0]) / dt
    assert np.all(np.abs(shift[inner] / (eps * aa * m2) - 1) < 0.01)  # measured 0.99993
    assert np.array_equal(corr.get("v"), base.get("v"))


# --------------------------------------------------------------------------
# dpcreg! + update_dpcreg! (src/gpukernels_sph.jl:1485-1632)
# --------------------------------------------------------------------------


def _pair(O, S, r, axis=0, vi=(0.0, 0.0), vj=(0.0, 0.0), pt=(2, 2), rho=(1000.0, 1000.0), dt=1e-12, **over):
    x, y = [0.5, 0.5], [0.5, 0.5]
    (x if axis == 0 else y)[1] += r
    v = (np.array([vi[0], vj[0]]), np.array([vi[1], vj[1]]))
    return _run(O, S, x, y, rho, pt, v=v, dt=dt, **over)


def test_dpc_reflects_a_close_approaching_pair(O, S):
    """u = r/l0 below dpckernlim, approaching (Δv·Δx < 0): each side gets
    ∓ (Δv·Δx)/(r²+η²) Δx, so the normal relative velocity v_n becomes
    v_n (η² - r²)/(r² + η²) (an elastic reflection as η → 0) and the
    tangential velocities are untouched."""
    l0, r, u, wi, wj = 0.01, 0.003, 0.7, 0.2, -0.1
    _, sph = _pair(O, S, r, vi=(u, wi), vj=(-u, wj), dpc_l0=l0, **QUIET)
    v = sph.get("v")
    eta2 = (0.1 * l0) ** 2
    assert v[0][0] - v[0][1] == pytest.approx(2 * u * (eta2 - r * r) / (r * r + eta2), rel=1e-9)
    assert v[0][0] + v[0][1] == pytest.approx(0.0, abs=1e-12)  # momentum of the pair
    assert v[1][0] == pytest.approx(wi, rel=1e-8) and v[1][1] == pytest.approx(wj, rel=1e-8)


def test_dpc_pressure_branch_floor_cap_and_half_length_normalisation(O, S):
    """At rest (Δv·Δx = 0) the pair is pushed apart by 2Δt Pᵇ/((ρi+ρj)(r²+η²)) Δx
    with Pᵇ = χ max(min(λ|Pi+Pj|, λPmax), Pmin), χ = √(W(u)/W(½)).  At u = ½,
    χ = 1; with P = 0 (ρ = ρ0) Pᵇ is the floor Pmin; with a large pressure it
    is the cap λPmax whatever the pressure."""
    l0, dt = 0.01, 1e-5
    r = 0.5 * l0
    eta2 = (0.1 * l0) ** 2
    for pmin in (500.0, 1000.0):
        _, sph = _pair(O, S, r, dt=dt, dpc_l0=l0, dpc_pmin=pmin, **QUIET)
        v = sph.get("v")
        expect = 2 * dt * pmin / (2000.0 * (r * r + eta2)) * r
        assert -v[0][0] == pytest.approx(expect, rel=1e-9) and v[0][1] == pytest.approx(expect, rel=1e-9)
    caps = []
    for rho in (1010.0, 1020.0):
        # the pair's own pressure moves it during the step: the DPC share is the
        # difference to the same step without DPC (applied after the update)
        kw = dict(rho=(rho, rho), dt=1e-6, dpc_lambda=1.0, dpc_pmax=2000.0, dpc_pmin=1.0, **QUIET)
        _, with_dpc = _pair(O, S, r, dpc_l0=l0, **kw)
        _, without = _pair(O, S, r, **kw)
        dv = with_dpc.get("v")[0][1] - without.get("v")[0][1]
        caps.append(dv * (with_dpc.get("rho")[0] + with_dpc.get("rho")[1]))  # Δv ∝ Pᵇ/(ρi+ρj)
    assert caps[0] == pytest.approx(caps[1], rel=1e-6)
    assert caps[0] == pytest.approx(2 * 1e-6 * 2000.0 / (r * r + eta2) * r, rel=1e-6)


# --------------------------------------------------------------------------
# ∂v∂tpF! (src/gpukernels_sph.jl:1373-1409) and fbmolforce! (:1875-1913)
# --------------------------------------------------------------------------


def test_surface_tension_changes_sign_at_two_thirds_h(O, S):
    """s cos(1.5π r/2h)/(r + 0.1h) Δx/m0: repulsive below r = 2h/3, attractive
    above (cos(1.5π r/2h) = 0 at 2h/3).  Catches a wrong cosine argument, 2h
    read as 2·h on the wrong side of the division, and the sign of Δx."""
    p, _ = S.standard_params(0.02, 2)
    rz = 2 * p.h / 3
    acc = {}
    for f in (0.5, 0.999, 1.0, 1.001):
        _, sph = _pair(O, S, f * rz, s=0.1, **QUIET)
        acc[f] = sph.get("acc")[0]  # x-acceleration of both particles; particle 0 sits at -x of particle 1
    assert acc[0.5][0] < 0 < acc[0.5][1] and acc[0.999][0] < 0  # pushed apart
    assert acc[1.001][0] > 0 > acc[1.001][1]  # pulled together
    assert abs(acc[1.0][0]) < 1e-9 * abs(acc[0.5][0])
    r = 0.5 * rz
    assert -acc[0.5][0] == pytest.approx(0.1 * math.cos(math.pi / 4) / (r + 0.1 * p.h) * r / p.m0, rel=1e-12)


def test_wall_force_zero_crossing_and_cutoff_use_bound_l(O, S):
    """Pass 2's wall force (bound_D, bound_l = r0) is d(q¹² - q⁴)/(r² + η²) Δx
    with q = r0/(r + η²), η² = 0.01 r0²: repulsive for r < r0 - η², attractive
    up to r0, zero from r0 on.  A fluid particle above a ptype -1 particle; r0
    = 1.5 dx differs from pass 1's hard-coded 2dx, so the zero crossing also
    shows that the reported acceleration is pass 2's."""
    rb = 1.5 * 0.02
    rz = rb - 0.01 * rb * rb
    acc = {}
    for r in (rz - 1e-6, rz + 1e-6, 0.5 * (rz + rb), rb + 1e-7):
        _, sph = _pair(O, S, r, axis=1, pt=(-1, 2), bound_l=rb, **QUIET)
        acc[r] = sph.get("acc")[1][1]  # fluid particle's vertical acceleration
    a = list(acc.values())
    assert a[0] > 0 > a[1] and a[2] < 0  # pushed up, then pulled down
    assert a[3] == 0.0
    # gravity-free, pressure-free: the fluid particle alone carries the force
    r = 0.6 * rb
    _, sph = _pair(O, S, r, axis=1, pt=(-1, 2), bound_l=rb, bound_D=0.4, **QUIET)
    q = rb / (r + 0.01 * rb * rb)
    assert sph.get("acc")[1][1] == pytest.approx(0.4 * (q ** 12 - q ** 4) / (r * r + 0.01 * rb * rb) * r, rel=1e-12)


# --------------------------------------------------------------------------
# sph∑W! / sph∑∇W! (src/gpukernels_sph.jl:129-156, 322-359)
# --------------------------------------------------------------------------


def test_sumW_is_a_partition_of_unity(O, S):
    """(Σ_j W_ij + W(0))·dx² → ∫W = 1 inside a regular lattice (the reference
    sums over pairs, without the self term).  Catches one-sided accumulation
    (half), u = r/h for r/H, and a wrong σ."""
    x, y, xc, yc, inner = _lattice(S)
    p, sph = _run(O, S, x, y, np.full(len(x), 1000.0), np.full(len(x), 2), **QUIET)
    w0 = O.wendland_W(2, 0.0, 1 / p.H)
    unity = (sph.get("sumW")[inner] + w0) * p.dx ** 2
    assert np.all(np.abs(unity - 1) < 0.01)  # measured 1.0031


def test_sumgradW_is_the_gradient_of_sumW(O, S):
    """Σ_j ∇W_ij with ∇W = Δx W'/r, Δx = x_i - x_j, is ∂/∂x_i of Σ_j W_ij: a
    central difference of sumW over a shift of one particle reproduces it, at
    a corner, on an edge and inside.  Catches Δx = x_j - x_i, adding (not
    subtracting) the j side, and ∇W without the 1/r."""
    n = 16
    g = 1.0 + np.arange(n) * 0.02
    X, Y = np.meshgrid(g, g, indexing="ij")
    x0, y0 = X.ravel(), Y.ravel()
    rho, pt = np.full(n * n, 1000.0), np.full(n * n, 2)
    _, base = _run(O, S, x0, y0, rho, pt, **QUIET)
    sg = base.get("sumgradW")
    d = 1e-6 * 0.02
    for k in (0, 5, n * n // 2 + 8):
        for ax in (0, 1):
            vals = []
            for sgn in (1, -1):
                x, y = x0.copy(), y0.copy()
                (x if ax == 0 else y)[k] += sgn * d
                _, s = _run(O, S, x, y, rho, pt, **QUIET)
                vals.append(s.get("sumW")[k])
            fd = (vals[0] - vals[1]) / (2 * d)
            assert fd == pytest.approx(sg[ax][k], rel=1e-6, abs=1e-3)
    assert sg[0][0] > 1e4 and sg[1][0] > 1e4  # the corner's sum rises toward its neighbours


def test_sum_gates_exclude_fluid_boundary_pairs(O, S):
    """!(A ⊻ B): a fluid particle's sumW counts only fluid neighbours."""
    x, y, xc, yc, inner = _lattice(S)
    pt = np.where(x < xc, 2, 0)
    p, sph = _run(O, S, x, y, np.full(len(x), 1000.0), pt, **QUIET)
    _, ref = _run(O, S, x, y, np.full(len(x), 1000.0), np.full(len(x), 2), **QUIET)
    sw, full = sph.get("sumW"), ref.get("sumW")
    left = inner & (x < xc - p.H)
    assert np.allclose(sw[left], full[left], rtol=1e-13)
    edge = inner & (x < xc) & (x > xc - p.dx)
    assert np.all(sw[edge] < 0.8 * full[edge])


# --------------------------------------------------------------------------
# cspmcorr! (src/gpukernels_sph.jl:1673-1719)
# --------------------------------------------------------------------------


def test_cspm_is_the_shepard_ratio_of_the_neighbours_density(O, S):
    """ρ_i ← Σ_j m0 W_ij / Σ_j m0 W_ij/ρ_j (no self term): a particle whose
    neighbours all have density c gets exactly c, whatever its own density;
    particles outside the gate (ptype ≤ 0 here) get 0/0 and keep theirs.
    Catches ρ_i for ρ_j, a self term, and renormalising boundary particles."""
    n = 16
    g = 1.0 + np.arange(n) * 0.02
    X, Y = np.meshgrid(g, g, indexing="ij")
    x, y = X.ravel(), Y.ravel()
    rho = np.full(n * n, 1002.0)
    k = n * n // 2 + 8
    rho[k] = 1005.0
    pt = np.full(n * n, 2)
    pt[:n] = 0
    rho[:n] = 1001.0
    _, sph = _run(O, S, x, y, rho, pt, g=0.0, alpha=0.0, cspmn=1)
    out = sph.get("rho")
    assert out[k] == pytest.approx(1002.0, rel=1e-11)
    assert np.all(out[:n] == 1001.0)
    # a uniform field is returned unchanged
    _, uni = _run(O, S, x, y, np.full(n * n, 1002.0), np.full(n * n, 2), g=0.0, alpha=0.0, cspmn=1)
    assert np.abs(uni.get("rho") / 1002.0 - 1).max() < 1e-11


# --------------------------------------------------------------------------
# мaxpairs (src/gpukernels_cls.jl:109-286)
# --------------------------------------------------------------------------


@pytest.mark.parametrize("dim,n,dist,csf,seed", [(2, 700, 0.07, 1.0, 1), (2, 500, 0.05, 1.7, 2), (3, 600, 0.16, 1.0, 3),
                                                 (3, 400, 0.2, 1.3, 4)])
def test_maxpairs_counts_pairs_in_adjacent_cells(O, S, dim, n, dist, csf, seed):
    """Σ_cells n(n-1)/2 + n·Σ_{half stencil} m counts every unordered pair of
    particles whose cells touch (Chebyshev distance ≤ 1) exactly once, when
    the half stencil holds one of each ± offset pair: equal to a brute-force
    count over all pairs."""
    pts = S.uniform_points(seed, n, dim)
    cl = O.CellList(pts, dist * csf, dist)
    pc, _ = cl.cells()
    adj = np.ones((n, n), bool)
    for c in pc.astype(np.int64):
        adj &= np.abs(c[:, None] - c[None, :]) <= 1
    assert cl.maxpairs() == int(np.triu(adj, 1).sum())
