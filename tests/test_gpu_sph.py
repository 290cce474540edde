"""GPU SPH step vs the CPU oracle.

Tolerances.  FP64 path: per-pair terms are evaluated in the reference's
operation order without FMA contraction, so the only difference from the
oracle is the order of the per-particle sums (the reference's own FP64 atomic
order is nondeterministic); fields agree to ~1e-13 relative after tens of
steps.  FP32-split path: FP32 pair terms summed in FP64, density as ρ-ρ0, FP64
positions; bounds follow the acceptance target of 1e-5 relative density and
1e-4·dx positions after 1000 steps."""
import math
import os

import numpy as np
import pytest
from cases import golden_cases
from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "rho", "P", "acc", "drho", "sumW", "sumgradW")


def gpu_fields(G, prob):
    return {"x": np.array(G.get_points(prob)), "v": np.array(G.get_velocity(prob)), "rho": G.get_density(prob),
            "P": G.get_pressure(prob), "acc": np.array(G.get_acceleration(prob)), "drho": G.get_drho(prob),
            "sumW": G.get_sumw(prob), "sumgradW": np.array(G.get_sumgradw(prob))}


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# FP64 tolerances (relative to the field's max magnitude)
TOL64 = {"x": 1e-13, "v": 1e-10, "rho": 1e-13, "P": 1e-10, "acc": 1e-9, "drho": 1e-8, "sumW": 1e-12,
         "sumgradW": 1e-9}


@pytest.mark.parametrize("name", sorted(golden_cases()))
def test_f64_matches_golden(G, name):
    sc, steps, ts = golden_cases()[name]
    ref = np.load(os.path.join(GOLDEN, f"oracle_{name}.npz"))
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    assert np.array_equal(np.unique(G.neighborlistview(cl), axis=0), ref["pairs0"])
    prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, v=sc.v, precision="f64")
    upd, mx = G.stepsolve_(prob, steps, timestepping=ts)
    assert upd == ref["updaten"]
    assert mx == pytest.approx(float(ref["maxcdx"]), rel=1e-12)
    assert prob.dt == pytest.approx(float(ref["dt"]), rel=1e-12)
    assert prob.etime == pytest.approx(float(ref["etime"]), rel=1e-12)
    got = gpu_fields(G, prob)
    for f in FIELDS:
        assert rel_err(got[f], ref[f]) < TOL64[f], (f, rel_err(got[f], ref[f]))
    assert np.array_equal(np.unique(G.neighborlistview(cl), axis=0), ref["pairs_end"])


@pytest.mark.parametrize("name", sorted(golden_cases()))
def test_f32split_matches_golden(G, name):
    sc, steps, ts = golden_cases()[name]
    ref = np.load(os.path.join(GOLDEN, f"oracle_{name}.npz"))
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, v=sc.v, precision="f32split")
    upd, _ = G.stepsolve_(prob, steps, timestepping=ts)
    assert upd == ref["updaten"]
    got = gpu_fields(G, prob)
    dx = sc.params.dx
    assert np.abs(got["x"] - ref["x"]).max() < 1e-5 * dx
    assert np.abs(got["rho"] / ref["rho"] - 1).max() < 1e-6
    assert rel_err(got["v"], ref["v"]) < 1e-4


def _pair(G, O, sc, precision="f64"):
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    oc = O.CellList(sc.points, sc.cellsize, sc.dist)
    prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, v=sc.v, precision=precision)
    ref = O.SPHProblem(oc, sc.params, sc.rho, sc.ptype, v=sc.v)
    return cl, oc, prob, ref


def test_config0_f64_50_steps(G, O, S):
    """The 2D dam break at dx=0.02 (6,491 particles), reference physics."""
    sc = S.config0()
    cl, oc, prob, ref = _pair(G, O, sc)
    for n in (1, 9, 40):
        assert G.stepsolve_(prob, n) == pytest.approx(ref.stepsolve(n), rel=1e-12)
    assert np.abs(np.array(G.get_points(prob)) - ref.get("x")).max() < 1e-12 * sc.params.dx
    assert np.abs(G.get_density(prob) / ref.get("rho") - 1).max() < 1e-13
    assert rel_err(np.array(G.get_velocity(prob)), ref.get("v")) < 1e-11
    assert rel_err(np.array(G.get_acceleration(prob)), ref.get("acc")) < 1e-9
    assert rel_err(G.get_pressure(prob), ref.get("P")) < 1e-10
    assert np.array_equal(np.unique(G.neighborlistview(cl), axis=0), np.unique(oc.neighborlistview(), axis=0))


def test_config0_f32split_50_steps(G, O, S):
    sc = S.config0()
    cl, oc, prob, ref = _pair(G, O, sc, "f32split")
    for n in (10, 40):
        assert G.stepsolve_(prob, n)[0] == ref.stepsolve(n)[0]
    assert np.abs(np.array(G.get_points(prob)) - ref.get("x")).max() < 1e-5 * sc.params.dx
    assert np.abs(G.get_density(prob) / ref.get("rho") - 1).max() < 1e-6


def test_reference_fixture_with_timestepping(G, O, S):
    """The reference's Dp0.02 fixture (ptype 0/1 walls) with adaptive Δt and the examples' dist = 1.1H."""
    pts, rho, ptype = S.load_dualsphysics_csv(os.path.join(GOLDEN, "FluidPoints_Dp0.02.csv"),
                                              os.path.join(GOLDEN, "BoundaryPoints_Dp0.02.csv"))
    p, dist = S.standard_params(0.02, 2, dist_factor=1.1)
    sc = S.Scene(2, pts, rho, ptype, p, dist, (dist, dist))
    cl, oc, prob, ref = _pair(G, O, sc)
    r1 = G.stepsolve_(prob, 30, timestepping=True)
    r2 = ref.stepsolve(30, True)
    assert r1[0] == r2[0] and r1[1] == pytest.approx(r2[1], rel=1e-10)
    assert prob.dt == pytest.approx(ref.dt, rel=1e-10)
    assert np.abs(np.array(G.get_points(prob)) - ref.get("x")).max() < 1e-12 * p.dx
    assert np.abs(G.get_density(prob) / ref.get("rho") - 1).max() < 1e-12
    assert np.array_equal(np.unique(G.neighborlistview(cl), axis=0), np.unique(oc.neighborlistview(), axis=0))
    assert cl.pairsn == oc.pairsn


def test_3d_dam_break_all_physics(G, O, S):
    """3D dam break (the config with viscosity, surface tension, DPC, XSPH) at a coarse spacing."""
    sc = S.config2(dx=0.04)
    sc.params.cspmn = 2
    cl, oc, prob, ref = _pair(G, O, sc)
    assert G.stepsolve_(prob, 3, timestepping=True)[0] == ref.stepsolve(3, True)[0]
    assert np.abs(np.array(G.get_points(prob)) - ref.get("x")).max() < 1e-12 * sc.params.dx
    assert np.abs(G.get_density(prob) / ref.get("rho") - 1).max() < 1e-12
    assert rel_err(np.array(G.get_velocity(prob)), ref.get("v")) < 1e-10


def test_3d_f32split_particles_outside_the_grid(G, O, S):
    """3D FP32-split partner records hold positions relative to their cell's
    origin; particles outside the grid (binned into its edge cells, the grid
    here ending below the top of the column and short of the far wall) sit
    several cells from that origin and keep their neighbour lists and forces
    to the FP32-split tolerances against the FP64 oracle."""
    sc = S.config2(dx=0.04)
    top = 0.6  # the fluid column reaches y = 1.1: its upper part lies above the grid
    kw = dict(MIN=tuple(float(np.min(p)) for p in sc.points),
              MAX=(float(np.max(sc.points[0])) - 0.2, top, float(np.max(sc.points[2]))))
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist, **kw)
    oc = O.CellList(sc.points, sc.cellsize, sc.dist, **kw)
    assert np.sum((sc.points[1] > top) & (sc.ptype > 0)) > 1000
    prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, v=sc.v, precision="f32split")
    ref = O.SPHProblem(oc, sc.params, sc.rho, sc.ptype, v=sc.v)
    assert G.stepsolve_(prob, 4, timestepping=True)[0] == ref.stepsolve(4, True)[0]
    dx = sc.params.dx
    assert np.abs(np.array(G.get_points(prob)) - ref.get("x")).max() < 1e-5 * dx
    assert np.abs(G.get_density(prob) / ref.get("rho") - 1).max() < 1e-6
    assert rel_err(np.array(G.get_velocity(prob)), ref.get("v")) < 1e-4


def test_deterministic(G, S):
    sc = S.config0()
    out = []
    for _ in range(2):
        cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
        prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, precision="f32split")
        G.stepsolve_(prob, 20, timestepping=True)
        out.append((np.array(G.get_points(prob)), G.get_density(prob), np.array(G.get_velocity(prob))))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["f64", "f32split"])
@pytest.mark.parametrize("case", ["dam2d_allphys", "dam3d_allphys"])
def test_stored_membership_equals_stencil_scan(G, monkeypatch, prec, case):
    """Partners read from the per-particle membership store and partners found
    by scanning the stencil (the path of particles whose windows exceed the
    store's capacity, forced here with GCLSPH_NMASK_CAP) visit the same pairs in
    the same order: the steps are bit-identical, including steps that skip the
    rebuild (dist = 1.1 H) and every correction."""
    sc, steps, ts = golden_cases()[case]

    def run(cap):
        if cap is None:
            monkeypatch.delenv("GCLSPH_NMASK_CAP", raising=False)
        else:
            monkeypatch.setenv("GCLSPH_NMASK_CAP", str(cap))
        cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
        prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, v=sc.v, precision=prec)
        upd, mx = G.stepsolve_(prob, steps + 3, timestepping=ts)
        G.update_(cl)  # a user rebuild between calls refreshes the store too
        G.stepsolve_(prob, 2, timestepping=ts)
        return [upd, mx] + [np.asarray(gpu_fields(G, prob)[f]) for f in FIELDS]

    ref = run(None)
    for cap in (0, 2):
        got = run(cap)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["f64", "f32split"])
@pytest.mark.parametrize("dim", [2, 3])
def test_membership_with_candidates_exactly_at_dist(G, S, monkeypatch, prec, dim):
    """A lattice of spacing 0.25 with dist = cell size = 0.75 (all exact in
    binary) puts many candidates at exactly |x_a - x_b| = dist, where the test
    r² < dist² fails by the smallest margin: the membership store (in 3D
    bounded to |x_a - x_b| < dist by bisection over the x-ordered rows) and the
    full stencil scan (GCLSPH_NMASK_CAP=0) must give bit-identical steps."""
    import dataclasses

    if dim == 2:
        sc = S.dam_break_2d(0.25, 1.5, 1.5, 3.0, 2.5)
    else:
        sc = S.dam_break_3d(0.25, (1.5, 1.5, 1.0), (3.0, 2.5, 2.0))
    sc = dataclasses.replace(sc, dist=0.75, cellsize=(0.75,) * dim,
                             params=dataclasses.replace(sc.params, nui=0.0))
    x = np.asarray(sc.points[0])
    ties = np.count_nonzero(np.abs(x[:, None] - x[None, :]) == 0.75)
    assert ties > 1000  # the boundary case is exercised

    def run(cap):
        if cap is None:
            monkeypatch.delenv("GCLSPH_NMASK_CAP", raising=False)
        else:
            monkeypatch.setenv("GCLSPH_NMASK_CAP", str(cap))
        cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
        prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, precision=prec)
        upd, mx = G.stepsolve_(prob, 4, timestepping=True)
        return [upd, mx] + [np.asarray(gpu_fields(G, prob)[f]) for f in FIELDS]

    ref = run(None)
    got = run(0)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["f64", "f32split"])
@pytest.mark.parametrize("case", ["dam2d", "dam2d_allphys", "dam3d_allphys"])
def test_ptype_values_act_by_class(G, prec, case):
    """The pair terms ask of a particle type only whether it is > 0, > 1, < 0 or
    <= 1 (the ptype tests of src/gpukernels_sph.jl), so types moved inside
    their class (2 -> 7, -1 -> -4; 0 and 1 are classes of their own) must give
    bit-identical steps.  The 2D FP32-split partner records carry only the
    class of a partner's type, read back as -1, 0, 1 or 2; this holds them to
    it, with fluid (2), wall (0, 1) and ptype -1 particles in the scenes."""
    sc, steps, ts = golden_cases()[case]

    def run(pt):
        cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
        prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, pt, v=sc.v, precision=prec)
        upd, mx = G.stepsolve_(prob, steps + 3, timestepping=ts)
        return [upd, mx] + [np.asarray(gpu_fields(G, prob)[f]) for f in FIELDS]

    pt = np.asarray(sc.ptype)
    moved = np.where(pt > 1, pt + 5, np.where(pt < 0, pt - 3, pt)).astype(pt.dtype)
    assert (moved != pt).any()
    ref = run(pt)
    got = run(moved)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_user_update_and_set_points_between_calls(G, O, S):
    """update!(prob.system), partialupdate! and edits of prob.system.points between stepsolve! calls."""
    sc = golden_cases()["dam2d_allphys"][0]
    cl, oc, prob, ref = _pair(G, O, sc)
    G.stepsolve_(prob, 3, timestepping=True)
    ref.stepsolve(3, True)
    G.update_(cl)
    oc.update()
    assert cl.pairsn == oc.pairsn
    assert np.array_equal(np.unique(G.neighborlistview(cl), axis=0), np.unique(oc.neighborlistview(), axis=0))
    x = np.array(G.get_points(prob))
    x[0] += 1e-4
    cl.points = x
    oc.set_points(x)
    G.partialupdate_(cl)
    oc.partialupdate()
    assert np.array_equal(np.unique(G.neighborlistview(cl), axis=0), np.unique(oc.neighborlistview(), axis=0))
    G.stepsolve_(prob, 3, timestepping=True)
    ref.stepsolve(3, True)
    assert np.abs(np.array(G.get_points(prob)) - ref.get("x")).max() < 1e-12 * sc.params.dx
    assert np.abs(G.get_density(prob) / ref.get("rho") - 1).max() < 1e-12


def test_timesolve_batches(G, O, S):
    sc = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5)
    cl, oc, prob, ref = _pair(G, O, sc)
    G.timesolve_(prob, batch=7, timeframe=2e-4, verbose=False)
    while ref.etime <= 2e-4:
        ref.stepsolve(7)
    assert prob.etime == pytest.approx(ref.etime, rel=1e-12)
    assert np.abs(np.array(G.get_points(prob)) - ref.get("x")).max() < 1e-12 * sc.params.dx


def test_timesolve_progress_lines(G, O, S, capsys):
    """timesolve!'s per-batch progress (src/sphproblem.jl:456,498,510): one
    line per batch with iter, time, Δt, updn and dxpncu of that batch, then a
    closing line; the values are those stepsolve! returned."""
    sc = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5)
    cl, oc, prob, ref = _pair(G, O, sc)
    G.timesolve_(prob, batch=7, timeframe=1e-4, timestepping=True, verbose=True)
    err = capsys.readouterr().err.strip().splitlines()
    lines = [l for l in err if "iter:" in l]
    i = 0
    while ref.etime <= 1e-4:
        updn, dxpncu = ref.stepsolve(7, True)
        i += 1
        fields = dict(kv.split(": ") for kv in lines[i - 1].split("  ") if kv.count(": ") == 1)
        assert int(fields["iter"]) == i
        assert float(fields["time"]) == pytest.approx(ref.etime, rel=1e-12)
        assert float(fields["Δt"]) == pytest.approx(ref.dt, rel=1e-9)
        assert int(fields["updn"]) == updn
        assert float(fields["dxpncu"]) == pytest.approx(dxpncu, rel=1e-9, abs=1e-300)
    assert len(lines) == i and err[-1].endswith("done")


def test_gamma_other_than_7(G, O, S):
    sc = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5)
    sc.params.gamma = 5
    cl, oc, prob, ref = _pair(G, O, sc)
    G.stepsolve_(prob, 3)
    ref.stepsolve(3)
    assert np.abs(G.get_density(prob) / ref.get("rho") - 1).max() < 1e-12
    with pytest.raises(NotImplementedError):
        cl2 = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
        G.SPHProblem.from_params(cl2, sc.params, sc.rho, sc.ptype, precision="f32split")


def test_argument_errors(G, S):
    sc = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5)
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    with pytest.raises(ValueError, match="length"):
        G.SPHProblem.from_params(cl, sc.params, sc.rho[:-1], sc.ptype)
    cl32 = G.GPUCellList(sc.points, sc.cellsize, sc.dist, precision="f32")
    with pytest.raises(NotImplementedError):
        G.SPHProblem.from_params(cl32, sc.params, sc.rho, sc.ptype)


def test_ptype_getter_and_stepsolve_zero(G, O, S):
    sc = S.dam_break_2d(0.05, 0.5, 1.0, 2.0, 1.5)
    cl, oc, prob, ref = _pair(G, O, sc)
    assert np.array_equal(G.get_ptype(prob), sc.ptype)
    assert G.stepsolve_(prob, 0) == (0, 0.0)
    ref.stepsolve(0)
    assert rel_err(G.get_sumw(prob), ref.get("sumW")) < 1e-12


def test_unimplemented_sph_kernel_is_refused(G, S):
    """Only WendlandC2 is evaluated; another kernel object is an error, not
    a silent substitution."""
    sc = S.config0()
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    p = sc.params

    class CubicSpline:
        dim = 2

    args = (cl, p.dx, p.h, p.H, CubicSpline(), sc.rho, sc.ptype, p.rho0, p.m0, p.dt, p.alpha, p.c0, p.gamma,
            p.delta, p.CFL)
    with pytest.raises(NotImplementedError, match="WendlandC2"):
        G.SPHProblem(*args)
    with pytest.raises(ValueError, match="dimension"):
        G.SPHProblem(*args[:4], G.WendlandC2(np.float64, 3), *args[5:])
    prob = G.SPHProblem(*args[:4], G.WendlandC2(np.float64, 2), *args[5:])
    assert prob.dim == 2


def test_full_size_f32split_vs_f64_1000_steps(G, S):
    """The 4,047,991-particle 2D dam break after 1000 adaptive steps.

    Density: FP32-split within 1e-5 relative of FP64 everywhere.  Positions:
    the run is chaotic at the floor-adjacent fluid row (the r^-12 wall force),
    so a pointwise bound is measured against the FP64 path's own sensitivity:
    FP64 started from positions perturbed by ~1 ulp.  FP32-split must stay
    within that chaos floor, and 99.99 % of particles within 1e-4·dx."""
    sc = S.config3()

    def run(prec, pert=0.0):
        pts = sc.points
        if pert:
            pts = tuple(p * (1 + pert * (S.uniform01(5, len(p), k) - 0.5)) for k, p in enumerate(pts))
        cl = G.GPUCellList(pts, sc.cellsize, sc.dist)
        prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, precision=prec)
        G.stepsolve_(prob, 1000, timestepping=True)
        out = (np.array(G.get_points(prob)), G.get_density(prob), prob.etime, np.array(G.get_velocity(prob)))
        del prob
        return out

    x64, r64, t64, v64 = run("f64")
    xp, _, _, vp = run("f64", 4e-16)
    x32, r32, t32, v32 = run("f32split")
    # deterministic sums (no floating-point atomics): a second run of the
    # 4 M-particle scene ends bit-identical
    again = run("f32split")
    assert np.array_equal(again[0], x32) and np.array_equal(again[1], r32) and np.array_equal(again[3], v32)
    dx = sc.params.dx
    chaos = np.abs(xp - x64).max(axis=0)
    floor = chaos.max()
    err = np.abs(x32 - x64).max(axis=0)
    vchaos = np.abs(vp - v64).max(axis=0)
    verr = np.abs(v32 - v64).max(axis=0)
    print(f"density max rel {np.abs(r32 / r64 - 1).max():.3e}; position/dx max {err.max() / dx:.3e} "
          f"p99.99 {np.quantile(err, 0.9999) / dx:.3e}; FP64 chaos floor/dx {floor / dx:.3e} "
          f"p99.99 {np.quantile(chaos, 0.9999) / dx:.3e}; velocity max {verr.max():.3e} "
          f"p99.99 {np.quantile(verr, 0.9999):.3e} (FP64 floor {vchaos.max():.3e} / "
          f"{np.quantile(vchaos, 0.9999):.3e}) m/s")
    assert np.abs(r32 / r64 - 1).max() < 1e-5
    assert np.quantile(err, 0.9999) < 1e-4 * dx
    # FP32-split deviates less than FP64 does from its own 1-ulp perturbation,
    # in positions and in velocities
    assert np.quantile(err, 0.9999) <= np.quantile(chaos, 0.9999)
    assert err.max() <= floor
    assert np.quantile(verr, 0.9999) <= np.quantile(vchaos, 0.9999)
    assert verr.max() <= vchaos.max()
    assert t32 == pytest.approx(t64, rel=1e-6)


@pytest.mark.parametrize("case", ["dam2d_allphys", "dam3d_allphys"])
def test_graph_replay_equals_eager_launches(G, monkeypatch, case):
    """Calls of several steps replay a captured step graph (kept across calls
    while nothing it bakes in changes); GCLSPH_NO_GRAPH=1 launches every kernel
    eagerly.  Several calls with eager work in between (a single-step call,
    parameter edits, getters) give bit-identical states."""
    sc, steps, ts = golden_cases()[case]

    def run(no_graph):
        if no_graph:
            monkeypatch.setenv("GCLSPH_NO_GRAPH", "1")
        else:
            monkeypatch.delenv("GCLSPH_NO_GRAPH", raising=False)
        cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
        prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, v=sc.v, precision="f32split")
        out = []
        for n in (4, 1, 3, 3):
            out.append(G.stepsolve_(prob, n, timestepping=ts))
            out.append(np.array(G.get_points(prob)))
            out.append(G.get_density(prob))
        prob.alpha = 0.02  # a changed constant re-captures the step
        out.append(G.stepsolve_(prob, 3, timestepping=ts))
        out += [np.asarray(gpu_fields(G, prob)[f]) for f in FIELDS]
        return out

    ref = run(True)
    got = run(False)
    for a, b in zip(got, ref):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_config2_3d_f32split_vs_f64_500_steps(G, S):
    """The 3D dam break with viscosity, surface tension, DPC, XSPH and CSPM
    (883,376 particles at dx = 0.01) after 500 adaptive steps: FP32-split
    against FP64, per field (measured: positions 5.6e-6·dx, density 8.8e-8
    relative, velocity 5.2e-6 of the largest speed)."""
    sc = S.config2(dx=0.01)

    def run(prec):
        cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
        prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, precision=prec)
        G.stepsolve_(prob, 500, timestepping=True)
        out = (np.array(G.get_points(prob)), G.get_density(prob), np.array(G.get_velocity(prob)), prob.etime)
        del prob
        return out

    x64, r64, v64, t64 = run("f64")
    x32, r32, v32, t32 = run("f32split")
    dx = sc.params.dx
    ex = np.abs(x32 - x64).max() / dx
    er = np.abs(r32 / r64 - 1).max()
    ev = np.abs(v32 - v64).max() / np.abs(v64).max()
    print(f"config 2, 500 steps: position/dx {ex:.3e}, density rel {er:.3e}, velocity rel {ev:.3e}")
    assert ex < 1e-4
    assert er < 1e-5
    assert ev < 1e-4
    assert t32 == pytest.approx(t64, rel=1e-6)


@pytest.mark.parametrize("prec", ["f64", "f32split"])
def test_pass_variants_bit_identical(G, monkeypatch, prec):
    """The pair passes exist in variants (partner prefetch, occupancy target,
    128, 256 or 512 threads per block, and an occupancy cap: variant = codegen
    + 4·blocksize + 12·cap) that the step driver times on live steps and picks
    from; every variant, forced or autotuned, must give bit-identical steps."""
    sc, steps, ts = golden_cases()["dam2d_allphys"]

    def run(variant):
        if variant is None:
            monkeypatch.delenv("GCLSPH_PASS_VARIANT", raising=False)
        else:
            monkeypatch.setenv("GCLSPH_PASS_VARIANT", variant)
        cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
        prob = G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, v=sc.v, precision=prec)
        out = [G.stepsolve_(prob, 30, timestepping=ts), G.stepsolve_(prob, 3, timestepping=ts)]
        if variant is None:  # the tuner ran on the first call and recorded its choice
            pv = prob.pass_variants()
            assert 0 <= pv["pass1"] < 24 and 0 <= pv["pass2"] < 24
            assert pv["pass1_threads"] in (128, 256, 512) and pv["pass2_threads"] in (128, 256, 512)
        return out + [np.asarray(gpu_fields(G, prob)[f]) for f in FIELDS]

    ref = run("0,0")
    for v in ("1,1", "2,2", "3,3", "1,3", "4,4", "5,5", "6,6", "7,7", "8,8", "9,9", "10,10", "11,11", "4,9", "8,1",
              "12,12", "17,19", "22,23", "14,0", None):
        got = run(v)
        for a, b in zip(got, ref):
            assert np.array_equal(np.asarray(a), np.asarray(b)), v


@pytest.mark.parametrize("n", [1, 2, 63, 65, 257, 1000])
def test_tiny_and_ragged_problems(G, O, S, n):
    """Problems of 1 and 2 particles and counts that leave the last wave or
    block of every kernel part-empty (63, 65, 257, 1000): fluid on a jittered
    lattice (plus two ptype -1 particles from n = 63 on), 12 adaptive steps in
    several calls (eager single steps and captured multi-step calls), FP64
    against the oracle at the golden tolerances and FP32-split within 1e-5·dx
    of it."""
    dx = 0.02
    p, dist = S.standard_params(dx, 2)
    side = int(math.ceil(math.sqrt(n)))
    k = np.arange(n)
    x = 0.1 + (k % side) * dx + 0.1 * dx * (S.uniform01(7, n, 0) - 0.5)
    y = 0.1 + (k // side) * dx + 0.1 * dx * (S.uniform01(7, n, 1) - 0.5)
    ptype = np.full(n, 2, np.int32)
    if n >= 63:  # two wall particles 3·dx below the first row: in the lists, outside the wall force's reach
        ptype[:2] = -1
        y[:2] = 0.1 - 3 * dx
    rho = np.full(n, p.rho0) * (1 + 1e-3 * S.uniform01(7, n, 2))
    sc = S.Scene(2, (x, y), rho, ptype, p, dist, (dist, dist))
    cl, oc, p64, ref = _pair(G, O, sc)
    cl32 = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    p32 = G.SPHProblem.from_params(cl32, sc.params, sc.rho, sc.ptype, precision="f32split")
    for steps in (1, 1, 4, 6):
        G.stepsolve_(p64, steps, timestepping=True)
        G.stepsolve_(p32, steps, timestepping=True)
        ref.stepsolve(steps, True)
    got = gpu_fields(G, p64)
    for f in ("x", "v", "rho", "acc"):
        assert rel_err(got[f], ref.get(f)) <= TOL64[f], f
    assert np.abs(np.array(G.get_points(p32)) - ref.get("x")).max() < 1e-5 * dx
    assert np.abs(G.get_density(p32) / ref.get("rho") - 1).max() < 1e-6


def test_wall_force_fp32_and_fp64_branches(G, O, S):
    """FP32-split evaluates the wall force in FP32 while q = r0/(r + η²) < 32 and
    in FP64 for closer pairs.  One fluid particle 0.02·dx above a wall particle
    (q ≈ 96, FP64 branch) and one 0.9·dx above another (q ≈ 2.2, FP32 branch):
    FP32-split accelerations within 1e-5 of FP64 per particle, FP64 against the
    oracle.  Δt = 1e-20 keeps the half step from moving the particles."""
    dx = 0.02
    p, dist = S.standard_params(dx, 2)
    p.dt = 1e-20
    wx = np.arange(10) * dx
    x = np.concatenate([wx, [2 * dx, 6 * dx]])
    y = np.concatenate([np.zeros(10), [0.02 * dx, 0.9 * dx]])
    ptype = np.array([0] * 10 + [2, 2], dtype=np.int32)
    rho = np.full(12, p.rho0)
    sc = S.Scene(2, (x, y), rho, ptype, p, dist, (dist, dist))
    r0, eta2 = 2 * dx, 0.01 * (2 * dx) ** 2
    assert r0 / (0.02 * dx + eta2) > 32 > r0 / (0.9 * dx + eta2)
    cl, oc, p64, ref = _pair(G, O, sc)
    cl32 = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    p32 = G.SPHProblem.from_params(cl32, sc.params, sc.rho, sc.ptype, precision="f32split")
    G.stepsolve_(p64, 1)
    G.stepsolve_(p32, 1)
    ref.stepsolve(1)
    a64, a32 = np.array(G.get_acceleration(p64)), np.array(G.get_acceleration(p32))
    assert rel_err(a64, ref.get("acc")) < 1e-9
    for i in (10, 11):
        assert np.linalg.norm(a32[:, i] - a64[:, i]) < 1e-5 * np.linalg.norm(a64[:, i]), i
    assert np.linalg.norm(a64[:, 10]) > 1e6 * np.linalg.norm(a64[:, 11])  # the close pair dominates
