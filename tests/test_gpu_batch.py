"""Batched ensemble: problems stepped together (one launch per stage for all of
them) end bit-identical to the same problems stepped one by one, each keeping
its own Δt, rebuild schedule, time and corrections."""
import numpy as np
import pytest
from cases import golden_cases

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "rho", "P", "acc", "drho", "sumW", "sumgradW")


def fields(G, prob):
    return {"x": np.array(G.get_points(prob)), "v": np.array(G.get_velocity(prob)), "rho": G.get_density(prob),
            "P": G.get_pressure(prob), "acc": np.array(G.get_acceleration(prob)), "drho": G.get_drho(prob),
            "sumW": G.get_sumw(prob), "sumgradW": np.array(G.get_sumgradw(prob)), "dt": prob.dt,
            "etime": prob.etime}


def make(G, sc, precision, v=None):
    cl = G.GPUCellList(sc.points, sc.cellsize, sc.dist)
    return G.SPHProblem.from_params(cl, sc.params, sc.rho, sc.ptype, v=v if v is not None else sc.v,
                                    precision=precision)


def scenes_2d(S):
    out = [golden_cases()["dam2d"][0], golden_cases()["dam2d_allphys"][0], S.config0()]
    sc = S.dam_break_2d(0.03, 0.6, 0.9, 2.0, 1.5, boundary_ptype=-1, dist_factor=1.2, name="uneven")
    sc.params.cspmn = 3
    out.append(sc)
    return out


@pytest.mark.parametrize("prec", ["f64", "f32split"])
def test_batch_2d_matches_one_by_one(G, S, prec):
    scs = scenes_2d(S)
    steps = (7, 5)
    alone = []
    for sc in scs:
        p = make(G, sc, prec)
        res = [G.stepsolve_(p, n, timestepping=True) for n in steps]
        alone.append((res, fields(G, p)))
    probs = [make(G, sc, prec) for sc in scs]
    batch = G.SPHBatch(probs)
    got = [G.batch_stepsolve_(batch, n, timestepping=True) for n in steps]
    for i, p in enumerate(probs):
        res, ref = alone[i]
        for k, (upd, mx) in enumerate(got):
            assert (int(upd[i]), float(mx[i])) == res[k]
        f = fields(G, p)
        for name in ref:
            assert np.array_equal(f[name], ref[name]), (i, name)


def test_batch_3d_and_single_steps_interleaved(G, S):
    """3D problems in a batch; a problem stepped alone between batch calls keeps
    the same trajectory as one never batched."""
    scs = [golden_cases()["dam3d_allphys"][0], S.config2(dx=0.05)]
    ref = []
    for sc in scs:
        p = make(G, sc, "f64")
        G.stepsolve_(p, 2, timestepping=True)
        G.stepsolve_(p, 1, timestepping=True)
        G.stepsolve_(p, 2, timestepping=True)
        ref.append(fields(G, p))
    probs = [make(G, sc, "f64") for sc in scs]
    batch = G.SPHBatch(probs)
    G.batch_stepsolve_(batch, 2, timestepping=True)
    G.stepsolve_(probs[0], 1, timestepping=True)
    G.stepsolve_(probs[1], 1, timestepping=True)
    G.batch_stepsolve_(batch, 2, timestepping=True)
    for p, r in zip(probs, ref):
        f = fields(G, p)
        for name in r:
            assert np.array_equal(f[name], r[name]), name


def test_batch_argument_errors(G, S):
    sc = golden_cases()["dam2d"][0]
    a, b = make(G, sc, "f64"), make(G, sc, "f32split")
    with pytest.raises(ValueError, match="share dimension and precision"):
        G.SPHBatch([a, b])
    with pytest.raises(ValueError, match="twice"):
        G.SPHBatch([a, a])
    with pytest.raises(ValueError, match="at least one"):
        G.SPHBatch([])
    batch = G.SPHBatch([a])
    assert G.lib().sph_destroy(a._h) != 0  # held by the batch
    with pytest.raises(ValueError, match="one timelims pair"):
        G.batch_stepsolve_(batch, 1, timelims=[(1e-8, 1e-5), (1e-8, 1e-5)])
    del batch


def _oracle_pair(O, sc):
    oc = O.CellList(sc.points, sc.cellsize, sc.dist)
    return O.SPHProblem(oc, sc.params, sc.rho, sc.ptype, v=sc.v)


def test_config4_ensemble_64_scenes(G, O, S):
    """The 64-scene seeded ensemble (1.5 k-120 k particles per scene) as one
    FP32-split batch, 12 adaptive steps in two calls: every scene ends
    bit-identical to the same scene stepped alone (tile maps, descriptor array
    and captured step of the full-size batch).  An FP64 batch of the same
    scenes matches the CPU oracle on the smallest and the largest scene, and
    the FP32-split batch stays within the FP32-split bounds of it."""
    scs = S.config4()
    assert len(scs) == 64
    sizes = [sc.n for sc in scs]
    lo, hi = int(np.argmin(sizes)), int(np.argmax(sizes))
    assert sizes[lo] < 5000 and sizes[hi] > 100000
    steps = (10, 2)
    probs = [make(G, sc, "f32split") for sc in scs]
    batch = G.SPHBatch(probs)
    got = [G.batch_stepsolve_(batch, n, timestepping=True) for n in steps]
    for i, sc in enumerate(scs):
        p = make(G, sc, "f32split")
        res = [G.stepsolve_(p, n, timestepping=True) for n in steps]
        for k, (upd, mx) in enumerate(got):
            assert (int(upd[i]), float(mx[i])) == res[k], i
        ref, f = fields(G, p), fields(G, probs[i])
        for name in ref:
            assert np.array_equal(f[name], ref[name]), (i, name)
        del p
    p64 = [make(G, sc, "f64") for sc in scs]
    b64 = G.SPHBatch(p64)
    for n in steps:
        G.batch_stepsolve_(b64, n, timestepping=True)
    for i in (lo, hi):
        sc = scs[i]
        ref = _oracle_pair(O, sc)
        for n in steps:
            ref.stepsolve(n, True)
        dx = sc.params.dx
        assert p64[i].etime == pytest.approx(ref.etime, rel=1e-12)
        x64 = np.array(G.get_points(p64[i]))
        assert np.abs(x64 - ref.get("x")).max() < 1e-12 * dx, i
        assert np.abs(G.get_density(p64[i]) / ref.get("rho") - 1).max() < 1e-12, i
        x32 = np.array(G.get_points(probs[i]))
        assert np.abs(x32 - x64).max() < 1e-5 * dx, i
        assert np.abs(G.get_density(probs[i]) / G.get_density(p64[i]) - 1).max() < 1e-6, i
