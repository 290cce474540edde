"""gcl_neighborlist_noalloc (no counterpart in the reference: for callers
that cannot allocate device memory) against the allocating entry points it
is a variant of: GPUCellList(points, cellsize, dist; MIN, MAX) followed by
neighborlistview (gcl_create + gcl_get_pairs).  Same pairs in the same order
and the same count, bit for bit, on two calls in a row that share the
library's fixed scratch array (the second with fewer points and cells), in
2D and 3D, with a pair buffer too small for the list, and the size limits."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    """The test's own device buffers, through the HIP runtime libgclsph is
    linked against (found among the loaded libraries): the caller allocates
    the inputs and outputs, the entry point allocates nothing.  (A second HIP
    runtime in the process, such as the one bundled with PyTorch, cannot
    share the device with it.)"""

    def __init__(self, G):
        G.lib()
        path = next(l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64.so" in l)
        self.L = C.CDLL(path)
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def put(self, a: np.ndarray) -> int:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        self.bufs.append(p)
        return int(p.value)

    def get(self, ptr: int, like: np.ndarray) -> np.ndarray:
        out = np.empty_like(like)
        assert self.L.hipDeviceSynchronize() == 0
        assert self.L.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _reference_pairs(G, pts, cs, dist):
    mn = [float(p.min()) for p in pts]
    mx = [float(p.max()) for p in pts]
    cl = G.GPUCellList(pts, cs, dist, MIN=mn, MAX=mx)
    return G.neighborlistview(cl), mn, mx


def _noalloc(G, hip, pts, cs, dist, mn, mx, capacity):
    xs = [hip.put(np.asarray(p, dtype=np.float64)) for p in pts]
    sentinel = np.full(2 * capacity + 2, -7, dtype=np.int32)
    out = hip.put(sentinel)
    n = len(pts[0])
    print(f"noalloc call: dim={len(pts)} n={n} capacity={capacity}", flush=True)
    pairsn = G.neighborlist_noalloc(xs, n, cs, dist, mn, mx, out, capacity)
    buf = hip.get(out, sentinel)
    hip.free()
    return pairsn, buf


def test_noalloc_neighborlist_matches_gpucelllist(G, S):
    hip = Hip(G)
    cases = [
        (S.config1_points(), S.CONFIG1_DIST, S.CONFIG1_DIST),  # the reference's README system, 200,000 points
        (S.uniform_points(21, 30_000, 2), 0.03, 0.025),  # fewer points and cells: second call on the same scratch
        (S.uniform_points(22, 40_000, 3), 0.08, 0.08),
        (S.uniform_points(23, 5_000, 3), 0.15, 0.12),
    ]
    for pts, cs, dist in cases:
        ref, mn, mx = _reference_pairs(G, pts, cs, dist)
        pairsn, buf = _noalloc(G, hip, pts, cs, dist, mn, mx, len(ref) + 10)
        assert pairsn == len(ref) > 0
        got = buf[: 2 * pairsn].reshape(-1, 2)
        assert np.array_equal(got, ref)
        assert np.all(buf[2 * pairsn:] == -7)  # entries after the list are left as they were
    # a buffer too small for the list keeps its first `capacity` pairs, and the
    # count is the true one (as an overflowing GPUCellList pair buffer)
    pts, cs, dist = cases[1]
    ref, mn, mx = _reference_pairs(G, pts, cs, dist)
    pairsn, buf = _noalloc(G, hip, pts, cs, dist, mn, mx, 1000)
    assert pairsn == len(ref)
    assert np.array_equal(buf[:2000].reshape(-1, 2), ref[:1000])


def test_noalloc_limits(G, S):
    hip = Hip(G)
    pts = S.uniform_points(24, 1000, 2)
    xs = [hip.put(p) for p in pts]
    out = hip.put(np.zeros(16, dtype=np.int32))
    with pytest.raises(ValueError, match="NOALLOC_MAX_POINTS"):
        G.neighborlist_noalloc(xs, G.api.NOALLOC_MAX_POINTS + 1, 0.1, 0.1, (0, 0), (1, 1), out, 8)
    with pytest.raises(ValueError, match="NOALLOC_MAX_CELLS"):
        G.neighborlist_noalloc(xs, 1000, 1e-4, 1e-4, (0, 0), (1, 1), out, 8)
    hip.free()
