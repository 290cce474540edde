"""Seeded synthetic scenes for the cell-list / SPH workloads.

Geometry follows the reference's DualSPHysics-style fixtures
(test/input/BoundaryPoints_Dp0.02.csv, FluidPoints_Dp0.02.csv): a tank whose
walls and floor are `layers` rows of particles at spacing dx lying *inside* the
tank bounds (x in {0, dx, 2dx} and {L-2dx, L-dx, L}; floor y in {0, dx, 2dx}),
and a fluid lattice starting `layers`*dx from the corner.  With dx = 0.02 and a
4 m x 3 m tank this reproduces the fixture's 1,491 boundary particles exactly.

Coordinate 2 (index 1) is vertical in 2D and 3D, as in the reference
(gravity vector (0, g[, 0]), src/sphproblem.jl:10-17).

Random numbers come from a project-written SplitMix64 so that every language
in this repository can regenerate the same inputs bit for bit.
"""
from __future__ import annotations

import csv
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# ---------------------------------------------------------------------------
# RNG: SplitMix64 (Steele, Lea, Flood 2014), vectorised over a counter.
# ---------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n uint64 outputs of SplitMix64 started at `seed` (+ stream offset)."""
    with np.errstate(over="ignore"):
        gamma = np.uint64(0x9E3779B97F4A7C15)
        base = np.uint64((seed + stream * 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF)
        z = base + gamma * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """Doubles in [0, 1) from the top 53 bits."""
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """Standard normals by Box-Muller over two uniform streams."""
    u1 = uniform01(seed, n, 2 * stream + 1)
    u2 = uniform01(seed, n, 2 * stream + 2)
    u1 = np.where(u1 <= 0.0, 1.0 / 9007199254740992.0, u1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


# ---------------------------------------------------------------------------
# Physical parameter set (examples/timesolve.jl:11-24, README.md:103-120)
# ---------------------------------------------------------------------------
@dataclass
class SPHParams:
    dx: float
    h: float
    H: float
    rho0: float = 1000.0
    m0: float = 0.0
    dt: float = 1e-5
    alpha: float = 0.01
    c0: float = 0.0
    gamma: int = 7
    delta: float = 0.1
    CFL: float = 0.2
    g: float = 9.81
    s: float = 0.0
    nu: float = 0.0
    dpc_l0: float = 0.0
    dpc_pmin: float = 1.0
    dpc_pmax: float = 10000.0
    dpc_lambda: float = 0.01
    xsph_eps: float = 0.0
    cspmn: int = 15
    bound_D: float = 0.4
    bound_l: float = 0.0
    nui: float = 0.0
    etime: float = 0.0

    ORDER = ("dx", "h", "H", "rho0", "m0", "dt", "alpha", "c0", "gamma", "delta", "CFL", "g", "s", "nu",
             "dpc_l0", "dpc_pmin", "dpc_pmax", "dpc_lambda", "xsph_eps", "cspmn", "bound_D", "bound_l",
             "nui", "etime")

    def as_array(self) -> np.ndarray:
        return np.array([float(getattr(self, k)) for k in self.ORDER], dtype=np.float64)


def standard_params(dx: float, dim: int = 2, dt: float = 1e-5, dist_factor: float = 1.0) -> Tuple[SPHParams, float]:
    """Reference parameter set for lattice spacing dx; returns (params, dist)."""
    h = 1.2 * math.sqrt(2) * dx
    H = 2 * h
    rho0 = 1000.0
    m0 = rho0 * dx ** dim
    g = 9.81
    c0 = math.sqrt(g * 2) * 20
    dist = H * dist_factor
    p = SPHParams(dx=dx, h=h, H=H, rho0=rho0, m0=m0, dt=dt, c0=c0, g=g, bound_l=2 * dx, nui=dist - H)
    return p, dist


def hydrostatic_rho(y: np.ndarray, surface: float, rho0: float, g: float, c0: float, gamma: int = 7) -> np.ndarray:
    """rho0 (1 + rho0 g (surface - y) / B)^(1/gamma), B = c0^2 rho0 / gamma."""
    B = c0 * c0 * rho0 / gamma
    return rho0 * np.power(1.0 + rho0 * g * (surface - y) / B, 1.0 / gamma)


# ---------------------------------------------------------------------------
# Scenes
# ---------------------------------------------------------------------------
@dataclass
class Scene:
    dim: int
    points: Tuple[np.ndarray, ...]
    rho: np.ndarray
    ptype: np.ndarray
    params: SPHParams
    dist: float
    cellsize: Tuple[float, ...]
    v: Optional[Tuple[np.ndarray, ...]] = None
    name: str = ""
    info: Dict[str, float] = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.points[0].shape[0])


def _lattice(n: int, start: float, dx: float) -> np.ndarray:
    return start + dx * np.arange(n, dtype=np.float64)


def tank_boundary_2d(dx: float, width: float, height: float, layers: int = 3) -> Tuple[np.ndarray, np.ndarray]:
    nx = int(round(width / dx)) + 1
    ny = int(round(height / dx)) + 1
    xs, ys = [], []
    xrow = _lattice(nx, 0.0, dx)
    for k in range(layers):  # floor rows
        xs.append(xrow)
        ys.append(np.full(nx, k * dx))
    ycol = _lattice(ny - layers, layers * dx, dx)
    for k in range(layers):  # both walls above the floor rows
        xs.append(np.full(ycol.shape, k * dx))
        ys.append(ycol)
        xs.append(np.full(ycol.shape, width - k * dx))
        ys.append(ycol)
    return np.concatenate(xs), np.concatenate(ys)


def fixture_wall_types(x: np.ndarray, y: np.ndarray, dx: float, width: float, layers: int = 3) -> np.ndarray:
    """Wall typing of the reference's fixtures (test/input/BoundaryPoints_Dp0.02.csv):
    the layer next to the fluid is ptype 1 (takes part in density, pressure and
    viscosity), the outer layers are ptype 0 (only the repulsive wall force)."""
    inner = (np.isclose(y, (layers - 1) * dx) & (x >= (layers - 1) * dx - 1e-9) & (x <= width - (layers - 1) * dx + 1e-9))
    inner |= (np.isclose(x, (layers - 1) * dx) | np.isclose(x, width - (layers - 1) * dx)) & (y >= (layers - 1) * dx - 1e-9)
    return np.where(inner, 1, 0).astype(np.int32)


def _assert_distinct(points: Sequence[np.ndarray], dx: float):
    key = np.stack([np.round(p / dx * 4).astype(np.int64) for p in points], axis=1)
    if len(np.unique(key, axis=0)) != len(key):
        raise ValueError("scene has coincident particles")


def dam_break_2d(dx: float, fluid_w: float, fluid_h: float, tank_w: float, tank_h: float, layers: int = 3,
                 boundary_ptype="fixture", fluid_ptype: int = 2, dt: float = 1e-5, dist_factor: float = 1.0,
                 name: str = "") -> Scene:
    """boundary_ptype: an int for every wall particle, or "fixture" for the
    reference fixtures' typing (inner layer 1, outer layers 0)."""
    p, dist = standard_params(dx, 2, dt, dist_factor)
    nfx = int(round(fluid_w / dx))
    nfy = int(round(fluid_h / dx))
    fx = _lattice(nfx, layers * dx, dx)
    fy = _lattice(nfy, layers * dx, dx)
    FX, FY = np.meshgrid(fx, fy, indexing="ij")
    FX, FY = FX.ravel(), FY.ravel()
    bx, by = tank_boundary_2d(dx, tank_w, tank_h, layers)
    x = np.concatenate([FX, bx])
    y = np.concatenate([FY, by])
    nf = FX.shape[0]
    if isinstance(boundary_ptype, str):
        wt = fixture_wall_types(bx, by, dx, tank_w, layers)
    else:
        wt = np.full(bx.shape[0], boundary_ptype, np.int32)
    ptype = np.concatenate([np.full(nf, fluid_ptype, np.int32), wt])
    _assert_distinct((x, y), dx)
    surface = float(FY.max())
    rho = np.concatenate([hydrostatic_rho(FY, surface, p.rho0, p.g, p.c0, p.gamma), np.full(bx.shape[0], p.rho0)])
    return Scene(2, (x, y), rho, ptype, p, dist, (p.H * dist_factor, p.H * dist_factor), None, name,
                 {"fluid": float(nf), "boundary": float(bx.shape[0])})


def dam_break_3d(dx: float, fluid: Sequence[float], tank: Sequence[float], layers: int = 3,
                 boundary_ptype: int = -1, fluid_ptype: int = 2, dt: float = 1e-5, name: str = "") -> Scene:
    """fluid = (length x, height y, depth z); tank = (length, height, depth); y vertical.
    The walls lie inside the tank bounds (as in 2D), so a tank dimension smaller
    than the fluid block plus both wall stacks is widened to fit it."""
    p, dist = standard_params(dx, 3, dt)
    nfx, nfy, nfz = (int(round(a / dx)) for a in fluid)
    tank = list(tank)
    tank[0] = max(tank[0], (nfx - 1 + 2 * layers) * dx)
    tank[2] = max(tank[2], (nfz - 1 + 2 * layers) * dx)
    fx = _lattice(nfx, layers * dx, dx)
    fy = _lattice(nfy, layers * dx, dx)
    fz = _lattice(nfz, layers * dx, dx)
    FX, FY, FZ = (a.ravel() for a in np.meshgrid(fx, fy, fz, indexing="ij"))
    L, Ht, Dp = tank
    nx = int(round(L / dx)) + 1
    ny = int(round(Ht / dx)) + 1
    nz = int(round(Dp / dx)) + 1
    gx, gy, gz = (a.ravel() for a in np.meshgrid(_lattice(nx, 0.0, dx), _lattice(ny, 0.0, dx),
                                                 _lattice(nz, 0.0, dx), indexing="ij"))
    lay = layers * dx - 0.5 * dx
    shell = (gy < lay) | (gx < lay) | (gx > L - lay) | (gz < lay) | (gz > Dp - lay)
    bx, by, bz = gx[shell], gy[shell], gz[shell]
    nf = FX.shape[0]
    pts = (np.concatenate([FX, bx]), np.concatenate([FY, by]), np.concatenate([FZ, bz]))
    _assert_distinct(pts, dx)
    ptype = np.concatenate([np.full(nf, fluid_ptype, np.int32), np.full(bx.shape[0], boundary_ptype, np.int32)])
    surface = float(FY.max())
    rho = np.concatenate([hydrostatic_rho(FY, surface, p.rho0, p.g, p.c0, p.gamma), np.full(bx.shape[0], p.rho0)])
    return Scene(3, pts, rho, ptype, p, dist, (p.H, p.H, p.H), None, name,
                 {"fluid": float(nf), "boundary": float(bx.shape[0])})


def uniform_points(seed: int, n: int, dim: int = 2) -> Tuple[np.ndarray, ...]:
    return tuple(uniform01(seed, n, k) for k in range(dim))


def jitter(points: Tuple[np.ndarray, ...], seed: int, step: int, sigma: float = 0.001) -> Tuple[np.ndarray, ...]:
    d = len(points)
    return tuple(p + sigma * normal(seed, p.shape[0], stream=1000 + step * d + k) for k, p in enumerate(points))


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------
def config0() -> Scene:
    """2D dam break: 1.0 x 2.0 m fluid column at dx=0.02 (50x100) in a 4 x 3 m tank, ptype -1 walls."""
    return dam_break_2d(0.02, 1.0, 2.0, 4.0, 3.0, boundary_ptype=-1, name="dam-break-2d dx=0.02 (5,000 fluid)")


def config1_points(seed: int = 12345, n: int = 200_000) -> Tuple[np.ndarray, np.ndarray]:
    """200,000 points uniform in [0,1)^2; cellsize = dist = 0.016."""
    return uniform_points(seed, n, 2)


CONFIG1_DIST = 0.016


def config2(dx: float = 0.01) -> Scene:
    """3D dam break: 1.0 x 1.0 x 0.5 m fluid block (x, y-up, z) in a 3.2 x 1.5 x 0.5 m tank
    (widened in z to hold the 0.5 m block between its wall stacks), ptype -1 walls."""
    sc = dam_break_3d(dx, (1.0, 1.0, 0.5), (3.2, 1.5, 0.5), name=f"dam-break-3d dx={dx}")
    sc.params.nu = 0.2
    sc.params.s = 0.05
    sc.params.dpc_l0 = 0.005
    sc.params.dpc_lambda = 0.005
    sc.params.dpc_pmax = 36000.0
    sc.params.xsph_eps = 0.5
    return sc


def config3(dx: float = 0.002, fluid: float = 4.0, walls="fixture") -> Scene:
    """Large 2D dam break: 4 x 4 m column at dx=0.002 (4,000,000 fluid) in a 16 x 8 m tank.
    walls: -1 (every wall particle ptype -1, as config 0's walls) or "fixture" (typed as in the
    reference's fixtures: inner layer 1, outer layers 0)."""
    return dam_break_2d(dx, fluid, fluid, 16.0, 8.0, boundary_ptype=walls,
                        name=f"dam-break-2d dx={dx} ({fluid}x{fluid} m column, walls {walls})")


def config4(seed: int = 4, count: int = 64) -> list:
    """Ensemble of 2D dam breaks drawn from a seed: dx in {0.005, 0.01, 0.02}, a column 0.5-1.5 m wide and
    1.0-2.0 m high in a 4.0 x 3.0 m tank with 3 wall layers (1,500-120,000 particles per scene)."""
    u = [uniform01(seed, count, k) for k in range(3)]
    out = []
    for i in range(count):
        dx = (0.005, 0.01, 0.02)[min(int(u[0][i] * 3), 2)]
        w = 0.5 + float(u[1][i])
        h = 1.0 + float(u[2][i])
        out.append(dam_break_2d(dx, w, h, 4.0, 3.0, name=f"ensemble-{i} dx={dx} {w:.2f}x{h:.2f} m"))
    return out


def load_dualsphysics_csv(fluid_csv: str, boundary_csv: str, cols: Sequence[str] = ("Points:0", "Points:2")):
    """Reads the reference's CSV fixtures (src/auxillary.jl:3-11): points, Rhop, ptype."""
    pts: List[List[float]] = [[] for _ in cols]
    rho: List[float] = []
    ptype: List[int] = []
    for path in (fluid_csv, boundary_csv):
        with open(path, newline="") as fh:
            for row in csv.DictReader(fh):
                for k, c in enumerate(cols):
                    pts[k].append(float(row[c]))
                rho.append(float(row["Rhop"]))
                ptype.append(int(float(row.get("ptype", 0) or 0)))
    return tuple(np.array(a, dtype=np.float64) for a in pts), np.array(rho), np.array(ptype, dtype=np.int32)
