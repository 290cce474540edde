"""Writes the oracle_*.npz fixtures: CPU-oracle outputs on small seeded scenes.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "gpucelllistsph.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
sys.path.insert(0, os.path.join(ROOT, "tests"))
from cases import golden_cases  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def run_oracle(sc, steps, ts):
    cl = O.CellList(sc.points, sc.cellsize, sc.dist)
    pairs0 = np.unique(cl.neighborlistview(), axis=0)
    sph = O.SPHProblem(cl, sc.params, sc.rho, sc.ptype, v=sc.v)
    upd, mx = sph.stepsolve(steps, ts)
    out = {k: sph.get(k) for k in ("x", "v", "rho", "P", "acc", "drho", "sumW", "sumgradW")}
    out.update(pairs0=pairs0, pairs_end=np.unique(cl.neighborlistview(), axis=0), dt=sph.dt, etime=sph.etime,
               updaten=upd, maxcdx=mx, steps=steps)
    return out


def main():
    for name, (sc, steps, ts) in golden_cases().items():
        out = run_oracle(sc, steps, ts)
        np.savez_compressed(os.path.join(HERE, f"oracle_{name}.npz"), **out)
        print(name, sc.n, "particles", len(out["pairs0"]), "pairs", "updaten", out["updaten"])


if __name__ == "__main__":
    main()
