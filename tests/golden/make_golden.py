"""Write the golden fixtures under tests/golden/ from the reference's own
Fortran routines (oracle/_ref/ref_harness, built by ``make -C oracle ref``
from the pyOM2 sources).  Run in a container that has the reference:

    python tests/golden/make_golden.py

Each fixture stores the inputs it was made from (so tests do not depend on
the basin generator staying bit-stable) and the reference's outputs:
island boundary lists and directions (streamfunction_init), the psin
fields and line_psin matrix, the solution, iteration count and converged
flag of congrad_streamfunction, apply_op of that solution, and the island
line integrals of a seeded random (u, v) pair.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref_run  # noqa: E402
from pyom2_amd import basins  # noqa: E402


def tricky_basin():
    """Hand-drawn 36x28 map: a peninsula off the south wall, an island that
    touches another only at a corner (one 8-connected land mass), a one-cell
    island and a wide island with a notch."""
    nx, ny = 36, 28
    ocean = np.ones((ny, nx), bool)
    ocean[0, :] = ocean[-1, :] = False
    ocean[:, 0] = ocean[:, -1] = False
    ocean[1:9, 6] = False            # peninsula from the south wall
    ocean[6:9, 7:9] = False
    ocean[12:15, 12:15] = False      # block A
    ocean[15:17, 15:17] = False      # block B, corner contact with A
    ocean[20, 25] = False            # single cell island
    ocean[8:13, 24:31] = False       # wide island with a notch
    ocean[12, 27] = True
    ht = np.full((ny, nx), 3000.0)
    ht[:, 18:] = 1500.0
    grid = basins.make_grid(nx, ny, 25e3, 20e3)
    rng = np.random.default_rng(11)
    return basins._assemble("tricky", 11, nx, ny, grid, ht, ocean, rng, 16, 4, 1e-10)


CASES = {
    # the beta-plane workload itself: 128^2, eps 1e-12, 2000 iterations
    "c0_beta_plane_128": lambda: (basins.beta_plane_basin(seed=0), dict(eps=1e-12, max_iter=2000)),
    "islands_64": lambda: (basins.beta_plane_basin(seed=3, nx=64, ny=64, n_islands=2),
                           dict(eps=1e-12, max_iter=2000)),
    "spherical_96": lambda: (basins.spherical_basin(seed=2, n=96, n_islands=2),
                             dict(eps=1e-12, max_iter=3000)),
    "cyclic_40x32": lambda: (basins.beta_plane_basin(seed=5, nx=40, ny=32, n_islands=2,
                                                     cyclic_x=True),
                             dict(eps=1e-12, max_iter=2000)),
    "maxiter_48": lambda: (basins.beta_plane_basin(seed=7, nx=48, ny=40, n_islands=1),
                           dict(eps=1e-12, max_iter=7)),
    "warmstart_56": lambda: (basins.beta_plane_basin(seed=8, nx=56, ny=44, n_islands=1),
                             dict(eps=1e-12, max_iter=2000, warm=True)),
    "tricky_36x28": lambda: (tricky_basin(), dict(eps=1e-12, max_iter=2000)),
}


def make(name):
    b, kw = CASES[name]()
    rng = np.random.default_rng(1234)
    uloc = rng.normal(size=b.shape)
    vloc = rng.normal(size=b.shape)
    sol0 = np.zeros(b.shape)
    if kw.get("warm"):
        sol0 = rng.normal(size=b.shape) * 1e3
    r = ref_run.run(b, b.forc, sol0, kw["eps"], kw["max_iter"], uloc, vloc)
    bnd = r["boundary"]
    offs = np.concatenate([[0], np.cumsum([len(x[0]) for x in bnd])]).astype(np.int32)
    cat = lambda k: np.concatenate([x[k] for x in bnd]).astype(np.int32)
    out = dict(
        nx=b.nx, ny=b.ny, cyclic_x=int(b.cyclic_x), eps=kw["eps"], max_iter=kw["max_iter"],
        dxt=b.dxt, dxu=b.dxu, dyt=b.dyt, dyu=b.dyu, cost=b.cost, cosu=b.cosu,
        kbot=b.kbot, hur=b.hur, hvr=b.hvr, maskU=b.maskU, maskV=b.maskV, forc=b.forc,
        sol0=sol0, uloc=uloc, vloc=vloc,
        nisle=r["nisle"], offsets=offs, bi=cat(0), bj=cat(1), di=cat(2), dj=cat(3),
        line_psin=r["line_psin"], psin=r["psin"], sol=r["sol"], res=r["res"],
        iterations=r["iterations"], converged=int(r["converged"]), lines=r["lines"],
    )
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: nisle={r['nisle']} iterations={r['iterations']} converged={r['converged']} "
          f"-> {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    if not ref_run.build():
        sys.exit("reference sources not available: cannot build oracle/_ref")
    for n in (sys.argv[1:] or CASES):
        make(n)
