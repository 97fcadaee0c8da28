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

The ``press_*`` fixtures come from oracle/_ref/ref_press, the reference's
solve_pressure (surface-pressure mode, free surface and rigid lid).
The ``nh_*`` fixtures come from oracle/_ref/ref_nonhydro, the reference's
solve_non_hydrostatic (and its make_coeff_non_hydro coefficients).
The ``step_*`` fixtures come from oracle/_ref/ref_step, the reference's
whole solve_streamfunction on a basin with vertical levels and no islands:
inputs (levels, masks, the model state at tau/taum1) and outputs (p_hydro,
du/dv after the pressure gradient, u/v/psi/dpsi at taup1, iterations).
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


def cyclic_bridge_basin():
    """Zonally periodic 28x18 channel whose two walls are joined by a
    meridional ridge, so the land is one mass (no island system)."""
    nx, ny = 28, 18
    ocean = np.ones((ny, nx), bool)
    ocean[0, :] = ocean[-1, :] = False
    ocean[:, 9] = False
    ht = np.full((ny, nx), 3500.0)
    ht[:, 15:22] = 1200.0
    grid = basins.make_grid(nx, ny, 30e3, 25e3, cyclic_x=True)
    rng = np.random.default_rng(21)
    return basins._assemble("cyclic_bridge", 21, nx, ny, grid, ht, ocean, rng, 8, 3, 1e-10,
                            cyclic_x=True)


STEP_CASES = {
    "step_24x20_nz5": lambda: (basins.beta_plane_basin(seed=1, nx=24, ny=20), 5, 0),
    "step_cyclic_28x18_nz4": lambda: (cyclic_bridge_basin(), 4, 1),
}


def make_step(name):
    b, nz, seed = STEP_CASES[name]()
    b3, col = basins.column_basin(b, nz)
    mom = basins.momentum_fields(col, seed)
    eps, max_iter, dt = 1e-12, 2000, 3600.0
    r = ref_run.run_step(b3, col, mom, eps, max_iter, dt)
    out = dict(
        nx=b3.nx, ny=b3.ny, nz=nz, cyclic_x=int(b3.cyclic_x), eps=eps, max_iter=max_iter,
        dt_mom=dt, dxt=b3.dxt, dxu=b3.dxu, dyt=b3.dyt, dyu=b3.dyu, cost=b3.cost, cosu=b3.cosu,
        kbot=b3.kbot, hur=b3.hur, hvr=b3.hvr, dzt=col.dzt, dzw=col.dzw, kbot_levels=col.kbot,
        maskT=col.maskT, maskU=col.maskU, maskV=col.maskV,
        rho=mom.rho[1], du_tau=mom.du[1], du_taum1=mom.du[0], dv_tau=mom.dv[1],
        dv_taum1=mom.dv[0], du_mix=mom.du_mix, dv_mix=mom.dv_mix, u_tau=mom.u[1],
        v_tau=mom.v[1], psi_tau=mom.psi[1], dpsi_tau=mom.dpsi[1], dpsi_taum1=mom.dpsi[0],
        out_p_hydro=r["p_hydro"], out_du=r["du"], out_dv=r["dv"], out_u=r["u"], out_v=r["v"],
        out_psi=r["psi"], out_dpsi=r["dpsi"], iterations=r["iterations"],
    )
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: iterations={r['iterations']} -> {os.path.getsize(path) / 1024:.0f} KiB")


PRESS_CASES = {
    # free surface, a closed basin with an island, 5 levels
    "press_fs_24x20_nz5": lambda: (basins.beta_plane_basin(seed=1, nx=24, ny=20, n_islands=1),
                                   5, 2, 1),
    # rigid lid (a singular, consistent system), closed basin with an island
    "press_rigid_26x22_nz3": lambda: (basins.beta_plane_basin(seed=2, nx=26, ny=22, n_islands=1),
                                      3, 4, 0),
    # rigid lid, zonally periodic channel: the reference ends on its divergence
    # test (the residual norm grows 100-fold over its minimum)
    "press_rigid_cyclic_30x16_nz4": lambda: (basins.beta_plane_basin(seed=6, nx=30, ny=16,
                                                                     cyclic_x=True), 4, 3, 0),
}


NH_CASES = {
    # (seed, nx, ny, nz, cyclic, dx, depth, stretch, itt, dt)
    "nh_box_20x16_nz10": (1, 20, 16, 10, False, 50.0, 1000.0, 1.0, 1, 60.0),
    "nh_cyclic_22x14_nz8": (2, 22, 14, 8, True, 100.0, 800.0, 1.0, 0, 60.0),
    "nh_stretched_18x16_nz10": (1, 18, 16, 10, False, 100.0, 1000.0, 1.05, 1, 60.0),
}


def make_nonhydro(name):
    seed, nx, ny, nz, cyc, dx, depth, stretch, itt, dt = NH_CASES[name]
    b, col = basins.nonhydro_box(seed, nx, ny, nz, cyc, dx, depth, stretch)
    mW = basins.mask_w(col.maskT)
    nh = basins.nonhydro_fields(col, seed, b.nx if cyc else None)
    eps, max_iter = 1e-12, 5000
    r = ref_run.run_nonhydro(b, col, mW, nh, eps, max_iter, dt, itt)
    out = dict(
        nx=b.nx, ny=b.ny, nz=nz, cyclic_x=int(b.cyclic_x), eps=eps, max_iter=max_iter, dt_mom=dt,
        itt=itt, dxt=b.dxt, dxu=b.dxu, dyt=b.dyt, dyu=b.dyu, cost=b.cost, cosu=b.cosu,
        dzt=col.dzt, dzw=col.dzw, kbot_levels=col.kbot, maskT=col.maskT, maskU=col.maskU,
        maskV=col.maskV, maskW=mW, u_taup1=nh.u[2], v_taup1=nh.v[2], w_tau=nh.w[1],
        w_taup1=nh.w[2], dw_tau=nh.dw[1], dw_taum1=nh.dw[0], dw_mix=nh.dw_mix, p_tau=nh.p[1],
        p_taum1=nh.p[0], out_u=r["u"], out_v=r["v"], out_w=r["w"], out_p=r["p"],
        out_p_tau=r["p_tau"], out_p_taum1=r["p_taum1"], iterations=r["iterations"],
        ref_warning=int("WARNING" in r["log"]),
    )
    # the 7 non-zero planes of make_coeff_non_hydro's cf (below, S, W, C, E, N, above)
    out["cf7"] = r["cf"][[4, 10, 12, 13, 14, 16, 22]]
    out["cf_other_max"] = np.abs(np.delete(r["cf"], [4, 10, 12, 13, 14, 16, 22], axis=0)).max()
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: iterations={r['iterations']} warning={out['ref_warning']} -> "
          f"{os.path.getsize(path) / 1024:.0f} KiB")


def make_press(name):
    b0, nz, seed, fs = PRESS_CASES[name]()
    b, col = basins.column_basin(b0, nz)
    mom = basins.momentum_fields(col, seed)
    mom.psi[:] = mom.psi * 1e-6          # surface pressure, m^2/s^2
    eps, max_iter, dt = 1e-12, 3000, 1200.0
    r = ref_run.run_press(b, col, b.hu, b.hv, mom, fs, eps, max_iter, dt)
    out = dict(
        nx=b.nx, ny=b.ny, nz=nz, cyclic_x=int(b.cyclic_x), free_surface=fs, eps=eps,
        max_iter=max_iter, dt_mom=dt, dxt=b.dxt, dxu=b.dxu, dyt=b.dyt, dyu=b.dyu, cost=b.cost,
        cosu=b.cosu, hu=b.hu, hv=b.hv, dzt=col.dzt, dzw=col.dzw, kbot_levels=col.kbot,
        maskT=col.maskT, maskU=col.maskU, maskV=col.maskV,
        rho=mom.rho[1], du_tau=mom.du[1], du_taum1=mom.du[0], dv_tau=mom.dv[1],
        dv_taum1=mom.dv[0], du_mix=mom.du_mix, dv_mix=mom.dv_mix, u_tau=mom.u[1],
        v_tau=mom.v[1], psi_tau=mom.psi[1], psi_taum1=mom.psi[0],
        out_p_hydro=r["p_hydro"], out_du=r["du"], out_dv=r["dv"], out_u=r["u"], out_v=r["v"],
        out_psi=r["psi"], iterations=r["iterations"],
        ref_warning=int("WARNING" in r["log"]),
    )
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **out)
    print(f"{name}: iterations={r['iterations']} warning={out['ref_warning']} -> "
          f"{os.path.getsize(path) / 1024:.0f} KiB")


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
    for n in (sys.argv[1:] or list(CASES) + list(STEP_CASES) + list(PRESS_CASES) +
              list(NH_CASES)):
        (make_step if n in STEP_CASES else make_press if n in PRESS_CASES else
         make_nonhydro if n in NH_CASES else make)(n)
