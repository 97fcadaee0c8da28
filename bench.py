"""Benchmark: ms per 2048^2 streamfunction solve to a 1e-10 relative FP64
residual on one MI355X.

Workload: the 2048^2 spherical 0.05 deg basin with two seeded islands
(pyom2_amd.basins.spherical_basin, seed 0), a sequence of slowly drifting
forcings (AR(1), rho = 0.99) solved one after another with the reference's
warm start dpsi(taup1) = 2 dpsi(tau) - dpsi(taum1).  A step is one solve.
Inputs are resident in HBM before the timed region starts.

Paths: "mg" (the headline: the fastest) is FP64 flexible CG preconditioned
by an FP32 aggregation-multigrid V-cycle; "learned" is flexible CG (FP32
inner, FP64 residual replacement) preconditioned by the trained network
composed with one V-cycle per step; "poly" is FP32-inner / FP64-refined PCG
with a Chebyshev polynomial of the Jacobi-preconditioned operator (degree
4); "mixed" is FP32-inner / FP64-refined Jacobi-PCG; "plain" is FP64
Jacobi-PCG.

Prints one JSON line with the headline path's ms per solve, the other paths
beside it (the learned path's own figures under "learned"), every path's
solution against one oracle solve of the last timed forcing, the heaviest
kernel's roofline figure and the CPU oracle timed on this host.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# how oracle/Makefile builds the C restatement the CPU baseline times
ORACLE_FLAGS = "gcc -O3 -march=x86-64-v3 -ffp-contract=off"
# algorithmic bytes per point per fused PCG iteration: x, r (as u = Z r) and p
# read and written (6 words), hur/hvr-derived coefficients read (2 words),
# a 1-byte Dirichlet/land mask
BYTES_PER_PT = {"plain": 6 * 8 + 2 * 8 + 1, "mixed": 6 * 4 + 2 * 4 + 1,
                # polynomial-preconditioned FP32 sweep: x, r, z = M r and p
                # read and written, E and N, the mask
                "poly": 8 * 4 + 2 * 4 + 1}
# multigrid path: algorithmic bytes per point of the finest level for its
# heaviest kernels (FP32 level vectors, FP64 outer vectors): prolongation +
# post-smoothing reads x, b, E, N, C, the mask and a quarter of the coarse
# correction and writes y; pre-smoothing + restriction reads b, E, N, C and
# writes x and a quarter of the coarse right-hand side; outer pass A reads z,
# E, N, q, r, the mask and writes Az; outer pass B reads z, Az, p, q, r, x,
# the mask and writes p, q, r, x and the next V-cycle input
MG_KERNEL_BYTES = {"post_level0": 5 * 4 + 1 + 1 + 4, "pre_level0": 4 * 4 + 4 + 1,
                   "fcg_a": 4 + 2 * 8 + 2 * 8 + 1 + 8, "fcg_b": 4 + 8 + 4 * 8 + 1 + 4 * 8 + 4}
# the library's default smoothing degrees (psisolve.h): 5 for the
# streamfunction, 4 for the surface pressure
MG_DEFAULT_DEGREE = 5
PRESS_MG_DEFAULT_DEGREE = 4


def mg_kernel_names(degree):
    return {"post_level0": f"k_mg_post<{degree}, 8, 4> (finest level)",
            "pre_level0": f"k_mg_pre<{degree}, 8, 4> (finest level)",
            "fcg_a": "k_mg_fcg_a", "fcg_b": "k_mg_fcg_b"}
# HBM bytes per sweep launch at 2048^2 from rocprofv3 FETCH_SIZE / WRITE_SIZE
# (scaled by 2.0 / 1.0, the factors a 1 GiB device copy reads back): the
# Jacobi sweeps' wave-column march, profiles/k_cg_march_spherical_2048.md, and
# profiles/k_poly_pcg_spherical_2048.md
TRAFFIC_2048 = {"plain": 313.9e6, "mixed": 169.0e6, "poly": 175.5e6}


def sweep_kernel(n, fp64):
    """The Jacobi sweep kernel the library runs on an n x n grid: the
    wave-column march (strips of 60 columns x segments of 24 rows) when one
    round of its waves covers the grid (4 waves x 3 workgroups per CU on the
    MI355X's 256 CUs), else the LDS-tile sweep."""
    t = "double" if fp64 else "float"
    items = -(-n // 60) * -(-n // 24)
    return f"k_cg_march<{t}, 24, 6>" if items <= 4 * 3 * 256 else f"k_cg_sweep<{t}>"


# ... of the multigrid path's kernels (profiles/k_mg_spherical_2048.md)
MG_TRAFFIC_2048 = {"post_level0": 114.2e6, "pre_level0": 92.9e6, "fcg_a": 193.9e6,
                   "fcg_b": 347.6e6}
# ... of the surface-pressure CG kernels (profiles/k_sp_press_2048x40.md) and of the
# velocity update at 2048^2 x 40 levels (profiles/levels_kernels_step_2048x40.md)
PRESS_TRAFFIC_2048 = {"dir_apply": 283.2e6, "update": 206.4e6}
# ... of the non-hydrostatic CG kernels at 512^2 x 32 (profiles/k_nh_nonhydro_512x32.md)
NH_TRAFFIC_512x32 = {"dir": 216.9e6, "apply": 156.0e6, "update": 421.4e6}
VELOCITY_TRAFFIC_2048x40 = 16.288e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--modes", default="mg,learned,poly,mixed,plain",
                    help="first one is the headline")
    ap.add_argument("--weights", default="", help="learned preconditioner weights (.npz)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB (3 times): a "
                         "kernel of known traffic that calibrates HBM byte counters")
    ap.add_argument("--relres", type=float, default=1e-10)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-oracle", action="store_true",
                    help="solve workload: no oracle solve of the last timed forcing (no "
                         "psi_err / oracle_relres per path)")
    ap.add_argument("--cpu-iters", type=int, default=100,
                    help="CPU baseline: iterations per oracle run (repeated until the budget)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline: seconds of serial oracle work (half that threaded)")
    ap.add_argument("--check-every", type=int, default=0)
    ap.add_argument("--poly-degree", type=int, default=0, help="0: library default")
    ap.add_argument("--poly-lower", type=float, default=0.0)
    ap.add_argument("--mg-coarse-degree", type=int, default=0,
                    help="mg path: smoothing steps on the coarse levels (0: library default)")
    ap.add_argument("--inner-reduction", type=float, default=0.0,
                    help="mixed path: residual reduction per FP32 leg (0: library default)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--nz", type=int, default=40, help="levels of the step workload")
    ap.add_argument("--step-mode", default="mg", choices=["mg", "poly", "mixed", "plain"],
                    help="solve path of the step workload")
    ap.add_argument("--nh-n", type=int, default=512, help="nonhydro workload: nx = ny")
    ap.add_argument("--nh-nz", type=int, default=32, help="nonhydro workload: levels")
    ap.add_argument("--press-modes", default="mg,plain",
                    help="press workload: solve paths, the first the headline (mg: multigrid "
                         "to --relres; plain: the reference's CG and stopping rule)")
    ap.add_argument("--free-surface", type=int, default=1,
                    help="press workload: enable_free_surface (0: rigid lid)")
    ap.add_argument("--workload", default="solve",
                    choices=["solve", "ensemble", "step", "press", "nonhydro"],
                    help="solve: the 2048^2 two-island solve (headline); ensemble: 512 "
                         "batched basins of 64^2-512^2")
    ap.add_argument("--members", type=int, default=512)
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--ensemble-mode", default="poly", choices=["poly", "mixed"],
                    help="ensemble: FP32 legs preconditioned by the Chebyshev polynomial "
                         "(poly) or by Jacobi alone (mixed)")
    ap.add_argument("--tune", default="", help="fixed schedule tile_x,tile_y,tiles_per_item,"
                                                "items_per_grab[,groups_per_cu] instead of the "
                                                "autotuner")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def calibration_copy(mib, dev):
    """Three device copies of `mib` MiB: a kernel of known traffic that
    calibrates the profiler's byte counters."""
    if not mib:
        return
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] calibration copy: 3 x {mib} MiB")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_mib:
        import torch
        calibration_copy(args.copy_mib, torch.device("cuda:0"))
    if args.workload == "ensemble":
        return ensemble_main(args)
    if args.workload == "step":
        return step_main(args)
    if args.workload == "press":
        return press_main(args)
    if args.workload == "nonhydro":
        return nonhydro_main(args)
    import torch

    from pyom2_amd import basins
    from pyom2_amd.streamfunction import SolveOptions, StreamfunctionSolver, islands_trace

    dev = torch.device("cuda:0")
    t = time.time()
    # the 2048^2 basin's domain (102.4 deg of latitude about the equator) at
    # any resolution: spacing 0.05 deg * 2048 / n
    b = basins.spherical_basin(seed=0, n=args.n, n_islands=2, dlon=0.05 * 2048 / args.n)
    isl = islands_trace(b.kbot)
    log(f"[bench] basin {b.nx}x{b.ny} nisle={isl.nisle} built in {time.time() - t:.1f}s")
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=dev)
    nsteps = args.warmup + args.steps
    forcings = [torch.as_tensor(f, device=dev) for f in
                basins.ar1_forcings(b, nsteps, rho=0.99, seed=1, kmax=32)]
    torch.cuda.synchronize()
    log(f"[bench] {nsteps} forcings resident on the GPU")
    modes = args.modes.split(",")
    results = {}
    last_sol = {}
    for mode in modes:
        opts = SolveOptions(mode="mg" if mode == "learned" else mode, criterion="relres",
                            relres_target=args.relres, max_iterations=200000,
                            check_every=args.check_every, inner_reduction=args.inner_reduction,
                            poly_degree=args.poly_degree, poly_lower=args.poly_lower,
                            time_kernels=(mode == "mg"),
                            mg_coarse_degree=args.mg_coarse_degree)
        ls = None
        if mode == "learned":
            # flexible CG with the trained network composed with one
            # multigrid V-cycle per step (FP32 inner, FP64 residual
            # replacement), convergence polled after every step
            from pyom2_amd import learned
            net = learned.PrecondNet.load(args.weights or learned.DEFAULT_WEIGHTS, device=dev)
            ls = learned.LearnedSolver(s, net, compose_mg=True)
        dpsi = torch.zeros((3,) + b.shape, dtype=torch.float64, device=dev)
        taum1, tau, taup1 = 0, 1, 2
        per, choices = [], []
        for k in range(nsteps):
            if k == args.warmup:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            ts = time.perf_counter()
            torch.lerp(dpsi[taum1], dpsi[tau], 2.0, out=dpsi[taup1])   # 2 dpsi(tau) - dpsi(taum1)
            if ls is not None:
                its, conv, linfo = ls.solve(forcings[k], dpsi[taup1], relres_target=args.relres,
                                            max_iterations=2000, policy="learned", poll_every=1)
                choices.append(linfo["choice"])
                info = dict(relres_true=linfo["relres_true"], sweep_ms=0.0, sweep_launches=0,
                            outer_iterations=linfo["flushes"])
            else:
                its, conv, info = s.congrad_streamfunction(forcings[k], dpsi[taup1], opts)
            torch.cuda.synchronize()
            te = time.perf_counter()
            if not conv:
                raise SystemExit(f"{mode} solve {k} did not converge: {info}")
            if k >= args.warmup:
                ns = info.get("kernel_samples", 0)
                per.append(dict(ms=(te - ts) * 1e3, its=its, relres=info["relres_true"],
                                sweep_ms=info["sweep_ms"], launches=info["sweep_launches"],
                                outer=info["outer_iterations"], ksamples=ns,
                                ksum=np.array(info.get("kernel_us", [0.0] * 4)) * ns))
            if args.verbose:
                log(f"[bench] {mode} step {k}: {(te - ts) * 1e3:.2f} ms its={its} "
                    f"relres={info['relres_true']:.2e} outer={info['outer_iterations']}")
            taum1, tau, taup1 = tau, taup1, taum1
        torch.cuda.synchronize()
        total = time.perf_counter() - t0
        # the last timed solve's solution (level tau after the rotation), read
        # back after the timed region for the oracle comparison
        last_sol[mode] = dpsi[tau].cpu().numpy()
        its = np.array([p["its"] for p in per])
        sweep_ms = sum(p["sweep_ms"] for p in per)
        launches = sum(p["launches"] for p in per)
        npts = b.nx * b.ny
        # average duration of one fused sweep launch and the algorithmic bytes
        # it moves (early-exit launches after convergence move none)
        avg_launch_us = 1e3 * sweep_ms / max(launches, 1)
        bytes_per_launch = BYTES_PER_PT[mode if mode in BYTES_PER_PT else "mixed"] * npts * \
            its.sum() / max(launches, 1)
        achieved = bytes_per_launch / (avg_launch_us * 1e-6) / 1e9 if avg_launch_us > 0 else 0.0
        if mode == "mg":
            # the solve's heaviest kernels, timed inside the timed solves: HIP
            # events captured around them in the outer iteration's graph, read
            # at the host polls (the last iteration of every full batch)
            nsamp = sum(p["ksamples"] for p in per)
            if nsamp == 0:
                raise SystemExit("mg: no in-solve kernel timings were recorded")
            ksum = sum(p["ksum"] for p in per)
            kt = dict(zip(("pre_level0", "post_level0", "fcg_a", "fcg_b"),
                          (float(v) / nsamp for v in ksum)))
            top = max(kt, key=kt.get)
            mg_kernels = {k: dict(us=round(v, 3), bytes_per_pt=MG_KERNEL_BYTES[k],
                                  GBs=round(MG_KERNEL_BYTES[k] * npts / (v * 1e-6) / 1e9, 1))
                          for k, v in kt.items()}
            avg_launch_us = kt[top]
            achieved = MG_KERNEL_BYTES[top] * npts / (avg_launch_us * 1e-6) / 1e9
        results[mode] = dict(ms_per_solve=1e3 * total / args.steps,
                             solves_per_s=args.steps / total, iterations_mean=float(its.mean()),
                             relres_max=max(p["relres"] for p in per),
                             outer_mean=float(np.mean([p["outer"] for p in per])),
                             sweep_avg_launch_us=avg_launch_us, sweep_launches=launches,
                             achieved_GBs=achieved, frac=achieved / HBM_PEAK_GBS,
                             total_s=total)
        if mode == "mg":
            results[mode].update(us_per_iteration=1e3 * sweep_ms / max(launches, 1),
                                 kernels=mg_kernels, dominant=top, kernel_samples=nsamp)
        if ls is not None:
            results[mode].update(preconditioner="network + multigrid V-cycle",
                                 all_steps_learned=all(c == "learned" for c in choices))
            for key in ("sweep_avg_launch_us", "sweep_launches", "achieved_GBs", "frac"):
                results[mode].pop(key)
        log(f"[bench] {mode}: {results[mode]}")

    head = modes[0]
    r = results[head]
    # the roofline figure is the heaviest kernel of the headline path (of the
    # multigrid path when the learned one leads: its V-cycle is that path's)
    kpath = head if head != "learned" else next((m for m in ("mg", "poly", "mixed")
                                                    if m in results), None)
    rk = results.get(kpath, {}) if kpath else {}
    cpu = cpu_all = None
    oracle_ref = None
    if not args.skip_oracle:
        # every path's last solution against one oracle solve of the same
        # (last timed) forcing: the reference's CG and stopping rule
        # (congr_epsilon 1e-12) on all host threads, from a zero first guess
        oracle_ref = oracle_compare(b, isl, forcings[-1].cpu().numpy(), last_sol)
        for m, e in oracle_ref.pop("paths").items():
            results[m].update(e)
    if not args.skip_cpu:
        its_ref = results.get("plain", r)["iterations_mean"]
        cpu = cpu_baseline(b, isl, args.cpu_iters, its_ref, budget_s=args.cpu_seconds)
        if oracle_ref is not None:
            cpu_all = {"value": round(oracle_ref["ms"] * its_ref / oracle_ref["iterations"], 1),
                       "unit": "ms", "cores": oracle_ref["threads"], "kind": "port",
                       "sample": f"one whole solve of the last timed forcing from a zero first "
                                 f"guess at the reference's stopping rule: "
                                 f"{oracle_ref['iterations']} iterations in "
                                 f"{oracle_ref['ms'] / 1e3:.1f} s (OpenMP, "
                                 f"{oracle_ref['threads']} threads, C restatement "
                                 f"{ORACLE_FLAGS}), scaled to {its_ref:.0f} iterations "
                                 f"per solve"}
        else:
            cpu_all = cpu_baseline(b, isl, args.cpu_iters, its_ref, threaded=True,
                                   budget_s=args.cpu_seconds / 2)
    line = {
        "metric": f"ms per {args.n}\u00b2 psi solve to 1e-10 rel. FP64 residual: learned vs "
                  f"Jacobi PCG",
        "value": round(r["ms_per_solve"], 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_solve"], 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if head == "plain" else "f32-inner/f64-refined",
        "data": "synthetic",
        "config": {"workload": f"spherical basin {b.nx}x{b.ny}, {0.05 * 2048 / args.n:g} deg, "
                               f"2 islands, "
                               f"AR(1) rho=0.99 forcings, warm-started solves, relres<{args.relres:g}",
                   "path": head, "nisle": isl.nisle},
        "paths": {m: {k: (float(f"{v:.5g}") if isinstance(v, float) else v)
                      for k, v in d.items()} for m, d in results.items()},
        "roofline": ({"bound": "hbm",
                      "kernel": (mg_kernel_names(args.poly_degree or MG_DEFAULT_DEGREE)
                                 [rk["dominant"]] if kpath == "mg" else
                                 "k_poly_pcg<4, sweep>" if kpath == "poly" else
                                 sweep_kernel(b.nx, kpath == "plain")),
                      "path": kpath, "achieved": round(rk["achieved_GBs"], 1),
                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(rk["frac"], 4),
                      "traffic": (None if not b.nx == b.ny == 2048 else
                                  MG_TRAFFIC_2048.get(rk["dominant"]) if kpath == "mg" else
                                  TRAFFIC_2048.get(kpath)),
                      "traffic_unit": "bytes per launch",
                      "bytes_per_pt_iter": (MG_KERNEL_BYTES[rk["dominant"]] if kpath == "mg"
                                            else BYTES_PER_PT[kpath]),
                      "avg_launch_us": round(rk["sweep_avg_launch_us"], 3)} if rk else None),
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    line["solves_per_s"] = round(r["solves_per_s"], 2)
    if oracle_ref is not None:
        # the headline path's last solution against the oracle's solve of the
        # same forcing, and its FP64 true residual recomputed by the oracle
        line["psi_err"] = float(f"{r['psi_err']:.3g}")
        line["oracle_relres"] = float(f"{r['oracle_relres']:.3g}")
        line["oracle"] = {"iterations": oracle_ref["iterations"],
                          "converged": oracle_ref["converged"],
                          "relres": float(f"{oracle_ref['relres']:.3g}"),
                          "threads": oracle_ref["threads"],
                          "s": round(oracle_ref["ms"] / 1e3, 2),
                          "forcing": "the last timed forcing, zero first guess, "
                                     "congr_epsilon 1e-12 (the reference's stopping rule)"}
    if "plain" in results and head != "plain":
        line["speedup_vs_plain_fp64"] = round(results["plain"]["ms_per_solve"] / r["ms_per_solve"], 3)
    if "learned" in results:
        # the learned path's own figures (flexible CG, network + V-cycle)
        lr = results["learned"]
        line["learned"] = {"ms_per_solve": round(lr["ms_per_solve"], 3),
                           "iterations_mean": lr["iterations_mean"],
                           "relres_max": float(f"{lr['relres_max']:.3g}")}
        if "plain" in results:
            line["learned"]["speedup_vs_plain_fp64"] = round(
                results["plain"]["ms_per_solve"] / lr["ms_per_solve"], 2)
        if "mg" in results:
            line["learned"]["vs_mg_ms"] = round(lr["ms_per_solve"] / results["mg"]["ms_per_solve"], 2)
            line["learned"]["vs_mg_iterations"] = round(
                lr["iterations_mean"] / results["mg"]["iterations_mean"], 3)
        if "psi_err" in lr:
            line["learned"]["psi_err"] = float(f"{lr['psi_err']:.3g}")
    names = {"mg": "mg (FP64 flexible CG, FP32 multigrid V-cycle)",
             "learned": "learned (flexible CG, trained network + multigrid V-cycle)",
             "poly": "poly", "mixed": "mixed", "plain": "plain FP64 Jacobi-PCG"}
    line["config"]["path"] = names.get(head, head)
    print(json.dumps(finite(line)), flush=True)


def finite(v):
    """JSON-safe copy: non-finite floats become null."""
    if isinstance(v, dict):
        return {k: finite(x) for k, x in v.items()}
    if isinstance(v, (list, tuple)):
        return [finite(x) for x in v]
    if isinstance(v, float) and not np.isfinite(v):
        return None
    return v


def cpu_baseline(b, isl, iters, plain_its, threaded=False, budget_s=10.0):
    """The CPU oracle (C restatement of congrad_streamfunction; serial, the
    reference's own configuration, or with its grid loops on all OpenMP
    threads) on a bounded sample: `iters` PCG iterations of the same solve,
    scaled to the plain path's mean iteration count per solve."""
    from oracle import oracle as O
    cf = O.make_coeff(b)
    Z = O.make_inv(b.nx, b.ny, cf, O.Islands(isl.nisle, isl.offsets, isl.bi, isl.bj, isl.di,
                                             isl.dj, None))
    cores = O.threads() if threaded else 1
    n, dt = timed_oracle(b, cf, Z, iters, budget_s, threaded)
    per_it = dt / n
    how = f"OpenMP, {cores} threads" if threaded else "1 thread"
    return {"value": round(per_it * plain_its * 1e3, 1), "unit": "ms", "cores": cores,
            "kind": "port",
            "sample": f"{n} PCG iterations of the {b.nx}x{b.ny} solve from a zero first guess "
                      f"in {dt:.1f} s ({per_it * 1e3:.2f} ms/iteration, C restatement "
                      f"{ORACLE_FLAGS}, {how}) = {100.0 * n / plain_its:.1f}% of the "
                      f"{plain_its:.0f} iterations a solve takes, scaled to a whole solve"}


def oracle_compare(b, isl, forc, sols):
    """One oracle solve of `forc` (the reference's CG and stopping rule on all
    host threads, zero first guess) and, for every path's solution of the
    same forcing, max|psi - psi_ref| / max|psi_ref| and the FP64 true
    relative residual recomputed by the oracle."""
    from oracle import oracle as O
    oisl = O.Islands(isl.nisle, isl.offsets, isl.bi, isl.bj, isl.di, isl.dj, None)
    cf = O.make_coeff(b)
    Z = O.make_inv(b.nx, b.ny, cf, oisl)
    t = time.perf_counter()
    ref = O.congrad_threads(b.nx, b.ny, b.cyclic_x, cf, Z, forc, np.zeros(b.shape), 1e-12,
                            200000)
    ms = 1e3 * (time.perf_counter() - t)
    scale = np.abs(ref.sol).max()
    out = {"ms": ms, "iterations": ref.iterations, "converged": ref.converged,
           "threads": O.threads(), "relres": O.true_relres(b, cf, Z, forc, ref.sol), "paths": {}}
    for m, x in sols.items():
        out["paths"][m] = {"psi_err": float(np.abs(x - ref.sol).max() / scale),
                           "oracle_relres": O.true_relres(b, cf, Z, forc, x)}
    log(f"[bench] oracle: {ref.iterations} iterations in {ms / 1e3:.1f}s on {out['threads']} "
        f"threads, relres {out['relres']:.2e}; {out['paths']}")
    return out


def timed_oracle(b, cf, Z, iters, budget_s, threaded):
    """Runs the oracle's congrad from a zero first guess, `iters` iterations
    at a time, until `budget_s` seconds have passed; returns (iterations
    run, seconds)."""
    from oracle import oracle as O
    run = O.congrad_threads if threaded else O.congrad
    n, t0 = 0, time.perf_counter()
    while True:
        r = run(b.nx, b.ny, b.cyclic_x, cf, Z, b.forc, np.zeros(b.shape), 0.0, iters)
        n += r.iterations
        dt = time.perf_counter() - t0
        if dt >= budget_s:
            return n, dt


def ensemble_main(args):
    """512 independent basin solves per step, batched: N = nx*ny log-uniform
    in [64^2, 512^2], aspect in [0.5, 2], 0-3 islands, beta-plane
    construction with seeded bathymetry and forcing; every member from a zero
    first guess to a 1e-10 relative FP64 residual (FP32 inner PCG)."""
    import torch

    from pyom2_amd import basins
    from pyom2_amd.ensemble import EnsembleSolver
    from pyom2_amd.streamfunction import SolveOptions

    dev = torch.device("cuda:0")
    t = time.time()
    bs = basins.ensemble_basins(seed=0, count=args.members, lo=64, hi=512)
    log(f"[ensemble] {len(bs)} basins, {sum(b.nx * b.ny for b in bs) / 1e6:.1f} M points, "
        f"built in {time.time() - t:.1f}s")
    t = time.time()
    ens = EnsembleSolver(bs, device=dev)
    log(f"[ensemble] operators on the GPU in {time.time() - t:.1f}s, "
        f"{ens.active_points / 1e6:.1f} M active points")
    tune_rows = []
    if args.tune:
        v = [int(x) for x in args.tune.split(",")]
        ens.set_tune(tile_x=v[0], tile_y=v[1], tiles_per_item=v[2], items_per_grab=v[3],
                     groups_per_cu=v[4] if len(v) > 4 else None)
    elif not args.no_tune:
        t = time.time()
        best, tune_rows = ens.autotune(launches=24)
        log(f"[ensemble] autotune {time.time() - t:.1f}s -> {best}")
    forcs = [torch.as_tensor(b.forc, device=dev) for b in bs]
    opts = SolveOptions(mode=args.ensemble_mode, criterion="relres", relres_target=args.relres,
                        max_iterations=200000, poly_degree=args.poly_degree,
                        poly_lower=args.poly_lower)
    sols = ens.zeros()
    lat, its, per_step = [], [], []
    for k in range(args.warmup + args.steps):
        for x in sols:
            x.zero_()
        torch.cuda.synchronize()
        ts = time.perf_counter()
        res = ens.solve(forcs, sols, opts)
        torch.cuda.synchronize()
        te = time.perf_counter()
        if res.status != 0:
            raise SystemExit(f"ensemble step {k}: status {res.status}")
        if k >= args.warmup:
            per_step.append(te - ts)
            lat.extend(res.latency_ms.tolist())
            its.append(res.iterations)
        log(f"[ensemble] step {k}: {(te - ts) * 1e3:.1f} ms, its max {res.iterations.max()}, "
            f"relres max {res.relres_true.max():.2e}, launches {res.stats['launches']}")
    total = sum(per_step)
    lat = np.array(lat)
    its = np.array(its)
    npts = np.array([(b.nx * b.ny) for b in bs], np.float64)
    # algorithmic bytes: FP32 inner sweep bytes per point-iteration over the
    # members' interior points
    point_its = float((its * npts[None, :]).sum())
    achieved = BYTES_PER_PT[args.ensemble_mode] * point_its / total / 1e9
    cpu = cpu_all = None
    if not args.skip_cpu:
        its_jacobi = its.mean(axis=0)
        if args.ensemble_mode != "mixed":
            # the CPU runs Jacobi-PCG: one untimed batched Jacobi pass gives
            # each member's Jacobi iteration count
            for x in sols:
                x.zero_()
            rj = ens.solve(forcs, sols, SolveOptions(mode="mixed", criterion="relres",
                                                     relres_target=args.relres,
                                                     max_iterations=200000))
            its_jacobi = rj.iterations
        cpu = ensemble_cpu_baseline(bs, its_jacobi, args.cpu_iters, budget_s=args.cpu_seconds)
        cpu_all = ensemble_cpu_baseline(bs, its_jacobi, args.cpu_iters, threaded=True,
                                        budget_s=args.cpu_seconds / 2)
    solves_per_s = len(bs) * args.steps / total
    line = {
        "metric": "solves/s, 512-member batched psi ensemble (64^2-512^2) to 1e-10 rel. FP64 residual",
        "value": round(solves_per_s, 2),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-refined",
        "data": "synthetic",
        "config": {"workload": f"{len(bs)} beta-plane basins, N log-uniform in [64^2, 512^2], "
                               f"aspect 0.5-2, 0-3 islands, zero first guess, relres<{args.relres:g}",
                   "path": args.ensemble_mode,
                   "points": int(npts.sum()), "active_points": ens.active_points,
                   "tune": ens.tune},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3),
                       "max": round(float(lat.max()), 3)},
        "iterations": {"mean": round(float(its.mean()), 1), "max": int(its.max())},
        "roofline": {"bound": "hbm",
                     "kernel": "k_batch_poly" if args.ensemble_mode == "poly" else "k_batch_step",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "bytes_per_pt_iter": BYTES_PER_PT[args.ensemble_mode]},
        "autotune": tune_rows,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
    }
    print(json.dumps(finite(line)), flush=True)


def ensemble_cpu_baseline(bs, its_mean, iters, threaded=False, budget_s=10.0):
    """The CPU oracle on a bounded sample: `iters` Jacobi-PCG iterations on
    the largest member, as seconds per point-iteration, scaled to the
    ensemble's point-iterations of the batched Jacobi path (its iteration
    counts; one solve after the other, serial or with the grid loops on all
    OpenMP threads)."""
    from oracle import oracle as O
    big = max(bs, key=lambda b: b.nx * b.ny)
    isl = O.trace_islands(big)
    cf = O.make_coeff(big)
    Z = O.make_inv(big.nx, big.ny, cf, isl)
    cores = O.threads() if threaded else 1
    n, dt = timed_oracle(big, cf, Z, iters, budget_s, threaded)
    per_pt_it = dt / n / (big.nx * big.ny)
    total = per_pt_it * float(sum(n * (b.nx * b.ny) for n, b in zip(its_mean, bs)))
    how = f"OpenMP, {cores} threads" if threaded else "1 thread"
    return {"value": round(len(bs) / total, 3), "unit": "solves/s", "cores": cores,
            "kind": "port",
            "sample": f"{n} PCG iterations of the largest member ({big.nx}x{big.ny}) in "
                      f"{dt:.2f} s ({per_pt_it * 1e9:.2f} ns per point-iteration, C "
                      f"restatement, {how}), scaled to the ensemble's Jacobi point-iterations"}


# algorithmic bytes per point and level of the 3-D kernels around the solve:
# hydrostatic pressure (rho, maskT in; p_hydro out) + forcing (p_hydro; du, dv
# read and written; du_mix, dv_mix, maskU, maskV), and the velocity update
# (u, v, du, dv at tau; du, dv at taum1; du_mix, dv_mix, maskU, maskV in; u, v out)
LEVEL_BYTES = {"forcing": 24 + 72, "velocity": 96}


def step_main(args):
    """The whole solve_streamfunction per step (solve_streamfct.f90:10-120):
    hydrostatic pressure and forcing assembly over nz levels, the psi solve
    to a 1e-10 relative FP64 residual with the warm start, the island system,
    and the velocity update, on the 2048^2 spherical basin with two islands.
    The tendencies drift from step to step (AR(1), rho = 0.99, outside the
    timed calls); everything else is model state resident in HBM."""
    import torch

    from pyom2_amd import basins
    from pyom2_amd.streamfunction import (BarotropicState, Levels, MomentumState, SolveOptions,
                                          StreamfunctionSolver, islands_trace)

    dev = torch.device("cuda:0")
    calibration_copy(args.copy_mib, dev)
    nz = args.nz
    t = time.time()
    b0 = basins.spherical_basin(seed=0, n=args.n, n_islands=2, dlon=0.05 * 2048 / args.n)
    dzt, dzw, kb = basins.column_levels(b0, nz)
    isl = islands_trace(b0.kbot)
    # calc_topo's masks and depths (numerics.f90:215-259) on the device
    kb_d = torch.as_tensor(kb, device=dev)
    lvl = torch.arange(1, nz + 1, device=dev)[:, None, None]
    maskT = ((kb_d[None] != 0) & (kb_d[None] <= lvl)).double()
    maskU = maskT.clone()
    maskU[:, :, :-1] = torch.minimum(maskT[:, :, :-1], maskT[:, :, 1:])
    maskV = maskT.clone()
    maskV[:, :-1, :] = torch.minimum(maskT[:, :-1, :], maskT[:, 1:, :])
    hu = torch.zeros(b0.shape, dtype=torch.float64, device=dev)
    hv = torch.zeros_like(hu)
    for k in range(nz):
        hu = hu + maskU[k] * float(dzt[k])
        hv = hv + maskV[k] * float(dzt[k])
    hur = torch.where(hu != 0, 1.0 / torch.where(hu != 0, hu, 1.0), 0.0)
    hvr = torch.where(hv != 0, 1.0 / torch.where(hv != 0, hv, 1.0), 0.0)
    b = dataclasses.replace(b0, hur=hur.cpu().numpy(), hvr=hvr.cpu().numpy(),
                            maskU=maskU[-1].cpu().numpy(), maskV=maskV[-1].cpu().numpy())
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=dev)
    lev = Levels(dzt, dzw, maskT, maskU, maskV, dev)
    del maskT, maskU, maskV, hu, hv
    opts = SolveOptions(mode=args.step_mode, criterion="relres", relres_target=args.relres,
                        max_iterations=200000, poly_degree=args.poly_degree,
                        poly_lower=args.poly_lower)
    psin, line_psin, _ = s.streamfunction_init(opts)
    # synthetic model state of plausible magnitudes (seeded, on the device)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    sh3 = (nz,) + b.shape

    def rnd(scale, mask, levels=None):
        shape = sh3 if levels is None else (levels,) + sh3
        return scale * torch.randn(shape, generator=g, dtype=torch.float64, device=dev) * mask

    strat = torch.linspace(4.0, 0.0, nz, dtype=torch.float64, device=dev)[:, None, None]
    rho = (1024.0 + strat + 0.05 * torch.randn((3,) + sh3, generator=g, dtype=torch.float64,
                                                device=dev)) * lev.maskT
    ms = MomentumState(rho, rnd(0.1, lev.maskU, 3), rnd(0.1, lev.maskV, 3),
                       rnd(1e-6, lev.maskU, 3), rnd(1e-6, lev.maskV, 3), rnd(1e-7, lev.maskU),
                       rnd(1e-7, lev.maskV), torch.zeros(sh3, dtype=torch.float64, device=dev))
    st = BarotropicState.zeros(s, 3600.0, psin=psin.contiguous(), line_psin=line_psin)
    torch.cuda.synchronize()
    # rho, u, v, du, dv with 3 time levels; du_mix, dv_mix, p_hydro and 3 masks
    gb = 21 * nz * b.shape[0] * b.shape[1] * 8 / 1e9
    log(f"[step] basin {b.nx}x{b.ny} x {nz} levels, nisle={isl.nisle}, {gb:.1f} GB of 3-D "
        f"state resident, set up in {time.time() - t:.1f}s")
    c = float(np.sqrt(1 - 0.99 ** 2))
    per_step, its, solve_ms = [], [], []
    for k in range(args.warmup + args.steps):
        # drift the tendencies at the coming tau level (outside the timing)
        ms.du[st.tau] = 0.99 * ms.du[st.taum1] + rnd(c * 1e-6, lev.maskU)
        ms.dv[st.tau] = 0.99 * ms.dv[st.taum1] + rnd(c * 1e-6, lev.maskV)
        torch.cuda.synchronize()
        ts = time.perf_counter()
        n_it, conv, info = s.solve_streamfunction_3d(st, lev, ms, opts)
        torch.cuda.synchronize()
        te = time.perf_counter()
        if not conv:
            raise SystemExit(f"step {k}: solve did not converge ({info})")
        if k >= args.warmup:
            per_step.append(te - ts)
            its.append(n_it)
            solve_ms.append(info["elapsed_ms"])
        log(f"[step] {k}: {(te - ts) * 1e3:.2f} ms, solve {info['elapsed_ms']:.2f} ms, "
            f"{n_it} iterations, relres {info['relres_true']:.2e}")
        st.time_goes_by()
    # the 3-D kernels alone, timed with events on the solver's stream
    reps = 10
    fpx, fpy = s.zeros(), s.zeros()
    u_out = torch.empty(sh3, dtype=torch.float64, device=dev)
    v_out = torch.empty_like(u_out)

    def timed(fn):
        fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(s.stream):
            e0.record()
            for _ in range(reps):
                fn()
            e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    t_f = timed(lambda: s.streamfunction_forcing(lev, ms.rho[st.tau], ms.du[st.tau],
                                                 ms.dv[st.tau], ms.du_mix, ms.dv_mix, ms.p_hydro,
                                                 fpx, fpy))
    t_v = timed(lambda: s.velocity_update(lev, 3600.0, ms.u[st.tau], ms.v[st.tau], ms.du[st.tau],
                                          ms.dv[st.tau], ms.du[st.taum1], ms.dv[st.taum1],
                                          ms.du_mix, ms.dv_mix, st.psi[st.tau], u_out, v_out))
    pts = b.shape[0] * b.shape[1]
    kern = {}
    for name, ms_k in (("forcing", t_f), ("velocity", t_v)):
        gbs = LEVEL_BYTES[name] * pts * nz / (ms_k * 1e-3) / 1e9
        kern[name] = {"ms": round(ms_k, 4), "achieved": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4),
                      "bytes_per_pt_level": LEVEL_BYTES[name]}
    total = sum(per_step)
    cpu = None if args.skip_cpu else levels_cpu_baseline(nz, budget_s=args.cpu_seconds)
    line = {
        "metric": f"ms per {args.n}^2 x {nz}-level solve_streamfunction step (forcing, psi solve "
                  f"to {args.relres:g} rel. FP64 residual, velocity update)",
        "value": round(1e3 * total / args.steps, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64 (3-D parts); solve " + ("f64" if args.step_mode == "plain"
                                              else "f32-inner/f64-refined"),
        "data": "synthetic",
        "config": {"workload": f"spherical basin {b.nx}x{b.ny}, {nz} levels, 2 islands, "
                               f"drifting tendencies, warm start", "path": args.step_mode},
        "solve_ms_mean": round(float(np.mean(solve_ms)), 3),
        "iterations_mean": round(float(np.mean(its)), 1),
        "levels_kernels": kern,
        "roofline": {"bound": "hbm", "kernel": "k_velocity_update",
                     "achieved": kern["velocity"]["achieved"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": kern["velocity"]["frac"],
                     "traffic": (VELOCITY_TRAFFIC_2048x40 if b.nx == b.ny == 2048 and nz == 40
                                 else None)},
        "cpu_baseline_levels": cpu,
    }
    print(json.dumps(finite(line)), flush=True)


# algorithmic bytes per point of the surface-pressure CG kernels: k_sp_dir reads
# res, the old direction, 4 coefficient planes and the byte mask, writes the new
# direction and A p; k_sp_update reads psi, res, p, A p and the mask, writes psi
# and res.
PRESS_BYTES = {"dir_apply": 8 + 8 + 32 + 1 + 16, "update": 32 + 1 + 16}
PRESS_LEVEL_BYTES = {"forcing": 24 + 120, "velocity": 48}
# ... of the multigrid mode's kernels per interior point: the level kernels as
# in MG_KERNEL_BYTES; outer pass A reads z, the 4 coefficient planes, the two
# masks, q and r and writes Az; pass B as MG_KERNEL_BYTES
PRESS_MG_BYTES = {"post_level0": 26, "pre_level0": 21, "fcg_a": 4 + 32 + 1 + 1 + 16 + 8,
                  "fcg_b": 81}
# ... HBM bytes per launch at 2048^2 (profiles/k_spmg_press_2048x40.md)
PRESS_MG_TRAFFIC_2048 = {"post_level0": 113.3e6, "pre_level0": 92.3e6, "fcg_a": 265.2e6,
                         "fcg_b": 341.9e6}


def press_main(args):
    """The whole solve_pressure per step (solve_pressure.f90:9-84): hydrostatic
    pressure, forcing, congrad_surf_press with the reference's stopping rule
    (congr_epsilon 1e-12 absolute on the extrapolated step) from the
    2 psi(tau) - psi(taum1) first guess, velocity correction, on the 2048^2
    spherical basin with two islands and nz levels.  Tendencies drift from
    step to step (outside the timed calls)."""
    import torch

    from pyom2_amd import basins
    from pyom2_amd.pressure import SurfacePressureSolver
    from pyom2_amd.streamfunction import Levels, MomentumState, SolveOptions

    dev = torch.device("cuda:0")
    calibration_copy(args.copy_mib, dev)
    nz, fs, dt = args.nz, bool(args.free_surface), 3600.0
    t = time.time()
    b0 = basins.spherical_basin(seed=0, n=args.n, n_islands=2, dlon=0.05 * 2048 / args.n)
    dzt, dzw, kb = basins.column_levels(b0, nz)
    kb_d = torch.as_tensor(kb, device=dev)
    lvl = torch.arange(1, nz + 1, device=dev)[:, None, None]
    maskT = ((kb_d[None] != 0) & (kb_d[None] <= lvl)).double()
    maskU = maskT.clone()
    maskU[:, :, :-1] = torch.minimum(maskT[:, :, :-1], maskT[:, :, 1:])
    maskV = maskT.clone()
    maskV[:, :-1, :] = torch.minimum(maskT[:, :-1, :], maskT[:, 1:, :])
    hu = torch.zeros(b0.shape, dtype=torch.float64, device=dev)
    hv = torch.zeros_like(hu)
    for k in range(nz):
        hu = hu + maskU[k] * float(dzt[k])
        hv = hv + maskV[k] * float(dzt[k])
    s = SurfacePressureSolver(b0.nx, b0.ny, b0.dxt, b0.dxu, b0.dyt, b0.dyu, b0.cost, b0.cosu,
                              hu, hv, maskT[-1], free_surface=fs, dt_mom=dt, device=dev)
    host_op = (hu.cpu().numpy(), hv.cpu().numpy(), maskT[-1].cpu().numpy())
    lev = Levels(dzt, dzw, maskT, maskU, maskV, dev)
    del maskT, maskU, maskV, hu, hv
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    sh3 = (nz,) + b0.shape

    def rnd(scale, mask, levels=None):
        shape = sh3 if levels is None else (levels,) + sh3
        return scale * torch.randn(shape, generator=g, dtype=torch.float64, device=dev) * mask

    strat = torch.linspace(4.0, 0.0, nz, dtype=torch.float64, device=dev)[:, None, None]
    rho = (1024.0 + strat + 0.05 * torch.randn((3,) + sh3, generator=g, dtype=torch.float64,
                                                device=dev)) * lev.maskT
    ms = MomentumState(rho, rnd(0.1, lev.maskU, 3), rnd(0.1, lev.maskV, 3),
                       rnd(1e-6, lev.maskU, 3), rnd(1e-6, lev.maskV, 3), rnd(1e-7, lev.maskU),
                       rnd(1e-7, lev.maskV), torch.zeros(sh3, dtype=torch.float64, device=dev))
    psi = torch.zeros((3,) + b0.shape, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    log(f"[press] basin {b0.nx}x{b0.ny} x {nz} levels, free surface {int(fs)}, "
        f"set up in {time.time() - t:.1f}s")
    c = float(np.sqrt(1 - 0.99 ** 2))
    du0, dv0 = ms.du.clone(), ms.dv.clone()
    modes = [m for m in args.press_modes.split(",") if m]

    def run(mode):
        """warmup + steps of the whole routine with the given solve path, from
        the same initial state and the same drift for every path."""
        g.manual_seed(1)
        ms.du.copy_(du0)
        ms.dv.copy_(dv0)
        psi.zero_()
        opts = (SolveOptions(mode="mg", criterion="relres", relres_target=args.relres,
                             max_iterations=10000, time_kernels=True,
                             poly_degree=args.poly_degree, poly_lower=args.poly_lower)
                if mode == "mg" else
                SolveOptions(congr_epsilon=1e-12, max_iterations=200000))
        taum1, tau, taup1 = 0, 1, 2
        per_step, its, solve_ms = [], [], []
        ksum, nsamp = np.zeros(4), 0
        for k in range(args.warmup + args.steps):
            ms.du[tau] = 0.99 * ms.du[taum1] + rnd(c * 1e-6, lev.maskU)
            ms.dv[tau] = 0.99 * ms.dv[taum1] + rnd(c * 1e-6, lev.maskV)
            torch.cuda.synchronize()
            ts = time.perf_counter()
            n_it, conv, info = s.solve_pressure(psi, lev, ms, taum1, tau, taup1, opts)
            torch.cuda.synchronize()
            te = time.perf_counter()
            if not conv:
                raise SystemExit(f"{mode} step {k}: surface pressure did not converge ({info})")
            if k >= args.warmup:
                per_step.append(te - ts)
                its.append(n_it)
                solve_ms.append(info["elapsed_ms"])
                ksum += np.array(info["kernel_us"]) * info["kernel_samples"]
                nsamp += info["kernel_samples"]
            log(f"[press {mode}] {k}: {(te - ts) * 1e3:.2f} ms, solve {info['elapsed_ms']:.2f} "
                f"ms, {n_it} iterations, est. error {info['est_error']:.2e}, relres "
                f"{info['relres_true']:.2e}")
            taum1, tau, taup1 = tau, taup1, taum1
        return {"ms_per_step": round(1e3 * sum(per_step) / args.steps, 3),
                "solve_ms_mean": round(float(np.mean(solve_ms)), 3),
                "iterations_mean": round(float(np.mean(its)), 1),
                "us_per_iteration": round(1e3 * float(np.mean(solve_ms)) /
                                          max(float(np.mean(its)), 1.0), 2),
                "_kernels": (ksum, nsamp)}

    paths = {m: run(m) for m in modes}
    head = modes[0]
    pts = b0.shape[0] * b0.shape[1]
    if head == "mg":
        # the heaviest kernels timed inside the timed solves (HIP events in the
        # iteration graph, the last iteration of every poll batch)
        ksum, nsamp = paths["mg"]["_kernels"]
        if nsamp == 0:
            raise SystemExit("press mg: no in-solve kernel timings were recorded")
        kt = dict(zip(("pre_level0", "post_level0", "fcg_a", "fcg_b"),
                      (float(v) / nsamp for v in ksum)))
        pts_i = b0.nx * b0.ny
        kern = {}
        for name, us in kt.items():
            gbs = PRESS_MG_BYTES[name] * pts_i / (us * 1e-6) / 1e9
            kern[name] = {"us": round(us, 2), "achieved": round(gbs, 1),
                          "frac": round(gbs / HBM_PEAK_GBS, 4),
                          "bytes_per_pt": PRESS_MG_BYTES[name]}
        dom = max(kt, key=lambda x: kt[x])
        names = dict(mg_kernel_names(args.poly_degree or PRESS_MG_DEFAULT_DEGREE),
                     fcg_a="k_spmg_fcg_a")
        roof = {"bound": "hbm", "kernel": names[dom], "path": "mg",
                "achieved": kern[dom]["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": kern[dom]["frac"],
                "traffic": (PRESS_MG_TRAFFIC_2048.get(dom) if b0.nx == b0.ny == 2048 else None)}
    else:
        kt = s.kernel_times(reps=20)
        kern = {}
        for name in ("dir_apply", "update"):
            gbs = PRESS_BYTES[name] * pts / (kt[name] * 1e-6) / 1e9
            kern[name] = {"us": round(kt[name], 2), "achieved": round(gbs, 1),
                          "frac": round(gbs / HBM_PEAK_GBS, 4),
                          "bytes_per_pt": PRESS_BYTES[name]}
        kern["fold_a_us"] = round(kt["fold_a"], 2)
        kern["fold_b_us"] = round(kt["fold_b"], 2)
        dom = max(("dir_apply", "update"), key=lambda n: kt[n])
        roof = {"bound": "hbm", "kernel": "k_sp_dir" if dom == "dir_apply" else "k_sp_update2",
                "path": "plain", "achieved": kern[dom]["achieved"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": kern[dom]["frac"],
                "traffic": (PRESS_TRAFFIC_2048[dom] if b0.nx == b0.ny == 2048 else None)}
    for v in paths.values():
        v.pop("_kernels", None)
    plain_its = paths.get("plain", {}).get("iterations_mean")
    cpu = None if (args.skip_cpu or plain_its is None) else press_cpu_baseline(
        b0, host_op, nz, plain_its, fs, budget_s=args.cpu_seconds)
    r = paths[head]
    line = {
        "metric": f"ms per {args.n}^2 x {nz}-level solve_pressure step (forcing, surface-"
                  f"pressure solve, velocity correction)",
        "value": r["ms_per_step"],
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64" if head == "plain" else "f64 outer / f32 V-cycle",
        "data": "synthetic",
        "config": {"workload": f"spherical basin {b0.nx}x{b0.ny}, {nz} levels, 2 islands, "
                               f"{'free surface' if fs else 'rigid lid'}, dt 3600 s, drifting "
                               f"tendencies, first guess 2 psi(tau) - psi(taum1)",
                   "path": head,
                   "stopping": ("FP64 true relative residual <= %g" % args.relres
                                if head == "mg" else
                                "the reference's rule: congr_epsilon 1e-12 on the extrapolated "
                                "step")},
        "paths": paths,
        "press_kernels": kern,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "plain" in paths and head != "plain":
        line["speedup_vs_reference_algorithm"] = round(
            paths["plain"]["ms_per_step"] / r["ms_per_step"], 2)
    print(json.dumps(finite(line)), flush=True)


# algorithmic bytes per point of the non-hydrostatic CG kernels: k_nh_p reads
# res, the old direction and the packed mask byte and writes the new
# direction; k_nh_ap reads it and the mask and writes A p (the operator is
# rebuilt from the mask and 1-D factors); k_nh_update as k_sp_update.
NH_BYTES = {"dir": 8 + 8 + 1 + 8, "apply": 8 + 1 + 8, "update": 32 + 1 + 16}
NH_NAMES = {"dir": "k_nh_p", "apply": "k_nh_ap", "update": "k_nh_update"}


def nh_apply_kernel(nz):
    """The A p kernel the library runs for nz levels: four or two levels per
    thread where nz is a multiple of them (k_nh_ap_l), else one (k_nh_ap)."""
    return "k_nh_ap_l<4>" if nz % 4 == 0 else "k_nh_ap_l<2>" if nz % 2 == 0 else "k_nh_ap"


def nonhydro_main(args):
    """The whole solve_non_hydrostatic per step (non_hydrostatic.f90:6-64): w
    step, forcing, congrad_non_hydro with the reference's stopping rule
    (congr_epsilon_non_hydro 1e-12 on the extrapolated step) from the
    2 p(tau) - p(taum1) first guess, velocity correction; a 512^2 x 32-level
    basin at 50 m horizontal and 25 m vertical spacing (basins.nonhydro_box),
    seeded state, tendencies of w redrawn each step (outside the timed calls)."""
    import torch

    from pyom2_amd import basins
    from pyom2_amd.nonhydro import NonHydrostaticSolver
    from pyom2_amd.streamfunction import SolveOptions

    dev = torch.device("cuda:0")
    calibration_copy(args.copy_mib, dev)
    n, nz, dt = args.nh_n, args.nh_nz, 60.0
    t = time.time()
    b, col = basins.nonhydro_box(0, n, n, nz, False, 50.0, 25.0 * nz, 1.0, n_islands=2)
    mW = basins.mask_w(col.maskT)
    nh = basins.nonhydro_fields(col, 0)
    s = NonHydrostaticSolver.from_column(b, col, mW, device=dev)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    u, v, w, dw, dw_mix, p = (T(a) for a in (nh.u, nh.v, nh.w, nh.dw, nh.dw_mix, nh.p))
    u0, v0 = u.clone(), v.clone()
    mWd = T(mW)
    opts = SolveOptions(congr_epsilon=1e-12, max_iterations=200000)
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    torch.cuda.synchronize()
    log(f"[nonhydro] basin {b.nx}x{b.ny} x {nz} levels, set up in {time.time() - t:.1f}s")
    taum1, tau, taup1 = 0, 1, 2
    per_step, its, solve_ms = [], [], []
    for k in range(args.warmup + args.steps):
        # u, v at taup1 as the hydrostatic step leaves them; new w tendencies
        u[taup1] = u0[taup1]
        v[taup1] = v0[taup1]
        dw[tau] = 1e-8 * torch.randn(col.shape, generator=g, dtype=torch.float64,
                                     device=dev) * mWd
        dw[tau][-1] = 0.0
        torch.cuda.synchronize()
        ts = time.perf_counter()
        n_it, conv, info = s.solve_non_hydrostatic(u, v, w, dw, dw_mix, p, dt, taum1, tau,
                                                   taup1, 1, opts)
        torch.cuda.synchronize()
        te = time.perf_counter()
        if not conv:
            raise SystemExit(f"step {k}: non-hydrostatic solve did not converge ({info})")
        if k >= args.warmup:
            per_step.append(te - ts)
            its.append(n_it)
            solve_ms.append(info["elapsed_ms"])
        log(f"[nonhydro] {k}: {(te - ts) * 1e3:.2f} ms, solve {info['elapsed_ms']:.2f} ms, "
            f"{n_it} iterations, est. error {info['est_error']:.2e}")
        taum1, tau, taup1 = tau, taup1, taum1
    kt = s.kernel_times(reps=20)
    pts = nz * b.shape[0] * b.shape[1]
    kern = {}
    for name in ("dir", "apply", "update"):
        gbs = NH_BYTES[name] * pts / (kt[name] * 1e-6) / 1e9
        kern[name] = {"kernel": nh_apply_kernel(nz) if name == "apply" else NH_NAMES[name],
                      "us": round(kt[name], 2),
                      "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                      "bytes_per_pt": NH_BYTES[name]}
    kern["fold_a_us"] = round(kt["fold_a"], 2)
    kern["fold_b_us"] = round(kt["fold_b"], 2)
    dom = max(("dir", "apply", "update"), key=lambda x: kt[x])
    total = sum(per_step)
    it_mean = float(np.mean(its))
    cpu = None if args.skip_cpu else nonhydro_cpu_baseline(b, col, mW, nh, it_mean,
                                                           budget_s=args.cpu_seconds)
    line = {
        "metric": f"ms per {n}^2 x {nz}-level solve_non_hydrostatic step (w step, forcing, "
                  f"congrad_non_hydro to congr_epsilon_non_hydro 1e-12, velocity correction)",
        "value": round(1e3 * total / args.steps, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * total / args.steps, 3),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"basin {b.nx}x{b.ny}, {nz} levels, 50 m x 25 m spacing, "
                               f"2 islands, dt 60 s, first guess 2 p(tau) - p(taum1)"},
        "solve_ms_mean": round(float(np.mean(solve_ms)), 3),
        "iterations_mean": round(it_mean, 1),
        "us_per_iteration": round(1e3 * float(np.mean(solve_ms)) / max(it_mean, 1.0), 2),
        "nonhydro_kernels": kern,
        "roofline": {"bound": "hbm", "kernel": NH_NAMES[dom], "achieved": kern[dom]["achieved"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": kern[dom]["frac"],
                     "traffic": (NH_TRAFFIC_512x32.get(dom) if (n, nz) == (512, 32) else None)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(finite(line)), flush=True)


def nonhydro_cpu_baseline(b, col, mW, nh, iterations, budget_s=10.0):
    """The oracle's congrad_non_hydro on the same operator (1 core) for about
    budget_s, per iteration, scaled to the GPU run's iteration count, plus the
    oracle's forcing and velocity correction once."""
    from oracle import oracle as O
    from pyom2_amd.streamfunction import AB_EPS_REFERENCE
    t0 = time.perf_counter()
    u, v, w, forc = O.non_hydro_forcing(b, col, mW, 60.0, AB_EPS_REFERENCE, nh.w[1], nh.dw[1],
                                        nh.dw[0], nh.dw_mix, nh.u[2], nh.v[2], nh.w[2])
    O.non_hydro_velocity(b, col, mW, 60.0, nh.p[1], u, v, w)
    lev_s = time.perf_counter() - t0
    cf = O.make_coeff_non_hydro(b, col, mW)
    k = 2
    while True:
        t2 = time.perf_counter()
        O.congrad_non_hydro(b, col, cf, forc, nh.p[1], 1e-30, k)
        el = time.perf_counter() - t2
        if el * 4 > budget_s or k >= 4096:
            break
        k *= 4
    per_it = el / k
    total = lev_s + per_it * iterations
    return {"value": round(1e3 * total, 1), "unit": "ms per step (extrapolated)", "cores": 1,
            "kind": "port",
            "sample": f"{k} CG iterations on the same {b.nx}x{b.ny}x{col.nz} operator "
                      f"({per_it * 1e3:.1f} ms each) x {iterations:.0f} iterations of the GPU "
                      f"run, plus forcing and velocity correction once ({lev_s:.1f} s); C "
                      f"restatement, 1 thread"}


def press_cpu_baseline(b, host_op, nz, iterations, fs, budget_s=10.0, n3=512):
    """The oracle's solve_pressure on the CPU (1 core): CG iterations on the
    benchmark's own 2048^2 operator for about budget_s, and the forcing and
    velocity parts on an n3^2 x nz basin scaled per point-level, summed to one
    step of the GPU run's iteration count."""
    from oracle import oracle as O
    from pyom2_amd import basins
    from pyom2_amd.streamfunction import AB_EPS_REFERENCE, GRAV_REFERENCE, RHO_0_REFERENCE
    dt = 3600.0
    hu, hv, mT = host_op
    cf = O.make_coeff_surf_press(b, hu, hv, mT, fs, GRAV_REFERENCE, dt)
    rng = np.random.default_rng(0)
    forc = np.where(mT != 0, 1e-9 * rng.standard_normal(b.shape), 0.0)
    k = 2
    while True:
        t2 = time.perf_counter()
        O.congrad_surf_press(b, cf, mT, forc, np.zeros(b.shape), 1e-30, k)
        el = time.perf_counter() - t2
        if el * 4 > budget_s or k >= 4096:
            break
        k *= 4
    per_it = el / k
    b3, col = basins.column_basin(basins.spherical_basin(seed=0, n=n3, n_islands=2,
                                                         dlon=0.05 * 2048 / n3), nz)
    m = basins.momentum_fields(col, 0)
    t0 = time.perf_counter()
    p, du, dv, _, _ = O.streamfunction_forcing(b3, col, m.rho[1], m.du[1], m.dv[1], m.du_mix,
                                               m.dv_mix, GRAV_REFERENCE / RHO_0_REFERENCE)
    u, v, f3 = O.pressure_forcing(b3, col, dt, AB_EPS_REFERENCE, m.u[1], m.v[1], du, dv,
                                  m.du[0], m.dv[0], m.du_mix, m.dv_mix, m.psi[1], fs,
                                  GRAV_REFERENCE)
    O.pressure_velocity(b3, col, dt, m.psi[1], u, v)
    lev_s = (time.perf_counter() - t0) * (b.shape[0] * b.shape[1]) / (b3.shape[0] * b3.shape[1])
    total = lev_s + per_it * iterations
    return {"value": round(1e3 * total, 1), "unit": "ms per step (extrapolated)", "cores": 1,
            "kind": "port",
            "sample": f"{k} CG iterations on the {b.nx}x{b.ny} operator ({per_it * 1e3:.1f} ms "
                      f"each) x {iterations:.0f} iterations of the GPU run, plus forcing and "
                      f"velocity correction on {b3.nx}x{b3.ny}x{nz} scaled per point-level "
                      f"({lev_s:.1f} s); C restatement, 1 thread"}


def levels_cpu_baseline(nz, budget_s=10.0, n=512):
    """The oracle's forcing assembly and velocity update (1 core) on an n^2
    basin with nz levels, repeated within the budget, as ns per point-level."""
    from oracle import oracle as O
    from pyom2_amd import basins
    from pyom2_amd.streamfunction import AB_EPS_REFERENCE, GRAV_REFERENCE, RHO_0_REFERENCE
    b0 = basins.spherical_basin(seed=0, n=n, n_islands=0, dlon=0.05 * 2048 / n)
    b, col = basins.column_basin(b0, nz)
    m = basins.momentum_fields(col, 0)
    t0 = time.perf_counter()
    reps = 0
    while True:
        O.streamfunction_forcing(b, col, m.rho[1], m.du[1], m.dv[1], m.du_mix, m.dv_mix,
                                 GRAV_REFERENCE / RHO_0_REFERENCE)
        O.velocity_update(b, col, 3600.0, AB_EPS_REFERENCE, m.u[1], m.v[1], m.du[1], m.dv[1], m.du[0],
                          m.dv[0], m.du_mix, m.dv_mix, m.psi[2])
        reps += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = (time.perf_counter() - t0) / reps
    ns = dt / (b.shape[0] * b.shape[1] * nz) * 1e9
    return {"value": round(ns, 3), "unit": "ns per point-level (forcing + velocity update)",
            "cores": 1, "kind": "port",
            "sample": f"{reps} x (forcing + velocity update) on {b.nx}x{b.ny}x{nz} in "
                      f"{dt * reps:.1f} s, C restatement, 1 thread"}


if __name__ == "__main__":
    main()
