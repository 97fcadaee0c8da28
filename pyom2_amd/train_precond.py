"""Train the learned preconditioner on the GPU and evaluate it.

    python -m pyom2_amd.train_precond --pairs mg --steps 12500 --basins 24 \
        --init pyom2_amd/weights/precond_unet.npz --out pyom2_amd/weights/precond_unet.npz

Training data: residual/error pairs on seeded basins of 256^2-1024^2 (the
spherical and beta-plane constructions, 0-3 islands): errors of the solver's
own iterates (x* - x_k with r_k = forc - A x_k) harvested once per basin --
multigrid flexible-CG iterates with --pairs mg (the errors a V-cycle leaves,
which the network composed with the V-cycle has to catch), plain FP64
Jacobi-PCG iterates with --pairs plain -- plus red-spectrum synthetic errors
e with r = A e drawn afresh every --refresh-every steps.  Loss: the A-norm
cosine (pyom2_amd.learned); optimiser: Adam.  The shipped weights are the
command above: 12,500 steps of batch 4 (50,000 pairs) on 24 basins, started
from weights trained on plain-CG pairs.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

from . import basins, learned
from .streamfunction import SolveOptions, StreamfunctionSolver, islands_trace

DEFAULT_WEIGHTS = learned.DEFAULT_WEIGHTS


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def training_basins(seed, count, sizes):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(sizes[k % len(sizes)])
        nisl = int(rng.integers(0, 4))
        s = int(rng.integers(0, 2**31 - 1))
        if k % 3 == 2:
            out.append(basins.beta_plane_basin(seed=s, nx=n, ny=n, n_islands=nisl))
        else:
            out.append(basins.spherical_basin(seed=s, n=n, n_islands=nisl))
    return out


# iterates harvested per basin: plain Jacobi-PCG, or multigrid flexible CG
# (whose errors are the modes the composed preconditioner must catch)
PAIR_ITERS = {"plain": (5, 20, 60, 150, 400, 1000, 2500), "mg": (1, 2, 3, 4, 5, 6, 8, 10)}


def build_sets(bs, rng, device="cuda", pairs="plain", n_synthetic=8):
    sets = []
    for b in bs:
        t = time.time()
        s = StreamfunctionSolver.from_basin(b, device=device)
        forc = torch.as_tensor(b.forc, device=device)
        ts = learned.make_trainset(s, forc, rng, n_synthetic=n_synthetic,
                                   cg_iters=PAIR_ITERS[pairs], mode=pairs)
        sets.append(ts)
        log(f"[data] {b.name} {b.nx}x{b.ny} nisle={s.nisle} pairs={ts.R.shape[0]} "
            f"({time.time() - t:.1f}s)")
    return sets


def refresh_synthetic(ts, rng, k_synth):
    """Replace the synthetic pairs of a set with fresh ones."""
    s = ts.solver
    n = ts.R.shape[0]
    act = (ts.inputs.Z32 != 0).double()
    for k in range(n - k_synth, n):
        e = torch.as_tensor(learned.smooth_error(rng, s), device=s.device) * act
        r = s.apply_op(e) * act
        ts.R[k].copy_(r)
        ts.inputs.build(ts.R[k].contiguous(), ts.X, k)
        out = learned.C.c_double()
        s.L.psi_nn_weighted_dot(s.ctx, learned._p(e.contiguous()), learned._p(r.contiguous()),
                                learned.C.byref(out))
        ts.ER[k] = out.value


def train_main(args):
    rng = np.random.default_rng(args.seed)
    bs = training_basins(args.seed, args.basins, [int(x) for x in args.sizes.split(",")])
    sets = build_sets(bs, rng, pairs=args.pairs, n_synthetic=args.synthetic)
    net = learned.PrecondNet(seed=args.seed)
    if args.init and os.path.exists(args.init):
        net = learned.PrecondNet.load(args.init)
    t = time.time()
    done = 0
    pairs_seen = 0
    while done < args.steps:
        chunk = min(args.refresh_every, args.steps - done)
        learned.train(net, sets, steps=chunk, batch=args.batch,
                      lr=args.lr * (0.5 ** (done // max(args.decay_every, 1))), seed=args.seed + done,
                      log_every=max(chunk // 2, 1), log=log)
        done += chunk
        pairs_seen += chunk * args.batch
        for ts in sets:
            refresh_synthetic(ts, rng, args.synthetic)
    log(f"[train] {done} steps, {pairs_seen} pairs seen, {time.time() - t:.1f}s")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    net.save(args.out)
    log(f"[train] weights -> {args.out}")
    return net


def zero_net(device="cuda"):
    """The network with every weight zero: the composed preconditioner without
    the network's term (its ablation)."""
    net = learned.PrecondNet(device=device)
    for k in net.w:
        net.w[k].zero_()
        net.b[k].zero_()
    net.repack()
    return net


def evaluate(net, n, seed=11, relres=1e-10, modes=("learned", "mixed", "plain")):
    b = basins.spherical_basin(seed=seed, n=n, n_islands=2, dlon=0.05 * 2048 / n)
    s = StreamfunctionSolver.from_basin(b)
    forc = torch.as_tensor(b.forc, device=s.device)
    res = {"n": n}
    for mode in modes:
        sol = s.zeros()
        torch.cuda.synchronize()
        t = time.perf_counter()
        if mode in ("learned", "learned+mg", "mg-fcg32"):
            ls = learned.LearnedSolver(s, zero_net() if mode == "mg-fcg32" else net,
                                       compose_mg=(mode != "learned"))
            its, conv, info = ls.solve(forc, sol, relres_target=relres, max_iterations=20000,
                                       policy="learned")
        else:
            its, conv, info = s.congrad_streamfunction(forc, sol, SolveOptions(
                mode=mode, criterion="relres", relres_target=relres, max_iterations=400000))
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t)
        res[mode] = {"its": its, "conv": conv, "ms": round(ms, 2),
                     "relres": info.get("relres_true")}
        log(f"[eval] n={n} {mode}: its={its} conv={conv} {ms:.1f} ms relres={info.get('relres_true')}")
    return res


def time_forward(n, reps=10, weights=DEFAULT_WEIGHTS):
    """Average time of the network's forward pass on an n x n solver grid's
    network layout (random input), HIP events around reps passes."""
    net = learned.PrecondNet.load(weights)
    H = (n + 4 + 7) // 8 * 8
    x = torch.randn(1, H, H, 16, device=net.dev).to(learned.BF16)
    net.forward(x, 1, H, H)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(net.stream):
        e0.record()
        for _ in range(reps):
            net.forward(x, 1, H, H)
        e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flop = sum(2 * 9 * ly.cin * ly.cout_pad * (H >> lvl) ** 2
               for ly, lvl in zip(learned.LAYERS, (0, 1, 2, 3, 2, 1, 0, 0)))
    log(f"[forward] {H}x{H}: {ms:.3f} ms, {flop / 1e9:.1f} GFLOP, "
        f"{flop / (ms * 1e-3) / 1e12:.0f} TFLOP/s")
    return ms


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--basins", type=int, default=12)
    ap.add_argument("--sizes", default="256,384,512,768,1024")
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--synthetic", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--decay-every", type=int, default=2000)
    ap.add_argument("--refresh-every", type=int, default=250)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default="")
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--eval", default="512,1024,2048")
    ap.add_argument("--eval-modes", default="learned,mixed,plain")
    ap.add_argument("--pairs", default="plain", choices=sorted(PAIR_ITERS),
                    help="iterates the (residual, error) pairs are harvested from")
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--time-forward", type=int, default=0,
                    help="only time the network's forward pass on an n x n grid")
    args = ap.parse_args()
    if args.time_forward:
        time_forward(args.time_forward)
        return
    if args.eval_only:
        net = learned.PrecondNet.load(args.out)
    else:
        net = train_main(args)
    out = [evaluate(net, int(n), modes=tuple(args.eval_modes.split(",")))
           for n in args.eval.split(",") if n]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
