"""The learned preconditioner at its evaluation size on the GPU: held-out
2048^2 spherical basins with two islands (seeds 11 and 12; the training
basins are 256^2-1024^2 with seeds drawn from numpy's generator at seed 0),
the shipped weights, flexible CG with the network composed with one
multigrid V-cycle per step.  Each solve is congrad_streamfunction's system
(solve_streamfct.f90:153-327) run to a 1e-10 FP64 relative residual.

  * the FP64 relative residual recomputed by the oracle <= 1.01e-10 and the
    solution within 1e-9 of max|psi| of a multigrid solve to 1e-12 (whose own
    residual the oracle confirms);
  * the network's gain over the same solver without it (every weight zero):
    fewer iterations;
  * the speed against plain FP64 Jacobi-PCG on the same GPU: at least 3x.
"""
import time

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pyom2_amd import basins, learned
from pyom2_amd.streamfunction import StreamfunctionSolver, islands_trace
from pyom2_amd.train_precond import zero_net

pytestmark = pytest.mark.gpu
DEV = "cuda"


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def H(t):
    return np.ascontiguousarray(t.detach().cpu().numpy())


@pytest.fixture(scope="module", params=[11, 12])
def heldout(request):
    seed = request.param
    t0 = time.time()
    b = basins.spherical_basin(seed=seed, n=2048, n_islands=2)
    isl = islands_trace(b.kbot)
    assert isl.nisle == 3
    oisl = O.trace_islands(b)
    cf = O.make_coeff(b)
    Z = O.make_inv(b.nx, b.ny, cf, oisl)
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=DEV)
    tight = s.zeros()
    its, conv, info = s.congrad_streamfunction(T(b.forc), tight, mode="mg", criterion="relres",
                                               relres_target=1e-12, max_iterations=200)
    x = H(tight)
    rel = O.true_relres(b, cf, Z, b.forc, x)
    print(f"\n[c4 held-out seed {seed}: mg to 1e-12 in {its} iterations, oracle relres "
          f"{rel:.2e}, {time.time() - t0:.1f}s]", flush=True)
    assert conv and rel <= 1.01e-12 + 3e-13
    yield b, cf, Z, x, s
    s.close()


def _learned_solve(s, b, net):
    ls = learned.LearnedSolver(s, net, compose_mg=True)
    f = T(b.forc)
    # the second solve is timed (the first builds the network's buffers)
    for _ in range(2):
        sol = s.zeros()
        torch.cuda.synchronize()
        t = time.perf_counter()
        its, conv, info = ls.solve(f, sol, relres_target=1e-10, max_iterations=400,
                                   policy="learned", poll_every=1)
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t)
    return its, conv, info, sol, ms


def test_learned_reaches_target_and_gains_over_no_network(heldout):
    b, cf, Z, tight, s = heldout
    net = learned.PrecondNet.load(learned.DEFAULT_WEIGHTS, device=DEV)
    its, conv, info, sol, ms = _learned_solve(s, b, net)
    x = H(sol)
    rel = O.true_relres(b, cf, Z, b.forc, x)
    err = float(np.abs(x - tight).max() / np.abs(tight).max())
    z_its, z_conv, _, _, z_ms = _learned_solve(s, b, zero_net(DEV))
    print(f"  learned: {its} iterations, {ms:.1f} ms, oracle relres {rel:.2e}, "
          f"max|psi-psi_tight|/max|psi_tight| {err:.2e}; without the network: {z_its} "
          f"iterations, {z_ms:.1f} ms", flush=True)
    assert conv and info["choice"] == "learned" and info["fallback_iterations"] == 0
    assert rel <= 1.01e-10
    assert err <= 1e-9
    assert z_conv and its < z_its


def test_learned_at_least_3x_faster_than_plain_fp64(heldout):
    b, cf, Z, tight, s = heldout
    net = learned.PrecondNet.load(learned.DEFAULT_WEIGHTS, device=DEV)
    its, conv, _, _, ms = _learned_solve(s, b, net)
    sol = s.zeros()
    torch.cuda.synchronize()
    t = time.perf_counter()
    p_its, p_conv, p_info = s.congrad_streamfunction(T(b.forc), sol, mode="plain",
                                                     criterion="relres", relres_target=1e-10,
                                                     max_iterations=100000)
    torch.cuda.synchronize()
    p_ms = 1e3 * (time.perf_counter() - t)
    print(f"  learned {ms:.1f} ms ({its} iterations) vs plain FP64 Jacobi-PCG {p_ms:.1f} ms "
          f"({p_its} iterations): {p_ms / ms:.1f}x", flush=True)
    assert conv and p_conv
    assert p_ms >= 3.0 * ms
