"""The multigrid-preconditioned path (PSI_MODE_MG) on the GPU, through the C ABI.

It has no counterpart in the reference: a solve is checked by what fixes
its answer -- the FP64 true relative residual recomputed by the oracle on the
CPU (<= 1.01e-10, after the cyclic halo is set as solve_streamfunction sets
it) and the reference's own solutions (<= 1e-7 relative; those stop at the
reference's eps criterion) -- and the hierarchy by the aggregation rule it
states (2 x 2 cells of active points per level, down to at most 1024 active
cells), recomputed here with numpy."""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pyom2_amd import _lib, basins
from pyom2_amd._lib import PsiError, PsiInfo
from pyom2_amd.streamfunction import (Islands, SolveOptions, StreamfunctionSolver,
                                      islands_trace)
from tests import fixtures

pytestmark = pytest.mark.gpu
DEV = "cuda"


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def N(t):
    return t.detach().cpu().numpy()


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _oracle_relres(b, isl, sol):
    cf = O.make_coeff(b)
    Z = O.make_inv(b.nx, b.ny, cf, O.Islands(isl.nisle, isl.offsets, isl.bi, isl.bj, isl.di,
                                             isl.dj, None))
    return O.true_relres(b, cf, Z, b.forc, O.setcyclic(b.nx, b.ny, b.cyclic_x, sol.copy())), Z


def _levels_numpy(act, cyclic):
    """Active cells per level of 2 x 2 aggregation of active cells."""
    out = [(act.shape[1], act.shape[0], int(act.sum()))]
    while out[-1][2] > 1024:
        ny, nx = act.shape
        c = np.zeros(((ny + 1) // 2, (nx + 1) // 2), bool)
        jj, ii = np.nonzero(act)
        c[jj // 2, ii // 2] = True
        if c.size >= act.size:
            break
        act = c
        out.append((act.shape[1], act.shape[0], int(act.sum())))
    return out


@pytest.fixture(scope="module", params=fixtures.NAMES)
def case(request):
    print(f"\n[mg case {request.param}]", flush=True)
    b, d = fixtures.load(request.param)
    isl = Islands(int(d["nisle"]), d["offsets"], d["bi"], d["bj"], d["di"], d["dj"])
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=DEV)
    yield b, d, isl, s
    s.close()


def test_mg_solves_every_fixture(case):
    b, d, isl, s = case
    sol = T(d["sol0"])   # the fixture's first guess (its Dirichlet values stay)
    its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode="mg", criterion="relres",
                                               relres_target=1e-10, max_iterations=500)
    assert conv and info["status"] == 0
    assert its <= 60
    rel, Z = _oracle_relres(b, isl, N(sol))
    assert rel <= 1.01e-10
    assert abs(info["relres_true"] - rel) <= 1e-3 * 1e-10 + 1e-6 * rel
    act = Z != 0
    ref = d["sol"]
    if not int(d["converged"]):
        # the fixture stops the reference at its iteration cap: converge the
        # oracle (the same algorithm) instead
        cf = O.make_coeff(b)
        ref = O.congrad(b.nx, b.ny, b.cyclic_x, cf, Z, b.forc, d["sol0"], 1e-14, 20000).sol
    assert relerr(N(sol)[act], ref[act]) <= 1e-7


def test_mg_hierarchy_follows_the_aggregation_rule(case):
    b, d, isl, s = case
    s.mg_setup()
    lev = s.mg_levels()
    _, Z = _oracle_relres(b, isl, np.zeros(b.shape))
    want = _levels_numpy(Z[2:-2, 2:-2] != 0, b.cyclic_x)
    assert lev == want
    assert lev[-1][2] <= 1024


def test_mg_warm_start_and_iteration_cap(case):
    b, d, isl, s = case
    if int(d["converged"]):
        sol = T(d["sol"])
        its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode="mg",
                                                   criterion="relres", relres_target=1e-6,
                                                   max_iterations=500)
        assert conv and its == 0      # the reference's solution already meets 1e-6
    sol = s.zeros()
    its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode="mg", criterion="relres",
                                               relres_target=1e-15, max_iterations=1)
    assert not conv and its == 2 and info["status"] == _lib.PSI_NOT_CONVERGED
    rel, _ = _oracle_relres(b, isl, N(sol))
    assert 0 < rel < 1.0              # the capped iterate is kept
    assert abs(info["relres_true"] - rel) <= 1e-6 * rel


def test_mg_legs_past_the_attainable_residual():
    """A target below what the FP64 true residual can reach: the recursive
    residual meets it, the confirmation after the leg refuses it, and new legs
    start from the true residual until the iteration cap -- the reported true
    residual is the oracle's for the returned iterate (or, if the target is
    met after all, the oracle confirms it)."""
    b, d = fixtures.load("spherical_96")
    isl = Islands(int(d["nisle"]), d["offsets"], d["bi"], d["bj"], d["di"], d["dj"])
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=DEV)
    sol = s.zeros()
    target = 1e-15
    its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode="mg", criterion="relres",
                                               relres_target=target, max_iterations=120)
    rel, _ = _oracle_relres(b, isl, N(sol))
    print(f"  its={its} conv={conv} legs={info['outer_iterations']} "
          f"relres_true={info['relres_true']:.3e} oracle={rel:.3e}", flush=True)
    # (at 1e-15 the residual itself is rounding: FP64 sums in another order
    # differ by about eps |A| |x| / |b|)
    assert abs(info["relres_true"] - rel) <= 1e-3 * rel + 5e-15
    if conv:
        assert rel <= 1.01 * target
    else:
        assert info["status"] == _lib.PSI_NOT_CONVERGED and its == 121
        assert info["outer_iterations"] >= 2     # at least one leg ended on a refused confirmation
        assert rel < 1e-12
    s.close()


@pytest.mark.parametrize("degree", [1, 2, 4])
def test_mg_smoothing_degrees(degree):
    b, d = fixtures.load("spherical_96")
    isl = Islands(int(d["nisle"]), d["offsets"], d["bi"], d["bj"], d["di"], d["dj"])
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=DEV)
    sol = s.zeros()
    its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode="mg", criterion="relres",
                                               relres_target=1e-10, max_iterations=500,
                                               poly_degree=degree)
    assert conv
    rel, _ = _oracle_relres(b, isl, N(sol))
    assert rel <= 1.01e-10
    s.close()


@pytest.mark.parametrize("coarse", [3, 8, 10])
def test_mg_coarse_smoothing_degree(coarse):
    """The coarse levels (at most 256 tiles of 64 x 16) smoothed with their
    own degree: same FP64 residual bound, and a degree above 10 refused."""
    b = basins.spherical_basin(seed=5, n=512, n_islands=2)
    s = StreamfunctionSolver.from_basin(b, device=DEV)
    isl = s.islands
    sol = s.zeros()
    its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode="mg", criterion="relres",
                                               relres_target=1e-10, max_iterations=500,
                                               mg_coarse_degree=coarse)
    print(f"  coarse degree {coarse}: {its} iterations", flush=True)
    assert conv
    rel, _ = _oracle_relres(b, isl, N(sol))
    assert rel <= 1.01e-10
    with pytest.raises(PsiError):
        s.congrad_streamfunction(T(b.forc), s.zeros(), mode="mg", criterion="relres",
                                 relres_target=1e-10, max_iterations=10, mg_coarse_degree=11)
    s.close()


def test_mg_argument_errors():
    b, d = fixtures.load("islands_64")
    isl = Islands(int(d["nisle"]), d["offsets"], d["bi"], d["bj"], d["di"], d["dj"])
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=DEV)
    L = _lib.lib()
    sol = s.zeros()
    info = PsiInfo()
    o = SolveOptions(mode="mg", criterion="relres").as_c()
    # the C ABI refuses a multigrid solve before psi_mg_setup
    rc = L.psi_congrad(s.ctx, C.byref(o), C.c_void_p(T(b.forc).data_ptr()),
                       C.c_void_p(sol.data_ptr()), C.byref(info))
    assert rc == _lib.PSI_BAD_ARG and b"psi_mg_setup" in L.psi_ctx_last_error(s.ctx)
    s.mg_setup()
    with pytest.raises(PsiError):
        s.congrad_streamfunction(T(b.forc), sol, mode="mg", criterion="reference")
    with pytest.raises(PsiError):
        s.congrad_streamfunction(T(b.forc), sol, mode="mg", criterion="relres", poly_degree=7)
    s.close()


def test_mg_spherical_1024_two_islands():
    """At a large size: FP64 true residual recomputed on the CPU, the
    iteration count, and agreement with the polynomial path."""
    print("\n[mg spherical 1024]", flush=True)
    b = basins.spherical_basin(seed=3, n=1024, n_islands=2)
    isl = islands_trace(b.kbot)
    assert isl.nisle == 3
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=DEV)
    sols = {}
    for mode in ("mg", "poly"):
        sol = s.zeros()
        its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode=mode, criterion="relres",
                                                   relres_target=1e-10, max_iterations=50000)
        print(f"  1024 {mode}: its={its} relres={info['relres_true']:.2e}", flush=True)
        assert conv
        rel, _ = _oracle_relres(b, isl, N(sol))
        assert rel <= 1.01e-10
        sols[mode] = N(sol)
        if mode == "mg":
            assert its <= 40
    assert relerr(sols["mg"], sols["poly"]) < 1e-6
    s.close()


def test_mg_in_solve_kernel_timing():
    """time_kernels records HIP events inside the iteration graph: the solve
    is bit-identical to the untimed one, the first poll batch (every replay of
    it iterating) gives the solve's one sample, and the four durations are
    positive and below the iteration's time."""
    print("\n[mg kernel timing 512]", flush=True)
    b = basins.spherical_basin(seed=4, n=512, n_islands=2)
    isl = islands_trace(b.kbot)
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=DEV)
    out = {}
    for timed in (False, True, True):
        sol = s.zeros()
        its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode="mg", criterion="relres",
                                                   relres_target=1e-10, max_iterations=500,
                                                   check_every=2, time_kernels=timed)
        assert conv
        if timed:
            assert info["kernel_samples"] == 1
            us_iter = 1e3 * info["sweep_ms"] / info["sweep_launches"]
            assert all(0.0 < u < us_iter for u in info["kernel_us"])
            np.testing.assert_array_equal(N(sol), out[False][1])
            assert its == out[False][0]
        else:
            assert info["kernel_samples"] == 0 and info["kernel_us"] == [0.0] * 4
        out[timed] = (its, N(sol))
    # the back-to-back timings of the same four kernels between solves
    kt = s.mg_kernel_times(reps=3)
    assert set(kt) == {"post_level0", "pre_level0", "fcg_a", "fcg_b"}
    assert all(u > 0.0 for u in kt.values())
    s.close()


def test_mg_cyclic_tall_tiles():
    """A cyclic basin large enough for the 64 x 32 tiles on its finest level
    (2048 x 512: 512 tiles), whose frames wrap across the seam: the FP64 true
    residual recomputed on the CPU meets the target."""
    print("\n[mg cyclic 2048x512]", flush=True)
    b = basins.beta_plane_basin(seed=5, nx=2048, ny=512, n_islands=2, cyclic_x=True)
    isl = islands_trace(b.kbot, cyclic_x=True)
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=DEV)
    sol = s.zeros()
    its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode="mg", criterion="relres",
                                               relres_target=1e-10, max_iterations=500)
    print(f"  its={its} relres={info['relres_true']:.2e}", flush=True)
    assert conv and its <= 60
    rel, _ = _oracle_relres(b, isl, N(sol))
    assert rel <= 1.01e-10
    s.close()


@pytest.mark.parametrize("nx,ny,cyclic", [(63, 50, False), (63, 50, True), (130, 66, False),
                                          (130, 66, True), (514, 258, True)])
def test_mg_odd_and_even_widths(nx, ny, cyclic):
    """The outer passes run on pairs of points when nx is even and point by
    point when it is odd; both, closed and cyclic (where the pair at the
    seam wraps its west or east neighbour), reach the target on the FP64 true
    residual recomputed on the CPU, and the 514-wide basin has frames that
    wrap across the seam on its 64 x 32 tiles."""
    print(f"\n[mg width {nx}x{ny} cyclic={cyclic}]", flush=True)
    b = basins.beta_plane_basin(seed=nx + ny, nx=nx, ny=ny, n_islands=1, cyclic_x=cyclic)
    isl = islands_trace(b.kbot, cyclic_x=cyclic)
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=DEV)
    sol = s.zeros()
    its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode="mg", criterion="relres",
                                               relres_target=1e-10, max_iterations=500)
    rel, _ = _oracle_relres(b, isl, N(sol))
    print(f"  its={its} relres={info['relres_true']:.2e} oracle {rel:.2e}", flush=True)
    assert conv and rel <= 1.01e-10
    s.close()
