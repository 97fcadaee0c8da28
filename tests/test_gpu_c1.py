"""The headline configuration at full size on the GPU, through the C ABI:
the 2048^2 spherical basin (0.05 deg, 64 seeded features on a 5000 m base,
land where H < 100 m joined to the rim, two elliptical islands) that the
benchmark times, against the CPU oracle.

  * plain FP64 Jacobi-PCG (congrad_streamfunction, solve_streamfct.f90:
    153-327) at the reference's own stopping rule, against the oracle's
    threaded restatement from the same zero first guess (22.7k iterations;
    the serial oracle would take most of an hour): iteration counts within
    2% (+3), max|psi - psi_ref| / max|psi_ref| <= 1e-9;
  * every fast path (multigrid, mixed precision, polynomial, plain at the
    relative-residual criterion) to a 1e-10 FP64 relative residual: the
    residual recomputed by the oracle <= 1.01e-10, and the solution within
    1e-9 of psi_ref (multigrid also at 1e-12);
  * three warm-started time steps of solve_streamfunction's 2-D body
    (solve_streamfct.f90:49-94) with AR(1)-drifting forcing (rho = 0.99)
    and a non-zero island history, with psin/line_psin from the device's own
    streamfunction_init: the curl forcing against the oracle's, the warm
    start kept on the Dirichlet points, the interior tendency's FP64
    residual recomputed by the oracle, the island amplitudes against the
    oracle's island system on the same tendency, psi(taup1) against the
    oracle's Adams-Bashforth step, and the circulation constraint around
    both islands.
"""
import math
import time

import numpy as np
import pytest
import torch

from oracle import oracle as O
from pyom2_amd import basins
from pyom2_amd.streamfunction import (AB_EPS_REFERENCE, BarotropicState, SolveOptions,
                                      StreamfunctionSolver, islands_trace)

pytestmark = pytest.mark.gpu
DEV = "cuda"
N = 2048


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def H(t):
    return np.ascontiguousarray(t.detach().cpu().numpy())


def relerr(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def c1():
    t0 = time.time()
    b = basins.spherical_basin(seed=0, n=N, n_islands=2)
    isl = islands_trace(b.kbot)
    oisl = O.trace_islands(b)
    assert isl.nisle == oisl.nisle == 3
    for a, k in ((isl.offsets, "offsets"), (isl.bi, "bi"), (isl.bj, "bj"), (isl.di, "di"),
                 (isl.dj, "dj")):
        np.testing.assert_array_equal(a, getattr(oisl, k))
    cf = O.make_coeff(b)
    Z = O.make_inv(b.nx, b.ny, cf, oisl)
    print(f"\n[c1 {N}^2 basin, nisle=3, {time.time() - t0:.1f}s; oracle congrad on "
          f"{O.threads()} threads]", flush=True)
    t0 = time.time()
    ref = O.congrad_threads(b.nx, b.ny, False, cf, Z, b.forc, np.zeros(b.shape), 1e-12, 60000)
    print(f"[c1 oracle: {ref.iterations} iterations, converged={ref.converged}, "
          f"{time.time() - t0:.1f}s]", flush=True)
    assert ref.converged
    s = StreamfunctionSolver.from_basin(b, islands=isl, device=DEV)
    yield b, oisl, cf, Z, ref, s
    s.close()


def test_plain_matches_oracle_at_reference_rule(c1):
    b, _, cf, Z, ref, s = c1
    sol = s.zeros()
    its, conv, info = s.congrad_streamfunction(T(b.forc), sol, congr_epsilon=1e-12,
                                               max_iterations=60000)
    err = relerr(H(sol), ref.sol)
    print(f"  plain (reference rule): {its} iterations vs oracle {ref.iterations}, "
          f"max|psi-psi_ref|/max|psi_ref| = {err:.2e}", flush=True)
    assert conv
    assert abs(its - ref.iterations) <= 3 + 0.02 * ref.iterations
    assert err <= 1e-9


@pytest.mark.parametrize("mode,target", [("mg", 1e-10), ("mg", 1e-12), ("mixed", 1e-10),
                                         ("poly", 1e-10), ("plain", 1e-10)])
def test_fast_paths_relres_and_error(c1, mode, target):
    b, _, cf, Z, ref, s = c1
    sol = s.zeros()
    its, conv, info = s.congrad_streamfunction(T(b.forc), sol, mode=mode, criterion="relres",
                                               relres_target=target, max_iterations=200000)
    x = H(sol)
    rel = O.true_relres(b, cf, Z, b.forc, x)
    err = relerr(x, ref.sol)
    print(f"  {mode} to {target:g}: {its} iterations, oracle relres {rel:.3e} "
          f"(device {info['relres_true']:.3e}), max|psi-psi_ref|/max|psi_ref| = {err:.2e}",
          flush=True)
    assert conv
    assert rel <= 1.01 * target
    # the device's and the oracle's residuals round differently (the device
    # forms W and S from E and N with a metric ratio): they agree to 0.1%
    # above the rounding floor of forc - A x, which is ~2e-13 here
    assert abs(info["relres_true"] - rel) <= 1e-3 * rel + 3e-13
    assert err <= 1e-9


def _smooth_field(rng, b, kmax=32, modes=16, amp=1e-6):
    f = basins.fourier_forcing(rng, b.nx, b.ny, modes, kmax, amp, np.ones(b.shape))
    return f


def test_three_warm_started_steps_with_islands(c1):
    b, oisl, cf, Z, ref, s = c1
    t0 = time.time()
    psin, line_psin, its0 = s.streamfunction_init(SolveOptions(congr_epsilon=1e-12,
                                                               max_iterations=60000))
    torch.cuda.synchronize()
    P = H(psin)
    print(f"  streamfunction_init: {its0} iterations, {time.time() - t0:.1f}s", flush=True)
    # psin: harmonic off the perimeters (residual against its own starting
    # residual), 1 on its own perimeter, 0 on the others'; line_psin = the
    # oracle's line integrals of psin's velocity, and symmetric
    for k in range(3):
        p0 = np.zeros(b.shape)
        bi, bj, _, _ = oisl.isle(k)
        p0[bj + 1, bi + 1] = 1.0
        r = O.apply_op(b.nx, b.ny, cf, P[k])[Z != 0]
        r0 = O.apply_op(b.nx, b.ny, cf, p0)[Z != 0]
        assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(r0)
        np.testing.assert_array_equal(P[k][bj + 1, bi + 1], 1.0)
        u, v = O.psi_velocity(b, P[k])
        for n in range(3):
            lp = O.line_integral(b, oisl, n, u, v)
            assert abs(line_psin[n, k] - lp) <= 1e-12 * np.abs(line_psin).max()
    assert np.abs(line_psin - line_psin.T).max() <= 1e-8 * np.abs(line_psin).max()

    init = O.InitResult(oisl, P, np.ascontiguousarray(line_psin), cf, Z, its0)
    dt = 1800.0
    st = BarotropicState.zeros(s, dt, psin=psin.contiguous(), line_psin=line_psin)
    rng = np.random.default_rng(11)
    # a non-zero history: earlier tendencies and island amplitudes
    st.dpsi[st.taum1] = T(_smooth_field(rng, b, amp=1e-2))
    st.dpsi[st.tau] = T(_smooth_field(rng, b, amp=1e-2))
    st.psi[st.tau] = T(_smooth_field(rng, b, amp=1e3))
    st.dpsin[1:, st.taum1] = [0.3, -0.2]
    st.dpsin[1:, st.tau] = [0.25, -0.15]
    rho = 0.99
    fx = _smooth_field(rng, b)
    fy = _smooth_field(rng, b)
    opts = SolveOptions(mode="mg", criterion="relres", relres_target=1e-12,
                        max_iterations=2000)
    forc_out = s.zeros()
    for step in range(3):
        xi, eta = _smooth_field(rng, b), _smooth_field(rng, b)
        fx = rho * fx + math.sqrt(1 - rho * rho) * xi
        fy = rho * fy + math.sqrt(1 - rho * rho) * eta
        fpx = np.ascontiguousarray(fx * b.maskU)
        fpy = np.ascontiguousarray(fy * b.maskV)
        before = (H(st.psi), H(st.dpsi), st.dpsin.copy())
        its, conv, info = s.solve_streamfunction(st, T(fpx), T(fpy), opts, forc_out=forc_out)
        torch.cuda.synchronize()
        assert conv
        psi1, dpsi1, dpsin1 = H(st.psi), H(st.dpsi), st.dpsin.copy()
        tm1, tau, tp1 = st.taum1, st.tau, st.taup1
        forc = O.curl_forcing(b, fpx, fpy)
        assert relerr(H(forc_out), forc) <= 1e-15
        warm = 2 * before[1][tau] - before[1][tm1]
        fixed = Z == 0
        np.testing.assert_array_equal(dpsi1[tp1][fixed], warm[fixed])
        # the interior tendency solves A x = forc with the warm start's
        # perimeter values: FP64 residual recomputed by the oracle
        x = O.setcyclic(b.nx, b.ny, False, dpsi1[tp1].copy())
        rel = O.true_relres(b, cf, Z, forc, x)
        assert rel <= 1.01e-12, rel
        # the island system on that tendency
        amp = O.island_solve(init, O.island_rhs(b, init, fpx, fpy, dpsi1[tp1]))
        np.testing.assert_allclose(dpsin1[1:, tau], amp, rtol=1e-10)
        np.testing.assert_array_equal(dpsin1[:, tm1], before[2][:, tm1])
        # Adams-Bashforth, in the oracle's order
        psi_o, dpsi_o = before[0].copy(), dpsi1.copy()
        O.ab_psi(psi_o, dpsi_o, dpsin1, P, tm1, tau, tp1, dt, AB_EPS_REFERENCE)
        assert relerr(psi1[tp1], psi_o[tp1]) <= 1e-13
        # the circulation around both islands equals that of the forcing
        tot = dpsi1[tp1] + sum(dpsin1[k, tau] * P[k] for k in (1, 2))
        u, v = O.psi_velocity(b, tot)
        circ = []
        for k in (1, 2):
            lhs = O.line_integral(b, oisl, k, u, v)
            rhs = O.line_integral(b, oisl, k, fpx, fpy)
            circ.append((lhs, rhs))
            assert abs(lhs - rhs) <= 1e-9 * abs(rhs)
        print(f"  step {step}: {its} iterations, relres {rel:.2e}, dpsin {dpsin1[1:, tau]}, "
              f"circulation {circ}", flush=True)
        st.time_goes_by()
