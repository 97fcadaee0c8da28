"""Pins the CPU oracle (oracle/psi_oracle.c) before anything is checked
against it: golden vectors from the reference's own Fortran routines,
closed forms, invariants and brute force on small grids."""
import math

import numpy as np
import pytest
import scipy.ndimage

from oracle import oracle as O
from oracle import ref_run
from pyom2_amd import basins
from tests import fixtures


@pytest.fixture(scope="module", params=fixtures.NAMES)
def golden(request):
    return fixtures.load(request.param)


def _islands_of(d):
    return O.Islands(int(d["nisle"]), d["offsets"], d["bi"], d["bj"], d["di"], d["dj"], None)


# ---------------------------------------------------------------------------
# golden vectors (reference Fortran, flang build) -- bit-exact
# ---------------------------------------------------------------------------
def test_islands_match_reference(golden):
    b, d = golden
    isl = O.trace_islands(b)
    assert isl.nisle == int(d["nisle"])
    for a, k in ((isl.offsets, "offsets"), (isl.bi, "bi"), (isl.bj, "bj"), (isl.di, "di"),
                 (isl.dj, "dj")):
        np.testing.assert_array_equal(a, d[k])


def test_congrad_matches_reference_bitwise(golden):
    b, d = golden
    isl = _islands_of(d)
    cf = O.make_coeff(b)
    Z = O.make_inv(b.nx, b.ny, cf, isl)
    r = O.congrad(b.nx, b.ny, b.cyclic_x, cf, Z, b.forc, d["sol0"], float(d["eps"]),
                  int(d["max_iter"]))
    assert r.iterations == int(d["iterations"])
    assert r.converged == bool(d["converged"])
    np.testing.assert_array_equal(r.sol, d["sol"])
    np.testing.assert_array_equal(O.apply_op(b.nx, b.ny, cf, d["sol"]), d["res"])


def test_psin_line_psin_lines_match_reference(golden):
    b, d = golden
    init = O.streamfunction_init(b, float(d["eps"]), 2000 if int(d["max_iter"]) < 100 else
                                 int(d["max_iter"]))
    if int(d["max_iter"]) >= 100:
        np.testing.assert_array_equal(init.psin, d["psin"])
        np.testing.assert_array_equal(init.line_psin, d["line_psin"])
    isl = _islands_of(d)
    lines = [O.line_integral(b, isl, k, d["uloc"], d["vloc"]) for k in range(isl.nisle)]
    np.testing.assert_array_equal(np.array(lines), d["lines"])


def test_reference_harness_on_fresh_case():
    """Where the reference build exists, compare on a case no fixture holds."""
    if not ref_run.available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    b = basins.spherical_basin(seed=21, n=80, ny=72, n_islands=3)
    r = ref_run.run(b, b.forc, np.zeros(b.shape), 1e-12, 3000)
    init = O.streamfunction_init(b, 1e-12, 3000)
    c = O.congrad(b.nx, b.ny, False, init.cf, init.Z, b.forc, np.zeros(b.shape), 1e-12, 3000)
    assert c.iterations == r["iterations"]
    np.testing.assert_array_equal(c.sol, r["sol"])
    np.testing.assert_array_equal(init.line_psin, r["line_psin"])


# ---------------------------------------------------------------------------
# closed forms and invariants
# ---------------------------------------------------------------------------
def _flat_box(nx=24, ny=18, H=3000.0, dx=10e3):
    ocean = np.ones((ny, nx), bool)
    ocean[0, :] = ocean[-1, :] = ocean[:, 0] = ocean[:, -1] = False
    grid = basins.make_grid(nx, ny, dx, dx)
    return basins._assemble("box", 0, nx, ny, grid, np.full((ny, nx), H), ocean,
                            np.random.default_rng(0), 4, 2, 1e-10)


def test_sine_modes_are_eigenvectors():
    """Constant depth, Cartesian: A = (1/H) Laplacian on the active interior
    with zero Dirichlet data; sine modes are exact eigenvectors."""
    H, dx = 3000.0, 10e3
    b = _flat_box(H=H, dx=dx)
    cf = O.make_coeff(b)
    isl = O.trace_islands(b)
    Z = O.make_inv(b.nx, b.ny, cf, isl)
    act = Z != 0
    # active psi points: Fortran i in [2, nx-2], j in [2, ny-2]
    m, n = b.nx - 3, b.ny - 3
    assert act.sum() == m * n
    for k, l in ((1, 1), (2, 3), (5, 1), (m, n)):
        ii = np.arange(1, m + 1)
        jj = np.arange(1, n + 1)
        mode = np.zeros(b.shape)
        mode[3:3 + n, 3:3 + m] = np.outer(np.sin(l * math.pi * jj / (n + 1)),
                                          np.sin(k * math.pi * ii / (m + 1)))
        lam = -(4.0 / (H * dx * dx)) * (math.sin(k * math.pi / (2 * (m + 1))) ** 2 +
                                         math.sin(l * math.pi / (2 * (n + 1))) ** 2)
        Am = O.apply_op(b.nx, b.ny, cf, mode)
        np.testing.assert_allclose(Am[act], lam * mode[act], rtol=0, atol=1e-12 * abs(lam))
        # and congrad solves A x = lam * mode back to the mode
        r = O.congrad(b.nx, b.ny, False, cf, Z, lam * mode, np.zeros(b.shape), 1e-16, 2000)
        assert r.converged
        np.testing.assert_allclose(r.sol[act], mode[act], rtol=0, atol=1e-9)


def test_area_weighted_symmetry_and_sign():
    """W A is symmetric negative definite on active points (W = area_z) for
    spherical metrics; A itself is not symmetric there."""
    b = basins.spherical_basin(seed=4, n=40, ny=36, n_islands=1)
    cf = O.make_coeff(b)
    isl = O.trace_islands(b)
    Z = O.make_inv(b.nx, b.ny, cf, isl)
    act = Z != 0
    rng = np.random.default_rng(1)
    W = b.area_z()
    u = np.where(act, rng.normal(size=b.shape), 0.0)
    v = np.where(act, rng.normal(size=b.shape), 0.0)
    Au = O.apply_op(b.nx, b.ny, cf, u)
    Av = O.apply_op(b.nx, b.ny, cf, v)
    a, c = (W * v * Au)[act].sum(), (W * u * Av)[act].sum()
    assert abs(a - c) <= 1e-12 * max(abs(a), abs(c))
    assert (W * u * Au)[act].sum() < 0
    vAu, uAv = (v * Au)[act].sum(), (u * Av)[act].sum()
    assert abs(vAu - uAv) > 1e-7 * abs(vAu)   # A alone is not symmetric


def test_row_sums_vanish_in_open_ocean():
    b = basins.beta_plane_basin(seed=2, nx=40, ny=30)
    cf = O.make_coeff(b)
    ones = np.ones(b.shape)
    A1 = O.apply_op(b.nx, b.ny, cf, ones)
    m4 = basins.psi_ocean_mask(b.maskT)
    inner = np.zeros(b.shape, bool)
    # points whose 4-cell neighbourhood is ocean in every direction
    inner[3:-3, 3:-3] = (m4[2:-4, 3:-3] * m4[4:-2, 3:-3] * m4[3:-3, 2:-4] * m4[3:-3, 4:-2] *
                         m4[3:-3, 3:-3]) > 0
    assert inner.sum() > 0
    np.testing.assert_allclose(A1[inner], 0.0, atol=1e-25)


def test_manufactured_and_dense_solution():
    """Brute force: the dense matrix of A on the active points, with Dirichlet
    data on one island, solved by LAPACK, against congrad."""
    b = basins.beta_plane_basin(seed=9, nx=22, ny=18, n_islands=1)
    cf = O.make_coeff(b)
    isl = O.trace_islands(b)
    assert isl.nisle == 2
    Z = O.make_inv(b.nx, b.ny, cf, isl)
    act = np.argwhere(Z != 0)
    nact = len(act)
    Adense = np.zeros((nact, nact))
    for q, (j, i) in enumerate(act):
        e = np.zeros(b.shape)
        e[j, i] = 1.0
        Adense[:, q] = O.apply_op(b.nx, b.ny, cf, e)[tuple(act.T)]
    # Dirichlet data: 1 on island 2's perimeter (a psin problem)
    g = np.zeros(b.shape)
    bi, bj, _, _ = isl.isle(1)
    g[bj + 1, bi + 1] = 1.0
    rhs = -O.apply_op(b.nx, b.ny, cf, g)[tuple(act.T)]
    xd = np.linalg.solve(Adense, rhs)
    r = O.congrad(b.nx, b.ny, False, cf, Z, np.zeros(b.shape), g, 1e-16, 5000)
    np.testing.assert_allclose(r.sol[tuple(act.T)], xd, rtol=0, atol=1e-10)
    assert np.all(r.sol[bj + 1, bi + 1] == 1.0)


def test_pcg_sign_invariants_and_forced_exits():
    b = basins.beta_plane_basin(seed=1, nx=40, ny=32)
    cf = O.make_coeff(b)
    Z = O.make_inv(b.nx, b.ny, cf, O.trace_islands(b))
    r = O.congrad(b.nx, b.ny, False, cf, Z, b.forc, np.zeros(b.shape), 1e-12, 1000, trace_cap=1000)
    t = r.trace
    assert np.all(t[:, 0] < 0) and np.all(t[:, 1] < 0) and np.all(t[:, 2] > 0)
    # max-iteration exit: iterations = max+1, converged false
    r3 = O.congrad(b.nx, b.ny, False, cf, Z, b.forc, np.zeros(b.shape), 1e-12, 3)
    assert (r3.iterations, r3.converged, r3.status) == (4, False, 1)
    # zero forcing and zero guess: immediate normal exit
    r0 = O.congrad(b.nx, b.ny, False, cf, Z, np.zeros(b.shape), np.zeros(b.shape), 1e-12, 100)
    assert (r0.iterations, r0.converged) == (0, True)
    # NaN forcing: error exit with a NaN estimate (where the reference halts)
    fn = b.forc.copy()
    fn[10, 10] = np.nan
    rn = O.congrad(b.nx, b.ny, False, cf, Z, fn, np.zeros(b.shape), 1e-12, 20)
    assert rn.status == 2 and not rn.converged


def test_line_integral_of_gradient_vanishes():
    """Discrete Stokes: the island integral of a masked T-point gradient field
    (u = maskU dphi/dx / cost, v = maskV dphi/dy as in solve_streamfct.f90:34-35)
    is 0 -- the reason surface pressure drops out of the island constraint."""
    b = basins.spherical_basin(seed=6, n=64, ny=56, n_islands=2)
    isl = O.trace_islands(b)
    rng = np.random.default_rng(3)
    phi = rng.normal(size=b.shape)
    u = np.zeros(b.shape)
    v = np.zeros(b.shape)
    u[:, :-1] = (phi[:, 1:] - phi[:, :-1]) / (b.dxu[None, :-1] * b.cost[:, None])
    v[:-1, :] = (phi[1:, :] - phi[:-1, :]) / b.dyu[:-1, None]
    u *= b.maskU   # velocities live on ocean faces only
    v *= b.maskV
    for k in range(isl.nisle):
        s = O.line_integral(b, isl, k, u, v)
        scale = O.line_integral(b, isl, k, np.abs(u), np.abs(v))
        assert abs(s) <= 1e-12 * abs(scale)


def test_island_count_is_8_connected_components():
    for seed in range(5):
        b = basins.beta_plane_basin(seed=seed, nx=50, ny=40, n_islands=3)
        isl = O.trace_islands(b)
        land = b.kbot == 0
        _, ncomp = scipy.ndimage.label(land, structure=np.ones((3, 3)))
        assert isl.nisle == ncomp
        # every boundary point is a psi point touching land and ocean
        for k in range(isl.nisle):
            bi, bj, di, dj = isl.isle(k)
            assert len(bi) > 0 and np.all(np.abs(di) + np.abs(dj) == 1)
            # the walk closes: last point + last direction = first point
            assert bi[-1] + di[-1] == bi[0] and bj[-1] + dj[-1] == bj[0]


def test_dense_solve_matches_lapack():
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 5):
        A = rng.normal(size=(n, n))
        rhs = rng.normal(size=n)
        x, info = O.dense_solve(A, rhs)
        assert info == 0
        np.testing.assert_allclose(x, np.linalg.solve(A, rhs), rtol=1e-12)
    _, info = O.dense_solve(np.array([[1.0, 2.0], [2.0, 4.0]]), np.array([1.0, 1.0]))
    assert info == 2


@pytest.mark.parametrize("name", ["c0_beta_plane_128", "spherical_96", "cyclic_40x32"])
def test_threaded_oracle_matches_serial(name):
    """The OpenMP variant (the benchmark's all-cores CPU baseline) runs the
    same algorithm: the same iterations up to reduction-order rounding and
    the reference's solution to 1e-9."""
    b, d = fixtures.load(name)
    isl = O.Islands(int(d["nisle"]), d["offsets"], d["bi"], d["bj"], d["di"], d["dj"], None)
    cf = O.make_coeff(b)
    Z = O.make_inv(b.nx, b.ny, cf, isl)
    r = O.congrad_threads(b.nx, b.ny, b.cyclic_x, cf, Z, b.forc, d["sol0"], float(d["eps"]),
                          int(d["max_iter"]))
    assert r.converged == bool(d["converged"])
    assert abs(r.iterations - int(d["iterations"])) <= 2 + 0.01 * int(d["iterations"])
    err = np.abs(r.sol - d["sol"]).max() / np.abs(d["sol"]).max()
    assert err <= 1e-9
    assert O.threads() >= 1
