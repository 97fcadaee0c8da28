"""Pins the oracle's island-system step (solve_streamfct.f90:56-94) without
the reference build, whose solve_streamfunction needs LAPACK's DGESV.

The reference splits the streamfunction tendency of a basin with islands
into an interior solve with the warm start's perimeter values plus one
harmonic field psin(isle) per island (streamfct_init.f90:269-303), and picks
the island amplitudes dpsin(2:nisle, tau) so that the circulation of the
barotropic velocity around every island k >= 2 equals the line integral of
the forcing (fpx, fpy) around it (DGESV of line_psin, :61-88).  Here the
same tendency is computed in one piece instead: a sparse direct solve of
the augmented system whose unknowns are the tendency at every active point
and one constant per island k >= 2 added on its perimeter, with the operator
rows on the active points and the circulation rows of the islands.  The
oracle's composition (its CG solves, its line integrals, its dense solve and
the Adams-Bashforth step) must agree with it, on basins with 3-4 land
masses, cyclic and closed, random forcing and a non-zero dpsi/dpsin
history; copies of the oracle with a planted slip (sign, island index, time
level, Adams-Bashforth weight) must not."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import oracle as O
from tests import fixtures

CASES = ["islands_64", "cyclic_40x32", "tricky_36x28", "spherical_96"]
AB_EPS = float(np.float32(0.1))    # the reference's single-precision AB_eps
DT = 3600.0
TAUM1, TAU, TAUP1 = 0, 1, 2


def _cyclic_map(b):
    """Field index -> the field index setcyclic_xy copies it from (itself for
    non-halo columns or closed basins)."""
    ny4, nx4 = b.shape
    idx = np.arange(ny4 * nx4).reshape(ny4, nx4)
    if b.cyclic_x:
        src = idx.copy()
        nx = b.nx
        for i in (1, 2):                      # Fortran i = nx+i <- i; 1-i <- nx-i+1
            src[:, nx + i + 1] = idx[:, i + 1]
            src[:, 1 - i + 1] = idx[:, nx - i + 1 + 1]
        return src.ravel()
    return idx.ravel()


def _operator(b, cf):
    """apply_op (solve_pressure.f90:254-274) as a sparse matrix over the whole
    field, rows of the interior only."""
    ny4, nx4 = b.shape
    rows, cols, vals = [], [], []
    idx = np.arange(ny4 * nx4).reshape(ny4, nx4)
    J, I = np.meshgrid(np.arange(2, b.ny + 2), np.arange(2, b.nx + 2), indexing="ij")
    for jj in (-1, 0, 1):
        for ii in (-1, 0, 1):
            c = cf[(ii + 1) + 3 * (jj + 1)][J, I]
            nz = c != 0
            rows.append(idx[J, I][nz]); cols.append(idx[J + jj, I + ii][nz]); vals.append(c[nz])
    n = ny4 * nx4
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


def _circulation_rows(b, isl):
    """Row k: the linear functional psi -> line integral around island k of
    the barotropic velocity of psi (streamfct_init.f90:294-300, 311-377),
    evaluated on unit fields near the perimeters (it reads psi at most one
    point away from a perimeter point)."""
    ny4, nx4 = b.shape
    near = np.zeros(b.shape, bool)
    for k in range(isl.nisle):
        bi, bj, _, _ = isl.isle(k)
        for dj in (-1, 0, 1, 2):
            for di in (-1, 0, 1, 2):
                jj = np.clip(bj + 1 + dj, 0, ny4 - 1)
                ii = np.clip(bi + 1 + di, 0, nx4 - 1)
                near[jj, ii] = True
    L = np.zeros((isl.nisle, ny4 * nx4))
    e = np.zeros(b.shape)
    for q in np.flatnonzero(near.ravel()):
        e.flat[q] = 1.0
        u, v = O.psi_velocity(b, e)
        e.flat[q] = 0.0
        for k in range(isl.nisle):
            L[k, q] = O.line_integral(b, isl, k, u, v)
    return L


def _perimeter(b, isl, k):
    p = np.zeros(b.shape)
    bi, bj, _, _ = isl.isle(k)
    ok = (bi >= -1) & (bi <= b.nx + 2) & (bj >= -1) & (bj <= b.ny + 2)
    p[bj[ok] + 1, bi[ok] + 1] = 1.0
    return O.setcyclic(b.nx, b.ny, b.cyclic_x, p)


def augmented_tendency(b, init, fpx, fpy, warm):
    """The streamfunction tendency of the step in one sparse direct solve:
    psi_t = warm on non-active points + sum_k c_k on island k's perimeter
    (k >= 2), unknown on active points; A psi_t = curl(fpx, fpy) on active
    points and circulation_k(psi_t) = line integral of (fpx, fpy) around k.
    Returns (psi_t, c[2..nisle])."""
    isl = init.islands
    n = b.shape[0] * b.shape[1]
    src = _cyclic_map(b)
    active = (init.Z != 0).ravel()
    act_idx = np.flatnonzero(active)
    col_of = -np.ones(n, int)
    col_of[act_idx] = np.arange(len(act_idx))
    m = isl.nisle - 1
    # T maps (y, c) to the field; t0 is the fixed part
    base = O.setcyclic(b.nx, b.ny, b.cyclic_x, warm.copy()).ravel()
    rows, cols = [], []
    t0 = np.zeros(n)
    for q in range(n):
        s = src[q]
        if active[s]:
            rows.append(q); cols.append(col_of[s])
        else:
            t0[q] = base[q]
    vals = np.ones(len(rows))
    perims = [_perimeter(b, isl, k).ravel() for k in range(1, isl.nisle)]
    for k, pk in enumerate(perims):
        nzq = np.flatnonzero((pk != 0) & ~active[src])
        rows += list(nzq); cols += [len(act_idx) + k] * len(nzq)
        vals = np.concatenate([vals, pk[nzq]])
    T = sp.csr_matrix((vals, (rows, cols)), shape=(n, len(act_idx) + m))
    A = _operator(b, init.cf)
    forc = O.curl_forcing(b, fpx, fpy).ravel()
    Aa = A[act_idx]
    top = Aa @ T
    rhs_top = forc[act_idx] - Aa @ t0
    L = _circulation_rows(b, isl)[1:]
    lf = np.array([O.line_integral(b, isl, k, fpx, fpy) for k in range(1, isl.nisle)])
    bottom = sp.csr_matrix(L @ T)
    rhs_bot = lf - L @ t0
    # scale the circulation rows to the operator rows' magnitude
    s = abs(top).max() / max(abs(bottom).max(), 1e-300)
    M = sp.vstack([top, bottom * s]).tocsc()
    sol = spla.spsolve(M, np.concatenate([rhs_top, rhs_bot * s]))
    psi_t = (T @ sol + t0).reshape(b.shape)
    return psi_t, sol[len(act_idx):]


def harmonic_fields(b, init):
    """psin(isle) by sparse direct solves: A psin = 0 on active points, 1 on
    island isle's perimeter, 0 on the other non-active points
    (streamfct_init.f90:269-287 solves the same systems by CG)."""
    isl = init.islands
    A = _operator(b, init.cf)
    src = _cyclic_map(b)
    active = (init.Z != 0).ravel()
    act_idx = np.flatnonzero(active)
    n = A.shape[0]
    E = sp.csr_matrix((np.ones(np.count_nonzero(active[src])),
                       (np.flatnonzero(active[src]),
                        np.searchsorted(act_idx, src[active[src]]))), shape=(n, len(act_idx)))
    out = np.zeros((isl.nisle,) + b.shape)
    for k in range(isl.nisle):
        g = _perimeter(b, isl, k).ravel() * ~active[src]
        y = spla.spsolve((A[act_idx] @ E).tocsc(), -(A[act_idx] @ g))
        out[k] = (E @ y + g).reshape(b.shape)
    return out


def _inputs(b, nisle, seed):
    rng = np.random.default_rng(seed)
    sc = 1e-6                                  # forcing of the fixtures' magnitude
    fpx = sc * rng.standard_normal(b.shape)
    fpy = sc * rng.standard_normal(b.shape)
    fpx = O.setcyclic(b.nx, b.ny, b.cyclic_x, fpx)
    fpy = O.setcyclic(b.nx, b.ny, b.cyclic_x, fpy)
    amp = 1e3
    psi = np.zeros((3,) + b.shape)
    psi[TAU] = amp * rng.standard_normal(b.shape)
    dpsi = np.zeros((3,) + b.shape)
    dpsi[TAUM1] = amp * 1e-3 * rng.standard_normal(b.shape)
    dpsi[TAU] = amp * 1e-3 * rng.standard_normal(b.shape)
    dpsin = np.zeros((nisle, 3))
    dpsin[1:, TAUM1] = amp * 1e-3 * rng.standard_normal(nisle - 1)
    dpsin[1:, TAU] = amp * 1e-3 * rng.standard_normal(nisle - 1)
    return fpx, fpy, psi, dpsi, dpsin


@pytest.fixture(scope="module", params=CASES)
def case(request):
    b, d = fixtures.load(request.param)
    init = O.streamfunction_init(b, 1e-12, 3000)
    assert init.islands.nisle >= 3
    psin_direct = harmonic_fields(b, init)
    return b, init, psin_direct


def run_step(b, init, seed=7):
    fpx, fpy, psi, dpsi, dpsin = _inputs(b, init.islands.nisle, seed)
    hist = (psi.copy(), dpsi.copy(), dpsin.copy())
    r = O.solve_streamfunction_2d(b, init, fpx, fpy, psi, dpsi, dpsin, TAUM1, TAU, TAUP1, DT,
                                  AB_EPS, eps=1e-14, max_iter=20000)
    return (fpx, fpy), hist, (psi, dpsi, dpsin), r


def check_step(b, init, psin_direct, forcing, hist, out):
    """The checks every correct island step passes; returns the list of the
    ones that fail (empty = consistent)."""
    fpx, fpy = forcing
    psi0, dpsi0, dpsin0 = hist
    psi, dpsi, dpsin = out
    fails = []
    warm = 2 * dpsi0[TAU] - dpsi0[TAUM1]
    psi_t, c = augmented_tendency(b, init, fpx, fpy, warm)
    scale = np.abs(psi_t).max()
    # the tendency the reference forms: interior solve + island amplitudes x psin
    got_t = dpsi[TAUP1] + sum(dpsin[k, TAU] * init.psin[k] for k in range(1, init.islands.nisle))
    if np.abs(got_t - psi_t).max() > 1e-8 * scale:
        fails.append("tendency")
    if np.abs(dpsin[1:, TAU] - c).max() > 1e-8 * np.abs(c).max():
        fails.append("amplitudes")
    # the previous levels are left alone
    if not (np.array_equal(dpsin[:, TAUM1], dpsin0[:, TAUM1]) and dpsin[0, TAU] == dpsin0[0, TAU]
            and np.array_equal(dpsi[TAU], dpsi0[TAU]) and np.array_equal(psi[TAU], psi0[TAU])):
        fails.append("history")
    # the circulation constraint around every island k >= 2
    u, v = O.psi_velocity(b, got_t)
    for k in range(1, init.islands.nisle):
        lhs = O.line_integral(b, init.islands, k, u, v)
        rhs = O.line_integral(b, init.islands, k, fpx, fpy)
        if abs(lhs - rhs) > 1e-9 * max(abs(rhs), 1e-300):
            fails.append(f"circulation {k + 1}")
    # Adams-Bashforth on the whole tendency, with the harmonic fields of the
    # direct solves
    prev_t = dpsi0[TAU] + sum(dpsin0[k, TAUM1] * psin_direct[k]
                              for k in range(1, init.islands.nisle))
    want = psi0[TAU] + DT * ((1.5 + AB_EPS) * psi_t - (0.5 + AB_EPS) * prev_t)
    if np.abs(psi[TAUP1] - want).max() > 1e-8 * np.abs(want - psi0[TAU]).max():
        fails.append("adams-bashforth")
    return fails


def test_harmonic_fields_match_direct_solves(case):
    b, init, psin_direct = case
    for k in range(init.islands.nisle):
        err = np.abs(init.psin[k] - psin_direct[k]).max()
        assert err <= 1e-9, (k, err)
    # line_psin, the circulation of psin(isle) around n, is symmetric: the
    # orientation of the matrix handed to the dense solve cannot matter
    lp = init.line_psin
    assert np.abs(lp - lp.T).max() <= 1e-9 * np.abs(lp).max()


def test_island_step_matches_augmented_solve(case):
    b, init, psin_direct = case
    forcing, hist, out, r = run_step(b, init)
    assert r.converged
    assert check_step(b, init, psin_direct, forcing, hist, out) == []


def _flip_sign(orig):
    return lambda b, init, fpx, fpy, d: -orig(b, init, fpx, fpy, d)


def _shift_islands(orig):
    def f(init, rhs):
        x, info = O.dense_solve(init.line_psin[:-1, :-1], rhs[:-1])
        return x
    return f


def _wrong_velocity_level(orig):
    # line_psi0 from the first guess instead of the solved tendency
    def f(b, init, fpx, fpy, d):
        return orig(b, init, fpx, fpy, np.zeros_like(d))
    return f


def _ab_swapped(orig):
    def f(psi, dpsi, dpsin, psin, taum1, tau, taup1, dt, ab):
        psi[taup1] = psi[tau] + dt * ((1.5 + ab) * dpsi[taup1] - (0.5 + ab) * dpsi[tau])
        for k in range(1, dpsin.shape[0]):
            psi[taup1] = psi[taup1] + dt * ((1.5 + ab) * dpsin[k, taum1] -
                                            (0.5 + ab) * dpsin[k, tau]) * psin[k]
    return f


def _ab_no_islands(orig):
    def f(psi, dpsi, dpsin, psin, taum1, tau, taup1, dt, ab):
        psi[taup1] = psi[tau] + dt * ((1.5 + ab) * dpsi[taup1] - (0.5 + ab) * dpsi[tau])
    return f


def _store_taup1(orig):
    # amplitudes written to the taup1 level; the AB step then reads a stale tau
    def f(b, init, fpx, fpy, psi, dpsi, dpsin, taum1, tau, taup1, dt, ab, eps, max_iter):
        forc = O.curl_forcing(b, fpx, fpy)
        dpsi[taup1] = 2 * dpsi[tau] - dpsi[taum1]
        r = O.congrad(b.nx, b.ny, b.cyclic_x, init.cf, init.Z, forc, dpsi[taup1], eps, max_iter)
        dpsi[taup1] = O.setcyclic(b.nx, b.ny, b.cyclic_x, r.sol)
        dpsin[1:, taup1] = O.island_solve(init, O.island_rhs(b, init, fpx, fpy, dpsi[taup1]))
        O.ab_psi(psi, dpsi, dpsin, init.psin, taum1, tau, taup1, dt, ab)
        return r
    return f


PLANTED = {
    "sign of line_forc - line_psi0": ("island_rhs", _flip_sign),
    "island index off by one": ("island_solve", _shift_islands),
    "line_psi0 of the first guess": ("island_rhs", _wrong_velocity_level),
    "AB weights of the island term swapped": ("ab_psi", _ab_swapped),
    "island term left out of AB": ("ab_psi", _ab_no_islands),
    "amplitudes stored at taup1": ("solve_streamfunction_2d", _store_taup1),
}


@pytest.mark.parametrize("slip", sorted(PLANTED))
def test_planted_slips_are_caught(case, slip, monkeypatch):
    b, init, psin_direct = case
    name, make = PLANTED[slip]
    orig = getattr(O, name)
    fpx, fpy, psi, dpsi, dpsin = _inputs(b, init.islands.nisle, 7)
    hist = (psi.copy(), dpsi.copy(), dpsin.copy())
    if name == "solve_streamfunction_2d":
        make(orig)(b, init, fpx, fpy, psi, dpsi, dpsin, TAUM1, TAU, TAUP1, DT, AB_EPS, 1e-14,
                   20000)
    else:
        monkeypatch.setattr(O, name, make(orig))
        O.solve_streamfunction_2d(b, init, fpx, fpy, psi, dpsi, dpsin, TAUM1, TAU, TAUP1, DT,
                                  AB_EPS, eps=1e-14, max_iter=20000)
    monkeypatch.undo()
    assert check_step(b, init, psin_direct, (fpx, fpy), hist, (psi, dpsi, dpsin)) != []
