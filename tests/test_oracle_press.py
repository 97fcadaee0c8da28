"""Pins the oracle's restatement of the surface-pressure solver
(solve_pressure.f90:9-312: forcing, make_coeff_surf_press,
congrad_surf_press, velocity correction) before the HIP path is checked
against it: the reference's routine (flang build of the pyOM2 sources,
golden fixtures and, where the build exists, fresh runs), and what the
mathematics fixes -- symmetry of the area-weighted operator, its null
space, the discrete continuity equation the corrected velocities satisfy,
and a dense direct solve -- chosen so that a wrong sign, index or weight
fails one of them."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import ref_run
from pyom2_amd import basins
from pyom2_amd.streamfunction import AB_EPS_REFERENCE as AB
from pyom2_amd.streamfunction import GRAV_REFERENCE as GRAV
from pyom2_amd.streamfunction import RHO_0_REFERENCE as RHO_0
from tests import fixtures

OUTS = ("p_hydro", "du", "dv", "psi", "u", "v")


def _whole(b, col, mom, dt, fs, eps=1e-12, max_iter=3000):
    p, r = O.solve_pressure(b, col, b.hu, b.hv, mom, 0, 1, 2, dt, AB, GRAV, RHO_0, fs, eps,
                            max_iter)
    return dict(p_hydro=p, du=mom.du[1], dv=mom.dv[1], psi=mom.psi[2], u=mom.u[2],
                v=mom.v[2]), r


@pytest.mark.parametrize("name", fixtures.PRESS_NAMES)
def test_whole_routine_matches_reference_bitwise(name):
    """Oracle composition == the reference's solve_pressure, bit for bit,
    including the iteration count and the kind of end (the cyclic rigid-lid
    fixture ends on the reference's divergence test, with its warning)."""
    b, col, mom, d = fixtures.load_press(name)
    out, r = _whole(b, col, mom, float(d["dt_mom"]), int(d["free_surface"]),
                    float(d["eps"]), int(d["max_iter"]))
    assert r.converged == (not int(d["ref_warning"]))
    assert r.status == (1 if int(d["ref_warning"]) else 0)
    assert r.iterations == int(d["iterations"])
    for k in OUTS:
        np.testing.assert_array_equal(out[k], d["out_" + k], err_msg=k)


def _case(seed, nx, ny, nz, cyclic, islands=1):
    b = basins.beta_plane_basin(seed=seed, nx=nx, ny=ny, n_islands=0 if cyclic else islands,
                                cyclic_x=cyclic)
    b3, col = basins.column_basin(b, nz)
    mom = basins.momentum_fields(col, seed)
    mom.psi[:] = mom.psi * 1e-6
    if cyclic:
        # a model state carries periodic halos (the step before set them)
        for a in (mom.rho, mom.u, mom.v, mom.du, mom.dv, mom.du_mix, mom.dv_mix, mom.psi):
            basins._setcyclic(a, b3.nx)
    return b3, col, mom


@pytest.mark.skipif(not ref_run.press_available(), reason="reference build not present")
@pytest.mark.parametrize("seed,nz,cyclic,fs", [(7, 1, False, 1), (8, 6, False, 0),
                                               (9, 3, True, 1)])
def test_fresh_cases_match_reference_bitwise(seed, nz, cyclic, fs):
    b, col, mom = _case(seed, 26, 18, nz, cyclic)
    ref = ref_run.run_press(b, col, b.hu, b.hv, mom, fs, 1e-12, 3000, 1800.0)
    out, r = _whole(b, col, mom, 1800.0, fs)
    assert r.iterations == ref["iterations"]
    for k in OUTS:
        np.testing.assert_array_equal(out[k], ref[k], err_msg=k)


def _area(b):
    return (b.cost[:, None] * b.dyt[:, None]) * b.dxt[None, :]


@pytest.mark.parametrize("fs", [0, 1])
def test_coefficients_symmetric_under_area_weights(fs):
    """cf(i,j) -> (i+1,j) times the T-cell area equals cf(i+1,j) -> (i,j)
    times its area (both are hu dyt / (dxu cos)), likewise north/south; the
    row sum is the free-surface term -1/(g dt^2) on ocean cells (zero for a
    rigid lid)."""
    b, col, _ = _case(3, 30, 22, 4, False, islands=2)
    dt = 900.0
    cf = O.make_coeff_surf_press(b, b.hu, b.hv, col.maskT[-1], fs, GRAV, dt)
    A = _area(b)
    cS, cW, cC, cE, cN = cf[1], cf[3], cf[4], cf[5], cf[7]
    assert not np.any(cf[[0, 2, 6, 8]])
    east = (A * cE)[2:-2, 2:-3]
    west = (A * cW)[2:-2, 3:-2]
    np.testing.assert_allclose(east, west, rtol=1e-13, atol=0)
    north = (A * cN)[2:-3, 2:-2]
    south = (A * cS)[3:-2, 2:-2]
    np.testing.assert_allclose(north, south, rtol=1e-13, atol=0)
    rows = cC + cE + cW + cN + cS
    want = -fs / (GRAV * dt * dt) * col.maskT[-1]
    want[:2], want[-2:], want[:, :2], want[:, -2:] = 0, 0, 0, 0
    scale = np.abs(cC).max()
    assert np.abs(rows - want).max() <= 1e-14 * scale
    assert cE.max() > 0 and cC.min() < 0


@pytest.mark.parametrize("fs,cyclic", [(1, False), (0, False), (1, True), (0, True)])
def test_corrected_velocities_satisfy_continuity(fs, cyclic):
    """After solve_pressure the new transport's divergence obeys the discrete
    continuity equation of the free surface eta = p/g:
        div(U_new) / dt - (p_tau - p_new) / (g dt^2) = forc - A p_new,
    the solver's residual, for any p_new; at convergence that residual is
    small next to the forcing.  This ties the forcing, the operator and the
    velocity correction together (each one's weights and signs)."""
    b, col, mom = _case(11, 28, 20, 5, cyclic)
    dt = 1200.0
    psi_tau = mom.psi[1].copy()
    _, r = O.solve_pressure(b, col, b.hu, b.hv, mom, 0, 1, 2, dt, AB, GRAV, RHO_0, fs,
                            1e-12, 5000)
    assert r.converged
    # the forcing of the solve (du, dv at tau now hold the hydrostatic gradient)
    u, v, forc = O.pressure_forcing(b, col, dt, AB, mom.u[1], mom.v[1], mom.du[1], mom.dv[1],
                                    mom.du[0], mom.dv[0], mom.du_mix, mom.dv_mix, psi_tau, fs,
                                    GRAV)
    p_new = mom.psi[2]
    cf = O.make_coeff_surf_press(b, b.hu, b.hv, col.maskT[-1], fs, GRAV, dt)
    res = np.zeros(b.shape)
    res[2:-2, 2:-2] = (forc - O.apply_op(b.nx, b.ny, cf, p_new))[2:-2, 2:-2]
    _, _, div_new = O.pressure_forcing(b, col, dt, 0.0, mom.u[2], mom.v[2], 0 * u, 0 * v,
                                       0 * u, 0 * v, 0 * u, 0 * v, p_new, 0, GRAV)
    lhs = div_new - fs * (psi_tau - p_new) / (GRAV * dt * dt) * col.maskT[-1]
    m = col.maskT[-1][2:-2, 2:-2] != 0
    scale = np.abs(forc[2:-2, 2:-2]).max()
    assert np.abs(lhs[2:-2, 2:-2] - res[2:-2, 2:-2])[m].max() <= 1e-11 * scale
    assert np.abs(res[2:-2, 2:-2])[m].max() <= 1e-6 * scale


def _dense(b, cf, act):
    """The operator restricted to active cells as a dense matrix."""
    idx = -np.ones(b.shape, int)
    jj, ii = np.nonzero(act)
    idx[jj, ii] = np.arange(len(jj))
    M = np.zeros((len(jj), len(jj)))
    for k, (dj, di, plane) in enumerate([(-1, 0, 1), (0, -1, 3), (0, 0, 4), (0, 1, 5),
                                         (1, 0, 7)]):
        c = cf[plane][jj, ii]
        t = idx[jj + dj, ii + di]
        ok = t >= 0
        M[np.arange(len(jj))[ok], t[ok]] += c[ok]
    return M, jj, ii


def test_free_surface_solution_matches_direct_solve():
    """congrad_surf_press on the free-surface system (nonsingular) reaches
    the dense direct solution of the same 5-point system."""
    b, col, _ = _case(5, 20, 16, 3, False, islands=1)
    dt = 600.0
    cf = O.make_coeff_surf_press(b, b.hu, b.hv, col.maskT[-1], 1, GRAV, dt)
    rng = np.random.default_rng(5)
    act = np.zeros(b.shape, bool)
    act[2:-2, 2:-2] = col.maskT[-1][2:-2, 2:-2] != 0
    forc = np.where(act, rng.standard_normal(b.shape), 0.0)
    r = O.congrad_surf_press(b, cf, col.maskT[-1], forc, np.zeros(b.shape), 1e-14, 5000)
    assert r.converged and r.status == 0
    M, jj, ii = _dense(b, cf, act)
    x = np.linalg.solve(M, forc[jj, ii])
    assert np.abs(r.sol[jj, ii] - x).max() <= 1e-9 * np.abs(x).max()


def test_rigid_lid_solution_up_to_a_constant():
    """Rigid lid: the operator annihilates constants on each connected ocean
    region; for a consistent forcing (zero area-weighted sum) the solve
    matches the least-squares solution up to a constant."""
    b, col, _ = _case(4, 22, 18, 2, False, islands=0)
    dt = 600.0
    cf = O.make_coeff_surf_press(b, b.hu, b.hv, col.maskT[-1], 0, GRAV, dt)
    act = np.zeros(b.shape, bool)
    act[2:-2, 2:-2] = col.maskT[-1][2:-2, 2:-2] != 0
    M, jj, ii = _dense(b, cf, act)
    assert np.abs(M.sum(axis=1)).max() <= 1e-14 * np.abs(M).max()
    rng = np.random.default_rng(4)
    f = 1e-6 * rng.standard_normal(len(jj))
    w = _area(b)[jj, ii]
    f -= (w * f).sum() / w.sum()           # consistent: range of M is orthogonal to w
    forc = np.zeros(b.shape)
    forc[jj, ii] = f
    r = O.congrad_surf_press(b, cf, col.maskT[-1], forc, np.zeros(b.shape), 1e-11, 5000)
    assert r.converged
    x = np.linalg.lstsq(M, f, rcond=None)[0]
    got = r.sol[jj, ii]
    d = (got - got.mean()) - (x - x.mean())
    assert np.abs(d).max() <= 1e-8 * np.abs(x - x.mean()).max()


def test_congrad_exits():
    """The reference's exits: a zero residual returns at once (iterations 0,
    normal end), the iteration cap is an error end with iterations =
    cap + 1 (the do-loop index after the loop)."""
    b, col, mom = _case(2, 20, 14, 2, False)
    cf = O.make_coeff_surf_press(b, b.hu, b.hv, col.maskT[-1], 1, GRAV, 1200.0)
    z = np.zeros(b.shape)
    r = O.congrad_surf_press(b, cf, col.maskT[-1], z, z, 1e-12, 100)
    assert r.iterations == 0 and r.converged and r.status == 0
    forc = O.apply_op(b.nx, b.ny, cf, np.where(col.maskT[-1] != 0, 1.0, 0.0))
    r = O.congrad_surf_press(b, cf, col.maskT[-1], forc, z, 1e-30, 3)
    assert r.iterations == 4 and not r.converged and r.status == 1
