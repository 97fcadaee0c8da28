"""The non-hydrostatic pressure solver on the GPU (psi_nh_*,
psi_solve_non_hydrostatic) through the C ABI, against the oracle
(non_hydrostatic.f90 restated in C) and the reference's fixtures.

Bar: the coefficients, apply_op_3D, the forcing (with the w step and the
cyclic halos it sets) and the velocity correction are bit-identical;
congrad_non_hydro sums its inner products in another order than the
reference's serial loops, so its iterates agree to rounding: iteration
counts within 2, p within 1e-9 of max|p| (the solve stops at an estimated
error of 1e-12 absolute on p of order 0.1 to 1).  At 256^2 x 16 the corrected
velocities satisfy the discrete continuity equation to the solver's residual."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from pyom2_amd import _lib, basins
from pyom2_amd._lib import PsiError
from pyom2_amd.nonhydro import NonHydrostaticSolver
from pyom2_amd.streamfunction import AB_EPS_REFERENCE as AB
from pyom2_amd.streamfunction import SolveOptions
from tests import fixtures

pytestmark = pytest.mark.gpu
DEV = "cuda"
P_TOL = 1e-9


def T(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=DEV)


def N(t):
    return t.detach().cpu().numpy()


def _state(nh):
    return [T(a) for a in (nh.u, nh.v, nh.w, nh.dw, nh.dw_mix, nh.p)]


@pytest.mark.parametrize("name", fixtures.NH_NAMES)
def test_coefficients_and_apply_bitwise(name):
    print(f"\n[nh coeff {name}]", flush=True)
    b, col, mW, nh, d = fixtures.load_nonhydro(name)
    s = NonHydrostaticSolver.from_column(b, col, mW, device=DEV)
    cf = O.make_coeff_non_hydro(b, col, mW)
    np.testing.assert_array_equal(N(s.make_coeff_non_hydro()), cf)
    np.testing.assert_array_equal(N(s.make_coeff_non_hydro()), d["cf7"])
    p = np.random.default_rng(1).standard_normal(col.shape)
    np.testing.assert_array_equal(N(s.apply_op_3d(T(p))), O.apply_op_3d(b, col.nz, cf, p))
    s.close()


def _check_whole(b, col, mW, nh, ref, dt, eps, max_iter, itt):
    s = NonHydrostaticSolver.from_column(b, col, mW, device=DEV)
    u, v, w, dw, dw_mix, p = _state(nh)
    forc = s.zeros()
    opts = SolveOptions(congr_epsilon=eps, max_iterations=max_iter)
    its, conv, info = s.solve_non_hydrostatic(u, v, w, dw, dw_mix, p, dt, 0, 1, 2, itt, opts,
                                              forc_out=forc)
    assert conv and info["status"] == _lib.PSI_OK
    assert abs(its - int(ref["iterations"])) <= 2
    # before the solve: the w step, the cyclic halos and the forcing, bit for bit
    u_s, v_s, w_s, forc_o = O.non_hydro_forcing(b, col, mW, dt, AB, nh.w[1], nh.dw[1], nh.dw[0],
                                                nh.dw_mix, nh.u[2], nh.v[2], nh.w[2])
    np.testing.assert_array_equal(N(forc), forc_o)
    # the solve
    p_new = N(p[2])
    scale = np.abs(ref["p"]).max()
    assert np.abs(p_new - ref["p"]).max() <= P_TOL * scale
    if itt == 0:
        np.testing.assert_array_equal(N(p[1]), p_new)
        np.testing.assert_array_equal(N(p[0]), p_new)
    else:
        np.testing.assert_array_equal(N(p[1]), nh.p[1])
    # the correction, given the GPU's p: bit for bit
    u_o, v_o, w_o = O.non_hydro_velocity(b, col, mW, dt, p_new, u_s, v_s, w_s)
    np.testing.assert_array_equal(N(u[2]), u_o)
    np.testing.assert_array_equal(N(v[2]), v_o)
    np.testing.assert_array_equal(N(w[2]), w_o)
    for k, a in (("u", u), ("v", v), ("w", w)):
        r = ref[k]
        assert np.abs(N(a[2]) - r).max() <= 1e-8 * max(np.abs(r).max(), 1e-300)
    s.close()
    return its


@pytest.mark.parametrize("name", fixtures.NH_NAMES)
def test_solve_non_hydrostatic_matches_reference(name):
    print(f"\n[nh fixture {name}]", flush=True)
    b, col, mW, nh, d = fixtures.load_nonhydro(name)
    ref = {"iterations": d["iterations"], "p": d["out_p"], "u": d["out_u"], "v": d["out_v"],
           "w": d["out_w"]}
    _check_whole(b, col, mW, nh, ref, float(d["dt_mom"]), float(d["eps"]), int(d["max_iter"]),
                 int(d["itt"]))


@pytest.mark.parametrize("nx,cyclic,itt", [(21, True, 0), (19, False, 1)])
def test_odd_row_length_against_oracle(nx, cyclic, itt):
    print(f"\n[nh odd nx={nx} cyclic={cyclic}]", flush=True)
    b, col = basins.nonhydro_box(5, nx, 15, 6, cyclic, 80.0, 600.0, 1.03)
    mW = O.mask_w(col.maskT)
    nh = basins.nonhydro_fields(col, 5, b.nx if cyclic else None)
    nh0 = basins.NonHydro(*[a.copy() for a in (nh.u, nh.v, nh.w, nh.dw, nh.dw_mix, nh.p)])
    r, _ = O.solve_non_hydrostatic(b, col, mW, nh0, 0, 1, 2, 60.0, AB, 1e-12, 5000, itt)
    assert r.converged
    ref = {"iterations": r.iterations, "p": nh0.p[2], "u": nh0.u[2], "v": nh0.v[2],
           "w": nh0.w[2]}
    _check_whole(b, col, mW, nh, ref, 60.0, 1e-12, 5000, itt)


def test_congrad_exits():
    """Zero residual: 0 iterations; the cap: cap + 1 and PSI_NOT_CONVERGED,
    the iterate kept; options the reference algorithm does not take."""
    print("\n[nh congrad exits]", flush=True)
    b, col = basins.nonhydro_box(2, 24, 18, 5, False, 60.0, 500.0, 1.0)
    mW = O.mask_w(col.maskT)
    s = NonHydrostaticSolver.from_column(b, col, mW, device=DEV)
    cf = O.make_coeff_non_hydro(b, col, mW)
    rng = np.random.default_rng(2)
    x0 = rng.standard_normal(col.shape) * col.maskT
    forc = O.apply_op_3d(b, col.nz, cf, rng.standard_normal(col.shape) * col.maskT)
    x = T(x0)
    its, conv, info = s.congrad_non_hydro(T(O.apply_op_3d(b, col.nz, cf, x0)), x,
                                          congr_epsilon=1e-12, max_iterations=100)
    assert its == 0 and conv
    np.testing.assert_array_equal(N(x), x0)
    for cap in (1, 2, 5):
        x = T(x0)
        its, conv, info = s.congrad_non_hydro(T(forc), x, congr_epsilon=1e-30,
                                              max_iterations=cap)
        r = O.congrad_non_hydro(b, col, cf, forc, x0, 1e-30, cap)
        assert its == cap + 1 == r.iterations and not conv
        assert info["status"] == _lib.PSI_NOT_CONVERGED
        assert np.abs(N(x) - r.sol).max() <= 1e-12 * np.abs(r.sol).max()
    x = T(x0)
    its, conv, info = s.congrad_non_hydro(T(forc), x, congr_epsilon=1e-12, max_iterations=5000)
    r = O.congrad_non_hydro(b, col, cf, forc, x0, 1e-12, 5000)
    assert conv and r.converged and abs(its - r.iterations) <= 2
    assert np.abs(N(x) - r.sol).max() <= P_TOL * np.abs(r.sol).max()
    with pytest.raises(PsiError):
        s.congrad_non_hydro(T(forc), x, mode="mg", criterion="relres")
    s.close()


def test_arguments_checked():
    b, col = basins.nonhydro_box(2, 12, 10, 3, False, 60.0, 300.0, 1.0)
    mW = O.mask_w(col.maskT)
    bad = mW.copy()
    bad[1, 4, 4] = 0.5
    with pytest.raises(PsiError):
        NonHydrostaticSolver.from_column(b, col, bad, device=DEV)
    b1, col1 = basins.nonhydro_box(2, 12, 10, 1, False, 60.0, 300.0, 1.0)
    with pytest.raises(PsiError):   # one level: no defined reference operator
        NonHydrostaticSolver.from_column(b1, col1, O.mask_w(col1.maskT), device=DEV)


def test_continuity_at_256():
    """256^2 x 16: forcing bit for bit; div(u, v, w)_new / dt = forc - A p to
    rounding, and that residual small next to the forcing."""
    print("\n[nh 256]", flush=True)
    b, col = basins.nonhydro_box(9, 256, 256, 16, False, 50.0, 800.0, 1.0, n_islands=3)
    mW = O.mask_w(col.maskT)
    nh = basins.nonhydro_fields(col, 9)
    s = NonHydrostaticSolver.from_column(b, col, mW, device=DEV)
    u, v, w, dw, dw_mix, p = _state(nh)
    forc = s.zeros()
    its, conv, info = s.solve_non_hydrostatic(u, v, w, dw, dw_mix, p, 60.0, forc_out=forc,
                                              congr_epsilon=1e-12, max_iterations=20000)
    print(f"  its={its} {info['elapsed_ms']:.1f} ms", flush=True)
    assert conv
    _, _, _, forc_o = O.non_hydro_forcing(b, col, mW, 60.0, AB, nh.w[1], nh.dw[1], nh.dw[0],
                                          nh.dw_mix, nh.u[2], nh.v[2], nh.w[2])
    np.testing.assert_array_equal(N(forc), forc_o)
    cf = O.make_coeff_non_hydro(b, col, mW)
    res = forc_o - O.apply_op_3d(b, col.nz, cf, N(p[2]))
    z = np.zeros(col.shape)
    _, _, _, div_new = O.non_hydro_forcing(b, col, mW, 60.0, 0.0, N(w[2]), z, z, z, N(u[2]),
                                           N(v[2]), N(w[2]))
    m = col.maskT.copy()
    m[:, :2], m[:, -2:], m[:, :, :2], m[:, :, -2:] = 0, 0, 0, 0
    m = m != 0
    scale = np.abs(forc_o[m]).max()
    assert np.abs(div_new - res)[m].max() <= 1e-10 * scale
    assert np.abs(res[m]).max() <= 1e-5 * scale
    s.close()


@pytest.mark.parametrize("nz,cyclic", [(16, False), (10, True), (7, False)])
def test_levels_per_thread_match_one_level_per_thread(nz, cyclic, monkeypatch):
    """A p on several levels per thread (k_nh_ap_l: four when nz is a
    multiple of 4, two when even) against one level per thread (k_nh_ap,
    PSI_NH_LEVELS=1): the same coefficients, stencil and partial sums, so the
    whole solve -- iterations, its end and p -- is bit for bit the same (also
    where the reference's CG ends on its own divergence test)."""
    print(f"\n[nh levels per thread nz={nz} cyclic={cyclic}]", flush=True)
    b, col = basins.nonhydro_box(11, 130, 96, nz, cyclic, 50.0, 800.0, 1.0, n_islands=2)
    mW = O.mask_w(col.maskT)
    nh = basins.nonhydro_fields(col, 11)
    out = {}
    for levels in ("1", "auto"):
        if levels == "1":
            monkeypatch.setenv("PSI_NH_LEVELS", "1")
        else:
            monkeypatch.delenv("PSI_NH_LEVELS", raising=False)
        s = NonHydrostaticSolver.from_column(b, col, mW, device=DEV)
        u, v, w, dw, dw_mix, p = _state(nh)
        its, conv, info = s.solve_non_hydrostatic(u, v, w, dw, dw_mix, p, 60.0,
                                                  congr_epsilon=1e-12, max_iterations=20000)
        print(f"  levels={levels}: its={its} conv={conv}", flush=True)
        out[levels] = (its, N(p[2]), N(u[2]), N(w[2]), conv)
        s.close()
    assert out["1"][0] == out["auto"][0] and out["1"][4] == out["auto"][4]
    for a, b_ in zip(out["1"][1:4], out["auto"][1:4]):
        np.testing.assert_array_equal(a, b_)
