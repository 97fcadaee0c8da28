"""CPU-side checks of the product: the C ABI library loads and exports every
entry point include/*.h declares, the host island tracer reproduces
the reference's boundary lists, and the host-side state logic."""
import os
import re

import numpy as np
import pytest

from pyom2_amd import _lib, basins
from tests import fixtures

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(psi[a-z]*_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == names


def test_library_is_built_for_gfx950():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


@pytest.mark.parametrize("name", fixtures.NAMES)
def test_island_trace_matches_reference(name):
    from pyom2_amd.streamfunction import islands_trace
    b, d = fixtures.load(name)
    isl = islands_trace(b.kbot, b.cyclic_x)
    assert isl.nisle == int(d["nisle"])
    for k in ("offsets", "bi", "bj", "di", "dj"):
        np.testing.assert_array_equal(getattr(isl, k), d[k])


def test_island_trace_has_no_fixed_limits():
    """A 600x500 map with 40 small islands exceeds none of the reference's
    hard limits here, but would overflow its 10000-entry flood-fill queue."""
    from pyom2_amd.streamfunction import islands_trace
    nx, ny = 600, 500
    ocean = np.ones((ny, nx), bool)
    ocean[0, :] = ocean[-1, :] = ocean[:, 0] = ocean[:, -1] = False
    rng = np.random.default_rng(0)
    n_isl = basins._ellipses(rng, ocean, 40, 3.0, 8.0, margin=4)
    grid = basins.make_grid(nx, ny, 10e3, 10e3)
    b = basins._assemble("many", 0, nx, ny, grid, np.full((ny, nx), 4000.0), ocean, rng, 4, 2,
                         1e-10)
    isl = islands_trace(b.kbot)
    assert isl.nisle == n_isl + 1
    # the basin rim walk covers the whole rim
    assert isl.offsets[1] >= 2 * (nx - 2 + ny - 2)


def test_time_level_rotation():
    from pyom2_amd.streamfunction import AB_EPS_REFERENCE, BarotropicState
    s = BarotropicState(None, None, np.zeros((1, 3)), None, np.zeros((1, 1)), 1.0)
    seen = []
    for _ in range(4):
        seen.append((s.taum1, s.tau, s.taup1))
        s.time_goes_by()
    assert seen == [(0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 1, 2)]
    # the Fortran default-real literal 0.1 widened to double
    assert AB_EPS_REFERENCE == float(np.float32(0.1)) and AB_EPS_REFERENCE != 0.1


def test_basin_generators():
    b = basins.beta_plane_basin(seed=0)
    assert (b.nx, b.ny) == (128, 128)
    m = basins.psi_ocean_mask(b.maskT)
    assert abs((b.forc * m).sum()) < 1e-12 * np.abs(b.forc).sum()
    assert np.all(b.hur[b.hu > 0] == 1.0 / b.hu[b.hu > 0])
    s = basins.spherical_basin(seed=1, n=256)
    assert s.n_islands_planned == 2
    assert np.all(s.cost > 0) and abs(s.cost[2 + 128] - 1.0) < 1e-3
    ens = basins.ensemble_basins(seed=0, count=6)
    for e in ens:
        assert 16 <= e.nx <= 1024 and 16 <= e.ny <= 1024


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("lower", [0.005, 0.02, 0.1])
def test_poly_preconditioner_is_the_chebyshev_polynomial(degree, lower):
    """The scalars of the polynomial preconditioner, run through their
    recurrence on a scalar eigenvalue l of Z A, give q(l) whose residual
    polynomial p(l) = 1 - l q(l) is the scaled Chebyshev polynomial
    T_{k+1}((b + a - 2l) / (b - a)) / T_{k+1}((b + a) / (b - a)) on [0, b],
    a = 2 lower, b = 2: p(0) = 1, |p| <= 1 / T_{k+1}(theta) on [a, b] with
    equality at the k+2 extrema, and 0 < l q(l) < 2 on (0, b]."""
    import ctypes as C
    L = _lib.lib()
    out = (C.c_float * (1 + 2 * degree))()
    assert L.psi_poly_coefficients(degree, lower, out) == 0
    out = np.array(out, np.float64)
    a, b = 2 * lower, 2.0

    def q(lam):
        v = 1.0                      # Z r for an eigenvector of Z A
        d = v * out[0]
        z, res = d, v - lam * d
        for j in range(1, degree + 1):
            d = out[2 * j - 1] * d + out[2 * j] * res
            z += d
            res = res - lam * d
        return z

    lam = np.linspace(0.0, b, 4001)[1:]
    p = 1 - lam * np.array([q(x) for x in lam])
    theta = (b + a) / (b - a)
    cheb = np.polynomial.chebyshev.Chebyshev.basis(degree + 1)
    expect = cheb((b + a - 2 * lam) / (b - a)) / cheb(theta)
    np.testing.assert_allclose(p, expect, atol=2e-5)
    inside = lam >= a
    assert np.abs(p[inside]).max() <= 1 / cheb(theta) * (1 + 1e-4)
    assert ((1 - p) > 0).all() and ((1 - p) < 2).all()
    assert L.psi_poly_coefficients(7, lower, out.ctypes.data_as(C.POINTER(C.c_float))) == 5
