"""Python face of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads ``libpsi_oracle.so`` (the C restatement in ``psi_oracle.c``) with
ctypes and offers numpy-level wrappers with the reference's names.  Only
tests, the smoke check and the benchmark's CPU-baseline leg import this
module; the product package never does.

Every wrapper cites the reference routine it restates (paths relative to
the pyOM2 source tree).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "libpsi_oracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        I, D, V = C.c_int, C.c_double, None
        L.orc_make_coeff.argtypes = [I, I, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_make_coeff.restype = V
        L.orc_apply_op.argtypes = [I, I, _dp, _dp, _dp]
        L.orc_apply_op.restype = V
        L.orc_absmax.argtypes = [I, I, _dp]
        L.orc_absmax.restype = D
        L.orc_dot.argtypes = [I, I, _dp, _dp]
        L.orc_dot.restype = D
        L.orc_setcyclic.argtypes = [I, I, I, _dp]
        L.orc_setcyclic.restype = V
        L.orc_make_inv.argtypes = [I, I, _dp, I, _ip, _ip, _ip, _dp]
        L.orc_make_inv.restype = V
        L.orc_congrad.argtypes = [I, I, I, _dp, _dp, _dp, _dp, D, I, C.POINTER(I), C.POINTER(I),
                                  C.POINTER(D), C.c_void_p, I]
        L.orc_congrad.restype = I
        L.orc_line_integral.argtypes = [I, I, I, _ip, _ip, _ip, _ip, _dp, _dp, _dp, _dp, _dp,
                                        C.POINTER(I)]
        L.orc_line_integral.restype = D
        L.orc_psi_velocity.argtypes = [I, I, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_psi_velocity.restype = V
        L.orc_isleperim.argtypes = [I, I, I, _ip, _ip, C.c_void_p, I]
        L.orc_isleperim.restype = I
        L.orc_trace_island.argtypes = [I, I, I, _ip, I, _ip, _ip, _ip, _ip]
        L.orc_trace_island.restype = I
        L.orc_dense_solve.argtypes = [I, _dp, _dp]
        L.orc_dense_solve.restype = I
        L.orc_congrad_omp.argtypes = [I, I, I, _dp, _dp, _dp, _dp, D, I, C.POINTER(I),
                                      C.POINTER(I), C.POINTER(D)]
        L.orc_congrad_omp.restype = I
        L.orc_threads.argtypes = []
        L.orc_threads.restype = I
        _lib = L
    return _lib


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Islands:
    """Boundary lists of all land masses (streamfct_init.f90:87-93)."""
    nisle: int
    offsets: np.ndarray   # int32, nisle+1
    bi: np.ndarray        # int32 boundary point i (Fortran index)
    bj: np.ndarray
    di: np.ndarray        # int32 line direction
    dj: np.ndarray
    allmap: np.ndarray    # int32 (ny+4, nx+4) land-mass labels / perimeters

    def isle(self, k):
        s, e = self.offsets[k], self.offsets[k + 1]
        return self.bi[s:e], self.bj[s:e], self.di[s:e], self.dj[s:e]


def make_coeff(b, hur=None, hvr=None):
    """make_coeff_streamfunction (solve_streamfct.f90:123-148) -> cf[9, ny+4, nx+4]
    with plane (ii+1)+3*(jj+1)."""
    cf = np.zeros((9,) + b.shape)
    lib().orc_make_coeff(b.nx, b.ny, _c(b.hur if hur is None else hur),
                         _c(b.hvr if hvr is None else hvr), _c(b.dxt), _c(b.dxu),
                         _c(b.dyt), _c(b.dyu), _c(b.cost), _c(b.cosu), cf)
    return cf


def apply_op(nx, ny, cf, p):
    """apply_op (solve_pressure.f90:254-274)."""
    res = np.zeros_like(p)
    lib().orc_apply_op(nx, ny, _c(cf), _c(p), res)
    return res


def dot_sfc(nx, ny, a, b):
    """dot_sfc (solve_streamfct.f90:348-362)."""
    return lib().orc_dot(nx, ny, _c(a), _c(b))


def absmax_sfc(nx, ny, a):
    """absmax_sfc (solve_streamfct.f90:331-345)."""
    return lib().orc_absmax(nx, ny, _c(a))


def setcyclic(nx, ny, cyclic_x, a):
    """setcyclic_xy (parallel_none.f90:188-211), in place."""
    assert a.flags.c_contiguous and a.dtype == np.float64
    lib().orc_setcyclic(nx, ny, int(cyclic_x), a)
    return a


def make_inv(nx, ny, cf, isl: Islands):
    """make_inv_sfc (solve_streamfct.f90:381-422)."""
    Z = np.zeros(cf.shape[1:])
    lib().orc_make_inv(nx, ny, _c(cf), isl.nisle, _c(isl.offsets, np.int32),
                       _c(isl.bi, np.int32), _c(isl.bj, np.int32), Z)
    return Z


@dataclasses.dataclass
class CGResult:
    sol: np.ndarray
    iterations: int
    converged: bool
    est_error: float
    status: int
    trace: np.ndarray | None = None


def congrad(nx, ny, cyclic_x, cf, Z, forc, sol, eps=1e-12, max_iter=1000, trace_cap=0):
    """congrad_streamfunction (solve_streamfct.f90:153-327); sol is copied."""
    s = _c(sol).copy()
    it, cv, est = C.c_int(), C.c_int(), C.c_double()
    tr = np.zeros((max(trace_cap, 1), 4))
    st = lib().orc_congrad(nx, ny, int(cyclic_x), _c(cf), _c(Z), _c(forc), s, float(eps),
                           int(max_iter), C.byref(it), C.byref(cv), C.byref(est),
                           tr.ctypes.data if trace_cap else None, int(trace_cap))
    return CGResult(s, it.value, bool(cv.value), est.value, st,
                    tr[:min(trace_cap, it.value)] if trace_cap else None)


def congrad_threads(nx, ny, cyclic_x, cf, Z, forc, sol, eps=1e-12, max_iter=1000):
    """The same algorithm with the grid loops on OpenMP threads (timing only;
    its reductions group sums differently from the serial oracle)."""
    s = _c(sol).copy()
    it, cv, est = C.c_int(), C.c_int(), C.c_double()
    st = lib().orc_congrad_omp(nx, ny, int(cyclic_x), _c(cf), _c(Z), _c(forc), s, float(eps),
                               int(max_iter), C.byref(it), C.byref(cv), C.byref(est))
    return CGResult(s, it.value, bool(cv.value), est.value, st)


def threads():
    """OpenMP threads the threaded variant uses."""
    return lib().orc_threads()


def line_integral(b, isl: Islands, k, uloc, vloc):
    """line_integral (streamfct_init.f90:311-377) for island k (0-based)."""
    bi, bj, di, dj = isl.isle(k)
    bad = C.c_int()
    v = lib().orc_line_integral(b.nx, b.ny, len(bi), _c(bi, np.int32), _c(bj, np.int32),
                                _c(di, np.int32), _c(dj, np.int32), _c(b.dxu), _c(b.dyu),
                                _c(b.cost), _c(uloc), _c(vloc), C.byref(bad))
    if bad.value:
        raise ValueError("missing line_dir in line integral")
    return v


def psi_velocity(b, psi):
    """Velocity (fpx, fpy) of a streamfunction field for the island integrals
    (streamfct_init.f90:294-300; solve_streamfct.f90:68-74)."""
    fpx = np.zeros(b.shape)
    fpy = np.zeros(b.shape)
    lib().orc_psi_velocity(b.nx, b.ny, _c(psi), _c(b.maskU), _c(b.maskV), _c(b.hur),
                           _c(b.hvr), _c(b.dxt), _c(b.dyt), _c(b.cosu), fpx, fpy)
    return fpx, fpy


def isleperim(kmt, cyclic):
    """isleperim (island_module.f90:7-127) on an (jmt, imt) int map.
    Returns (map, nisle)."""
    jmt, imt = kmt.shape
    m = np.zeros((jmt, imt), np.int32)
    n = lib().orc_isleperim(imt, jmt, int(cyclic), _c(kmt, np.int32), m, None, 0)
    if n < 0:
        raise MemoryError("isleperim")
    return m, n


def trace_islands(b) -> Islands:
    """The island part of streamfunction_init (streamfct_init.f90:35-267):
    label land masses, then walk each one's perimeter."""
    kmt = np.where(b.kbot > 0, 5, 0).astype(np.int32)
    kmt[:2, :] = 0
    kmt[-2:, :] = 0
    kmt[:, :2] = 0
    kmt[:, -2:] = 0
    kmt[2:-2, 2:-2] = np.where(b.kbot[2:-2, 2:-2] > 0, 5, 0)
    nx = b.nx
    if b.cyclic_x:
        kmt[:, nx + 2:nx + 4] = kmt[:, 2:4]
        kmt[:, 0:2] = kmt[:, nx:nx + 2]
    allmap, nisle = isleperim(kmt, b.cyclic_x)
    if b.cyclic_x:
        allmap[:, nx + 2:nx + 4] = allmap[:, 2:4]
        allmap[:, 0:2] = allmap[:, nx:nx + 2]
    offs = [0]
    BI, BJ, DI, DJ = [], [], [], []
    for isle in range(1, nisle + 1):
        k1 = np.where(allmap == isle, 0, 1).astype(np.int32)
        m1, _ = isleperim(k1, b.cyclic_x)
        cap = 8 * (b.nx + 4) * (b.ny + 4) + 16
        bi = np.zeros(cap, np.int32); bj = np.zeros(cap, np.int32)
        di = np.zeros(cap, np.int32); dj = np.zeros(cap, np.int32)
        n = lib().orc_trace_island(b.nx, b.ny, int(b.cyclic_x), m1, cap, bi, bj, di, dj)
        if n < 0:
            raise RuntimeError(f"island {isle}: perimeter walk failed ({n})")
        BI.append(bi[:n]); BJ.append(bj[:n]); DI.append(di[:n]); DJ.append(dj[:n])
        offs.append(offs[-1] + n)
    cat = (lambda xs: np.concatenate(xs).astype(np.int32) if xs else np.zeros(0, np.int32))
    return Islands(nisle, np.array(offs, np.int32), cat(BI), cat(BJ), cat(DI), cat(DJ), allmap)


def dense_solve(A, b):
    """DGESV restated (LU, partial pivoting); returns (x, info)."""
    A = _c(A).copy()
    x = _c(b).copy()
    info = lib().orc_dense_solve(A.shape[0], A, x)
    return x, info


@dataclasses.dataclass
class InitResult:
    islands: Islands
    psin: np.ndarray       # (nisle, ny+4, nx+4)
    line_psin: np.ndarray  # (nisle, nisle), [n, isle]
    cf: np.ndarray
    Z: np.ndarray
    iterations: list


def streamfunction_init(b, eps=1e-12, max_iter=1000):
    """streamfunction_init (streamfct_init.f90:7-304)."""
    isl = trace_islands(b)
    cf = make_coeff(b)
    Z = make_inv(b.nx, b.ny, cf, isl)
    forc = np.zeros(b.shape)
    psin = np.zeros((isl.nisle,) + b.shape)
    its = []
    for k in range(isl.nisle):
        p = np.zeros(b.shape)
        bi, bj, _, _ = isl.isle(k)
        ok = (bi >= -1) & (bi <= b.nx + 2) & (bj >= -1) & (bj <= b.ny + 2)
        p[bj[ok] + 1, bi[ok] + 1] = 1.0
        setcyclic(b.nx, b.ny, b.cyclic_x, p)
        r = congrad(b.nx, b.ny, b.cyclic_x, cf, Z, forc, p, eps, max_iter)
        psin[k] = setcyclic(b.nx, b.ny, b.cyclic_x, r.sol)
        its.append(r.iterations)
    line_psin = np.zeros((isl.nisle, isl.nisle))
    for isle in range(isl.nisle):
        fpx, fpy = psi_velocity(b, psin[isle])
        for n in range(isl.nisle):
            line_psin[n, isle] = line_integral(b, isl, n, fpx, fpy)
    return InitResult(isl, psin, line_psin, cf, Z, its)


def true_relres(b, cf, Z, forc, sol):
    """||forc - A sol||_2 / ||forc||_2 over active points (Z != 0)."""
    r = forc - apply_op(b.nx, b.ny, cf, sol)
    m = Z != 0
    return float(np.sqrt((r[m] ** 2).sum()) / max(np.sqrt((forc[m] ** 2).sum()), 1e-300))


def _levels_lib():
    L = lib()
    if not getattr(L, "_levels_set", False):
        I, D = C.c_int, C.c_double
        L.orc_streamfunction_forcing.argtypes = [I, I, I, I, D] + [_dp] * 18
        L.orc_streamfunction_forcing.restype = None
        L.orc_velocity_update.argtypes = [I, I, I, D, D] + [_dp] * 19
        L.orc_velocity_update.restype = None
        L._levels_set = True
    return L


def streamfunction_forcing(b, col, rho, du, dv, du_mix, dv_mix, fxa):
    """Hydrostatic pressure, pressure-gradient update of du/dv and the
    depth-integrated forcing of solve_streamfunction (solve_streamfct.f90:22-48).
    rho, du, dv: tau levels (nz, ny+4, nx+4); du, dv are returned updated
    (copies).  Returns (p_hydro, du, dv, fpx, fpy)."""
    du = _c(du).copy()
    dv = _c(dv).copy()
    p = np.zeros(col.shape)
    fpx = np.zeros(b.shape)
    fpy = np.zeros(b.shape)
    _levels_lib().orc_streamfunction_forcing(
        b.nx, b.ny, col.nz, int(b.cyclic_x), fxa, _c(b.dxu), _c(b.dyu), _c(b.cost),
        _c(col.dzt), _c(col.dzw), _c(col.maskT), _c(col.maskU), _c(col.maskV), _c(b.hur),
        _c(b.hvr), _c(rho), du, dv, _c(du_mix), _c(dv_mix), p, fpx, fpy)
    return p, du, dv, fpx, fpy


def velocity_update(b, col, dt_mom, AB_eps, u_tau, v_tau, du_tau, dv_tau, du_taum1, dv_taum1,
                    du_mix, dv_mix, psi_taup1):
    """u, v at taup1: Adams-Bashforth step, vertical-mean removal and the
    barotropic mode of psi(taup1) (solve_streamfct.f90:95-118)."""
    u = np.zeros(col.shape)
    v = np.zeros(col.shape)
    _levels_lib().orc_velocity_update(
        b.nx, b.ny, col.nz, dt_mom, AB_eps, _c(b.dxt), _c(b.dyt), _c(b.cosu), _c(col.dzt),
        _c(col.maskU), _c(col.maskV), _c(b.hur), _c(b.hvr), _c(u_tau), _c(v_tau), _c(du_tau),
        _c(dv_tau), _c(du_taum1), _c(dv_taum1), _c(du_mix), _c(dv_mix), _c(psi_taup1), u, v)
    return u, v


def curl_forcing(b, fpx, fpy):
    """forc = curl of (fpx, fpy) on the interior (solve_streamfct.f90:49-53)."""
    nx, ny = b.nx, b.ny
    forc = np.zeros(b.shape)
    j = np.arange(1, ny + 1)[:, None]
    i = np.arange(1, nx + 1)[None, :]
    forc[j + 1, i + 1] = ((fpy[j + 1, i + 2] - fpy[j + 1, i + 1]) / (b.cosu[j + 1] * b.dxu[i + 1]) -
                          (b.cost[j + 2] * fpx[j + 2, i + 1] - b.cost[j + 1] * fpx[j + 1, i + 1]) /
                          (b.cosu[j + 1] * b.dyu[j + 1]))
    return forc


def island_rhs(b, init, fpx, fpy, dpsi_new):
    """Right-hand side of the island system (solve_streamfct.f90:61-78):
    line_forc - line_psi0 for every island, line_forc(k) the line integral of
    (fpx, fpy) around island k, line_psi0(k) that of the barotropic velocity of
    the new interior streamfunction tendency.  Entry 0 (island 1, the basin
    rim, whose streamfunction is held fixed) is formed but not used."""
    n = init.islands.nisle
    lf = np.array([line_integral(b, init.islands, k, fpx, fpy) for k in range(n)])
    u, v = psi_velocity(b, dpsi_new)
    l0 = np.array([line_integral(b, init.islands, k, u, v) for k in range(n)])
    return lf - l0


def island_solve(init, rhs):
    """DGESV of line_psin(2:nisle, 2:nisle) dpsin = rhs(2:nisle)
    (solve_streamfct.f90:79-88); returns dpsin(2:nisle, tau)."""
    x, info = dense_solve(init.line_psin[1:, 1:], rhs[1:])
    if info != 0:
        raise ValueError("island system singular")
    return x


def ab_psi(psi, dpsi, dpsin, psin, taum1, tau, taup1, dt, ab):
    """Adams-Bashforth step of psi with the island terms (solve_streamfct.f90:
    90-94), in the reference's order; psi[taup1] is overwritten."""
    psi[taup1] = psi[tau] + dt * ((1.5 + ab) * dpsi[taup1] - (0.5 + ab) * dpsi[tau])
    for k in range(1, dpsin.shape[0]):
        psi[taup1] = psi[taup1] + dt * ((1.5 + ab) * dpsin[k, tau] -
                                        (0.5 + ab) * dpsin[k, taum1]) * psin[k]


def solve_streamfunction_2d(b, init, fpx, fpy, psi, dpsi, dpsin, taum1, tau, taup1, dt, ab,
                            eps=1e-12, max_iter=3000, congrad_fn=None):
    """solve_streamfunction's 2-D body (solve_streamfct.f90:49-94): curl, warm
    start, congrad, island system, Adams-Bashforth step of psi.  psi, dpsi
    (3, ny+4, nx+4) and dpsin (nisle, 3) are updated in place.  Returns the
    congrad result.  congrad_fn (default: the serial congrad) may be
    congrad_threads for grids where the serial oracle is too slow."""
    nx, ny = b.nx, b.ny
    forc = curl_forcing(b, fpx, fpy)
    dpsi[taup1] = 2 * dpsi[tau] - dpsi[taum1]
    r = (congrad_fn or congrad)(nx, ny, b.cyclic_x, init.cf, init.Z, forc, dpsi[taup1], eps,
                                max_iter)
    dpsi[taup1] = setcyclic(nx, ny, b.cyclic_x, r.sol)
    if init.islands.nisle > 1:
        dpsin[1:, tau] = island_solve(init, island_rhs(b, init, fpx, fpy, dpsi[taup1]))
    ab_psi(psi, dpsi, dpsin, init.psin, taum1, tau, taup1, dt, ab)
    return r


def solve_streamfunction(b, init, col, mom, dpsin, taum1, tau, taup1, dt, ab, grav, rho_0,
                         eps=1e-12, max_iter=3000):
    """The whole solve_streamfunction (solve_streamfct.f90:10-120, serial, no
    open boundaries) on a basins.Momentum state, updated in place (p_hydro
    is returned).  Returns (p_hydro, congrad result)."""
    fxa = grav / rho_0
    p, du, dv, fpx, fpy = streamfunction_forcing(b, col, mom.rho[tau], mom.du[tau], mom.dv[tau],
                                                 mom.du_mix, mom.dv_mix, fxa)
    mom.du[tau] = du
    mom.dv[tau] = dv
    r = solve_streamfunction_2d(b, init, fpx, fpy, mom.psi, mom.dpsi, dpsin, taum1, tau, taup1,
                                dt, ab, eps, max_iter)
    u, v = velocity_update(b, col, dt, ab, mom.u[tau], mom.v[tau], mom.du[tau], mom.dv[tau],
                           mom.du[taum1], mom.dv[taum1], mom.du_mix, mom.dv_mix, mom.psi[taup1])
    mom.u[taup1] = u
    mom.v[taup1] = v
    return p, r


# ---- the surface-pressure solver (solve_pressure.f90) -------------------------
def _press_lib():
    L = lib()
    if not getattr(L, "_press_set", False):
        I, D = C.c_int, C.c_double
        L.orc_make_coeff_surf_press.argtypes = [I, I] + [_dp] * 9 + [I, D, D, _dp]
        L.orc_make_coeff_surf_press.restype = None
        L.orc_congrad_surf_press.argtypes = [I, I, I, _dp, _dp, _dp, _dp, D, I, C.POINTER(I),
                                             C.POINTER(I), C.POINTER(D)]
        L.orc_congrad_surf_press.restype = I
        L.orc_pressure_forcing.argtypes = [I, I, I, I, D, D] + [_dp] * 17 + [I, D, _dp, _dp, _dp]
        L.orc_pressure_forcing.restype = None
        L.orc_pressure_velocity.argtypes = [I, I, I, D] + [_dp] * 8
        L.orc_pressure_velocity.restype = None
        L._press_set = True
    return L


def make_coeff_surf_press(b, hu, hv, maskT_top, free_surface, grav, dt_mom):
    """make_coeff_surf_press (solve_pressure.f90:87-152): 9 planes."""
    cf = np.zeros((9,) + b.shape)
    _press_lib().orc_make_coeff_surf_press(b.nx, b.ny, _c(hu), _c(hv), _c(maskT_top), _c(b.dxt),
                                           _c(b.dxu), _c(b.dyt), _c(b.dyu), _c(b.cost),
                                           _c(b.cosu), int(free_surface), grav, dt_mom, cf)
    return cf


def congrad_surf_press(b, cf, maskT_top, forc, psi, eps=1e-12, max_iter=1000):
    """congrad_surf_press (solve_pressure.f90:157-249); psi is copied."""
    s = _c(psi).copy()
    it, cv, est = C.c_int(), C.c_int(), C.c_double()
    st = _press_lib().orc_congrad_surf_press(b.nx, b.ny, int(b.cyclic_x), _c(cf), _c(maskT_top),
                                             _c(forc), s, float(eps), int(max_iter),
                                             C.byref(it), C.byref(cv), C.byref(est))
    return CGResult(s, it.value, bool(cv.value), est.value, st, None)


def pressure_forcing(b, col, dt, ab, u_tau, v_tau, du_tau, dv_tau, du_taum1, dv_taum1, du_mix,
                     dv_mix, psi_tau, free_surface, grav):
    """AB step of u, v to taup1, depth-integrated transport, its divergence
    and the free-surface term (solve_pressure.f90:31-58).  Returns
    (u_taup1, v_taup1, forc)."""
    u = np.zeros(col.shape)
    v = np.zeros(col.shape)
    forc = np.zeros(b.shape)
    _press_lib().orc_pressure_forcing(
        b.nx, b.ny, col.nz, int(b.cyclic_x), dt, ab, _c(b.dxt), _c(b.dyt), _c(b.cost),
        _c(b.cosu), _c(col.dzt), _c(col.maskT), _c(col.maskU), _c(col.maskV), _c(u_tau),
        _c(v_tau), _c(du_tau), _c(dv_tau), _c(du_taum1), _c(dv_taum1), _c(du_mix), _c(dv_mix),
        _c(psi_tau), int(free_surface), grav, u, v, forc)
    return u, v, forc


def pressure_velocity(b, col, dt, psi_taup1, u_taup1, v_taup1):
    """u, v -= dt grad psi(taup1) on the interior (solve_pressure.f90:64-70);
    returns updated copies."""
    u = _c(u_taup1).copy()
    v = _c(v_taup1).copy()
    _press_lib().orc_pressure_velocity(b.nx, b.ny, col.nz, dt, _c(b.dxu), _c(b.dyu), _c(b.cost),
                                       _c(col.maskU), _c(col.maskV), _c(psi_taup1), u, v)
    return u, v


def solve_pressure(b, col, hu, hv, mom, taum1, tau, taup1, dt, ab, grav, rho_0, free_surface,
                   eps=1e-12, max_iter=3000, cf=None):
    """The whole solve_pressure (solve_pressure.f90:9-72) on a basins.Momentum
    state (psi holds the surface pressure), updated in place.  Returns
    (p_hydro, congrad result)."""
    p, du, dv, _, _ = streamfunction_forcing(b, col, mom.rho[tau], mom.du[tau], mom.dv[tau],
                                             mom.du_mix, mom.dv_mix, grav / rho_0)
    mom.du[tau] = du
    mom.dv[tau] = dv
    u, v, forc = pressure_forcing(b, col, dt, ab, mom.u[tau], mom.v[tau], du, dv, mom.du[taum1],
                                  mom.dv[taum1], mom.du_mix, mom.dv_mix, mom.psi[tau],
                                  free_surface, grav)
    mom.psi[taup1] = 2 * mom.psi[tau] - mom.psi[taum1]
    if cf is None:
        cf = make_coeff_surf_press(b, hu, hv, col.maskT[-1], free_surface, grav, dt)
    r = congrad_surf_press(b, cf, col.maskT[-1], forc, mom.psi[taup1], eps, max_iter)
    mom.psi[taup1] = setcyclic(b.nx, b.ny, b.cyclic_x, r.sol)
    mom.u[taup1], mom.v[taup1] = pressure_velocity(b, col, dt, mom.psi[taup1], u, v)
    return p, r


# ---- the non-hydrostatic pressure solver (non_hydrostatic.f90) ------------------
NH_PLANES = 7   # below, south, west, centre, east, north, above


def _nh_lib():
    L = lib()
    if not getattr(L, "_nh_set", False):
        I, D = C.c_int, C.c_double
        L.orc_make_coeff_non_hydro.argtypes = [I, I, I] + [_dp] * 12
        L.orc_make_coeff_non_hydro.restype = None
        L.orc_apply_op_3d.argtypes = [I, I, I, _dp, _dp, _dp]
        L.orc_apply_op_3d.restype = None
        L.orc_congrad_non_hydro.argtypes = [I, I, I, I, _dp, _dp, _dp, _dp, D, I,
                                            C.POINTER(I), C.POINTER(I), C.POINTER(D)]
        L.orc_congrad_non_hydro.restype = I
        L.orc_non_hydro_forcing.argtypes = [I, I, I, I, D, D] + [_dp] * 14
        L.orc_non_hydro_forcing.restype = None
        L.orc_non_hydro_velocity.argtypes = [I, I, I, D] + [_dp] * 11
        L.orc_non_hydro_velocity.restype = None
        L._nh_set = True
    return L


def mask_w(maskT):
    """maskW of calc_topo (numerics.f90:246-249): min(maskT(k), maskT(k+1)),
    maskT at the top level."""
    w = np.array(maskT, dtype=np.float64, copy=True)
    w[:-1] = np.minimum(maskT[:-1], maskT[1:])
    return w


def make_coeff_non_hydro(b, col, maskW):
    """make_coeff_non_hydro (non_hydrostatic.f90:72-141): (7, nz, ny+4, nx+4)."""
    cf = np.zeros((NH_PLANES,) + col.shape)
    _nh_lib().orc_make_coeff_non_hydro(b.nx, b.ny, col.nz, _c(col.maskU), _c(col.maskV),
                                       _c(maskW), _c(b.dxt), _c(b.dxu), _c(b.dyt), _c(b.dyu),
                                       _c(b.cost), _c(b.cosu), _c(col.dzt), _c(col.dzw), cf)
    return cf


def apply_op_3d(b, nz, cf, p):
    """apply_op_3D (non_hydrostatic.f90:247-270)."""
    res = np.zeros(p.shape)
    _nh_lib().orc_apply_op_3d(b.nx, b.ny, nz, _c(cf), _c(p), res)
    return res


def congrad_non_hydro(b, col, cf, forc, p, eps=1e-12, max_iter=1000):
    """congrad_non_hydro (non_hydrostatic.f90:145-242); p is copied."""
    s = _c(p).copy()
    it, cv, est = C.c_int(), C.c_int(), C.c_double()
    st = _nh_lib().orc_congrad_non_hydro(b.nx, b.ny, col.nz, int(b.cyclic_x), _c(cf),
                                         _c(col.maskT), _c(forc), s, float(eps), int(max_iter),
                                         C.byref(it), C.byref(cv), C.byref(est))
    return CGResult(s, it.value, bool(cv.value), est.value, st, None)


def non_hydro_forcing(b, col, maskW, dt, ab, w_tau, dw_tau, dw_taum1, dw_mix, u_taup1, v_taup1,
                      w_taup1):
    """w(taup1) by Adams-Bashforth for k < nz, setcyclic_xyz of u, v, w at
    taup1 and the forcing (non_hydrostatic.f90:15-39).  u, v, w are updated
    copies; returns (u, v, w, forc)."""
    u, v, w = (_c(a).copy() for a in (u_taup1, v_taup1, w_taup1))
    forc = np.zeros(col.shape)
    _nh_lib().orc_non_hydro_forcing(b.nx, b.ny, col.nz, int(b.cyclic_x), dt, ab, _c(b.dxt),
                                    _c(b.dyt), _c(b.cost), _c(b.cosu), _c(col.dzt), _c(maskW),
                                    _c(w_tau), _c(dw_tau), _c(dw_taum1), _c(dw_mix), u, v, w,
                                    forc)
    return u, v, w, forc


def non_hydro_velocity(b, col, maskW, dt, p, u, v, w):
    """The pressure gradient out of u, v, w at taup1 (non_hydrostatic.f90:53-63);
    returns updated copies."""
    u, v, w = (_c(a).copy() for a in (u, v, w))
    _nh_lib().orc_non_hydro_velocity(b.nx, b.ny, col.nz, dt, _c(b.dxu), _c(b.dyu), _c(b.cost),
                                     _c(col.dzw), _c(col.maskU), _c(col.maskV), _c(maskW),
                                     _c(p), u, v, w)
    return u, v, w


def solve_non_hydrostatic(b, col, maskW, nh, taum1, tau, taup1, dt, ab, eps=1e-12,
                          max_iter=1000, itt=1, cf=None):
    """The whole solve_non_hydrostatic (non_hydrostatic.f90:6-64) on a state
    nh with u, v, w, dw (3 time levels), dw_mix and p_non_hydro (3 time
    levels), updated in place.  Returns the congrad result."""
    u, v, w, forc = non_hydro_forcing(b, col, maskW, dt, ab, nh.w[tau], nh.dw[tau], nh.dw[taum1],
                                      nh.dw_mix, nh.u[taup1], nh.v[taup1], nh.w[taup1])
    nh.p[taup1] = 2 * nh.p[tau] - nh.p[taum1]
    if cf is None:
        cf = make_coeff_non_hydro(b, col, maskW)
    r = congrad_non_hydro(b, col, cf, forc, nh.p[taup1], eps, max_iter)
    p = r.sol
    for k in range(col.nz):
        setcyclic(b.nx, b.ny, b.cyclic_x, p[k])
    nh.p[taup1] = p
    if itt == 0:
        nh.p[tau] = p
        nh.p[taum1] = p
    nh.u[taup1], nh.v[taup1], nh.w[taup1] = non_hydro_velocity(b, col, maskW, dt, p, u, v, w)
    return r, forc
