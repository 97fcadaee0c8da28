"""Host side of the MI355X streamfunction solver, mirroring pyOM2's interface.

The reference reaches the solver from its Python host as
``pyOM.time_step -> fortran.momentum() -> solve_streamfunction`` (py_src/pyOM.py:151,
for_src/main/momentum.f90:157-159) after ``fortran.streamfunction_init()``
at set-up (py_src/pyOM.py:91).  This module offers the same operations over
the C ABI of libpsisolve.so (include/psisolve.h) on device tensors:

* ``islands_trace``            -- island part of ``streamfunction_init``
* ``StreamfunctionSolver``     -- the per-basin solver state (the reference's
  module variables and SAVE'd coefficient cache) with methods named after
  the reference routines: ``congrad_streamfunction``, ``apply_op``,
  ``make_coeff_streamfunction``, ``make_inv_sfc``, ``dot_sfc``,
  ``absmax_sfc``, ``setcyclic_xy``, ``line_integral``,
  ``streamfunction_init``, ``solve_streamfunction``
* ``BarotropicState``          -- psi / dpsi time levels, dpsin, and
  ``time_goes_by`` (py_src/pyOM.py:199-207)
* ``Levels``, ``MomentumState`` -- vertical grid / 3-D masks and the 3-D model
  state for the whole routine (``streamfunction_forcing``,
  ``velocity_update``, ``solve_streamfunction_3d``)

Fields are float64 tensors of shape (ny+4, nx+4) on the GPU: element
``a(i, j)`` of pyOM2's Fortran array sits at ``[j+1, i+1]``.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import torch

from . import _lib
from ._lib import (PsiError, PsiFields, PsiGrid, PsiInfo, PsiIslands, PsiLevels, PsiMomentum,
                   PsiOpts)

# Fortran default-real literals widened to real*8 (main_module.f90:18-19, 52).
AB_EPS_REFERENCE = float(np.float32(0.1))
GRAV_REFERENCE = float(np.float32(9.81))
RHO_0_REFERENCE = 1024.0

MODES = {"plain": _lib.PSI_MODE_PLAIN, "mixed": _lib.PSI_MODE_MIXED, "mg": _lib.PSI_MODE_MG,
         "learned": _lib.PSI_MODE_LEARNED, "poly": _lib.PSI_MODE_POLY}
CRITERIA = {"reference": _lib.PSI_CRIT_REFERENCE, "relres": _lib.PSI_CRIT_RELRES}


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclasses.dataclass
class Islands:
    """Boundary lists of all land masses (streamfct_init.f90:87-93, 258)."""
    nisle: int
    offsets: np.ndarray
    bi: np.ndarray
    bj: np.ndarray
    di: np.ndarray
    dj: np.ndarray

    def isle(self, k):
        s, e = int(self.offsets[k]), int(self.offsets[k + 1])
        return self.bi[s:e], self.bj[s:e], self.di[s:e], self.dj[s:e]

    def as_c(self):
        self._keep = [np.ascontiguousarray(a, np.int32) for a in
                      (self.offsets, self.bi, self.bj, self.di, self.dj)]
        o, bi, bj, di, dj = self._keep
        return PsiIslands(self.nisle, _i32p(o), _i32p(bi), _i32p(bj), _i32p(di), _i32p(dj))


def islands_trace(kbot: np.ndarray, cyclic_x: bool = False) -> Islands:
    """Label land masses of kbot ((ny+4, nx+4), 0 = land) and walk their
    perimeters (streamfct_init.f90:35-267, island_module.f90:7-273).  Runs
    on the host, once per basin, as in the reference."""
    L = _lib.lib()
    kbot = np.ascontiguousarray(kbot, np.int32)
    ny, nx = kbot.shape[0] - 4, kbot.shape[1] - 4
    nisle = C.c_int32()
    npts = C.c_int32()
    cap_i, cap_p = 16, 4 * (nx + ny + 8)
    for _ in range(3):
        offs = np.zeros(cap_i, np.int32)
        bi, bj, di, dj = (np.zeros(cap_p, np.int32) for _ in range(4))
        rc = L.psi_islands_trace(nx, ny, int(cyclic_x), _i32p(kbot), C.byref(nisle), _i32p(offs),
                                 cap_i, _i32p(bi), _i32p(bj), _i32p(di), _i32p(dj), cap_p,
                                 C.byref(npts))
        if rc == _lib.PSI_OK:
            n, m = nisle.value, npts.value
            return Islands(n, offs[:n + 1].copy(), bi[:m].copy(), bj[:m].copy(), di[:m].copy(),
                           dj[:m].copy())
        if rc != _lib.PSI_BAD_ARG:
            raise PsiError(rc, "island perimeter walk failed")
        cap_i, cap_p = nisle.value + 1, max(npts.value, 1)
    raise PsiError(_lib.PSI_BAD_ARG, "island trace: buffer sizing failed")


@dataclasses.dataclass
class SolveOptions:
    """Solver options (congr_epsilon / congr_max_iterations of
    main_module.f90:189-190 plus the path selection)."""
    mode: str = "plain"
    criterion: str = "reference"
    congr_epsilon: float = 1e-12
    max_iterations: int = 1000
    relres_target: float = 1e-10
    check_every: int = 0
    inner_reduction: float = 0.0
    max_outer: int = 0
    poly_degree: int = 0
    poly_lower: float = 0.0
    time_kernels: bool = False
    mg_coarse_degree: int = 0

    def as_c(self):
        o = PsiOpts()
        o.mode = MODES[self.mode]
        o.criterion = CRITERIA[self.criterion]
        o.congr_epsilon = self.congr_epsilon
        o.max_iterations = self.max_iterations
        o.relres_target = self.relres_target
        o.check_every = self.check_every
        o.inner_reduction = self.inner_reduction
        o.max_outer = self.max_outer
        o.poly_degree = self.poly_degree
        o.poly_lower = self.poly_lower
        o.time_kernels = int(self.time_kernels)
        o.mg_coarse_degree = self.mg_coarse_degree
        return o


def _dev(t: torch.Tensor, shape=None):
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError("expected a contiguous float64 CUDA tensor")
    if shape is not None and tuple(t.shape[-2:]) != tuple(shape):
        raise ValueError(f"expected shape {shape}, got {tuple(t.shape)}")
    return C.c_void_p(t.data_ptr())


class StreamfunctionSolver:
    """Device-resident solver state for one basin (one C-ABI context)."""

    def __init__(self, nx, ny, dxt, dxu, dyt, dyu, cost, cosu, hur, hvr, maskU=None,
                 maskV=None, islands: Islands | None = None, cyclic_x=False,
                 device="cuda", stream: torch.cuda.Stream | None = None):
        self.L = _lib.lib()
        self.nx, self.ny, self.cyclic_x = int(nx), int(ny), bool(cyclic_x)
        self.shape = (self.ny + 4, self.nx + 4)
        self.device = torch.device(device)
        self._metrics = [np.ascontiguousarray(a, np.float64) for a in
                         (dxt, dxu, dyt, dyu, cost, cosu)]
        if any(m.shape[0] != n for m, n in zip(self._metrics, [nx + 4] * 2 + [ny + 4] * 4)):
            raise ValueError("metric arrays need nx+4 / ny+4 entries")
        self.metrics = dict(zip(("dxt", "dxu", "dyt", "dyu", "cost", "cosu"), self._metrics))
        t = lambda a: (torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64)
                       .to(self.device).contiguous() if a is not None else None)
        self.hur, self.hvr = t(hur), t(hvr)
        self.maskU, self.maskV = t(maskU), t(maskV)
        self.islands = islands
        self.nisle = islands.nisle if islands is not None else 0
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        g = PsiGrid(self.nx, self.ny, int(self.cyclic_x), *[_f64p(m) for m in self._metrics])
        f = PsiFields(self.hur.data_ptr(), self.hvr.data_ptr(),
                      self.maskU.data_ptr() if self.maskU is not None else None,
                      self.maskV.data_ptr() if self.maskV is not None else None)
        isl = islands.as_c() if islands is not None else None
        ctx = C.c_void_p()
        with torch.cuda.device(self.device):
            rc = self.L.psi_ctx_create(C.byref(ctx), C.byref(g), C.byref(f),
                                       C.byref(isl) if isl is not None else None,
                                       C.c_void_p(self.stream.cuda_stream))
        if rc != _lib.PSI_OK:
            raise PsiError(rc, self.L.psi_ctx_last_error(None).decode())
        self.ctx = ctx
        self._mg_ready = False

    @classmethod
    def from_basin(cls, b, islands: Islands | None = None, device="cuda", **kw):
        if islands is None:
            islands = islands_trace(b.kbot, b.cyclic_x)
        return cls(b.nx, b.ny, b.dxt, b.dxu, b.dyt, b.dyu, b.cost, b.cosu, b.hur, b.hvr,
                   b.maskU, b.maskV, islands, b.cyclic_x, device, **kw)

    def close(self):
        if getattr(self, "ctx", None):
            self.L.psi_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, ok=(_lib.PSI_OK,)):
        if rc not in ok:
            raise PsiError(rc, self.L.psi_ctx_last_error(self.ctx).decode())
        return rc

    def _sync_in(self):
        # order the caller's torch work on its current stream before ours
        cur = torch.cuda.current_stream(self.device)
        if cur != self.stream:
            self.stream.wait_stream(cur)

    def zeros(self):
        return torch.zeros(self.shape, dtype=torch.float64, device=self.device)

    # -- the solve ----------------------------------------------------------
    def congrad_streamfunction(self, forc: torch.Tensor, sol: torch.Tensor,
                               opts: SolveOptions | None = None, **kw):
        """congrad_streamfunction(forc, iterations, sol, converged)
        (solve_streamfct.f90:153-327); sol is updated in place.  Returns
        (iterations, converged, info dict).  Non-convergence is reported,
        not raised (the reference prints and continues); NaN raises
        (the reference halts)."""
        opts = self._prepare(opts or SolveOptions(**kw))
        info = PsiInfo()
        self._sync_in()
        rc = self.L.psi_congrad(self.ctx, C.byref(opts.as_c()), _dev(forc, self.shape),
                                _dev(sol, self.shape), C.byref(info))
        self._check(rc, (_lib.PSI_OK, _lib.PSI_NOT_CONVERGED, _lib.PSI_DIVERGED))
        d = info.as_dict()
        return d["iterations"], bool(d["converged"]), d

    def _prepare(self, opts: "SolveOptions"):
        # the multigrid path needs its hierarchy (built once per operator)
        if opts.mode == "mg" and not self._mg_ready:
            self.mg_setup()
        return opts

    def mg_setup(self):
        """Build the multigrid hierarchy (needed by mode "mg"; done on first use)."""
        self._sync_in()
        self._check(self.L.psi_mg_setup(self.ctx))
        self._mg_ready = True

    def mg_levels(self):
        """[(nx, ny, active cells)] of the multigrid levels, coarsest last."""
        cap = 32
        n = np.zeros(1, np.int32)
        nx, ny, na = (np.zeros(cap, np.int32) for _ in range(3))
        self._check(self.L.psi_mg_levels(self.ctx, _i32p(n), _i32p(nx), _i32p(ny), _i32p(na),
                                         cap))
        return [(int(nx[k]), int(ny[k]), int(na[k])) for k in range(min(int(n[0]), cap))]

    def mg_kernel_times(self, degree=0, reps=20):
        """Microseconds per launch of the multigrid path's heaviest kernels:
        {post_level0, pre_level0, fcg_a, fcg_b} (HIP events, between solves)."""
        us = np.zeros(4)
        self._sync_in()
        self._check(self.L.psi_mg_kernel_times(self.ctx, degree, reps, _f64p(us), 4))
        return dict(zip(("post_level0", "pre_level0", "fcg_a", "fcg_b"), us.tolist()))

    # -- building blocks -------------------------------------------------------
    def apply_op(self, p: torch.Tensor, res: torch.Tensor | None = None):
        """apply_op (solve_pressure.f90:254-274): res = A p."""
        res = self.zeros() if res is None else res
        self._sync_in()
        self._check(self.L.psi_apply_op(self.ctx, _dev(p, self.shape), _dev(res, self.shape)))
        return res

    def make_coeff_streamfunction(self):
        """make_coeff_streamfunction (solve_streamfct.f90:123-148): cf
        planes, shape (9, ny+4, nx+4) with plane (ii+1)+3*(jj+1)."""
        cf = torch.zeros((9,) + self.shape, dtype=torch.float64, device=self.device)
        self._check(self.L.psi_make_coeff(self.ctx, C.c_void_p(cf.data_ptr())))
        return cf

    def make_inv_sfc(self):
        """make_inv_sfc (solve_streamfct.f90:381-422): Z."""
        Z = self.zeros()
        self._check(self.L.psi_make_inv(self.ctx, C.c_void_p(Z.data_ptr())))
        return Z

    def dot_sfc(self, a, b):
        out = C.c_double()
        self._sync_in()
        self._check(self.L.psi_dot(self.ctx, _dev(a, self.shape), _dev(b, self.shape),
                                   C.byref(out)))
        return out.value

    def absmax_sfc(self, a):
        out = C.c_double()
        self._sync_in()
        self._check(self.L.psi_absmax(self.ctx, _dev(a, self.shape), C.byref(out)))
        return out.value

    def setcyclic_xy(self, a):
        self._sync_in()
        self._check(self.L.psi_setcyclic(self.ctx, _dev(a, self.shape)))
        return a

    def line_integral(self, isle: int, uloc, vloc):
        """line_integral (streamfct_init.f90:311-377), isle 1-based."""
        out = C.c_double()
        self._sync_in()
        self._check(self.L.psi_line_integral(self.ctx, int(isle), _dev(uloc, self.shape),
                                             _dev(vloc, self.shape), C.byref(out)))
        return out.value

    def psi_velocity(self, psi):
        fpx, fpy = self.zeros(), self.zeros()
        self._sync_in()
        self._check(self.L.psi_psi_velocity(self.ctx, _dev(psi, self.shape),
                                            _dev(fpx, self.shape), _dev(fpy, self.shape)))
        return fpx, fpy

    def true_relres(self, forc, sol):
        out = C.c_double()
        self._sync_in()
        self._check(self.L.psi_true_relres(self.ctx, _dev(forc, self.shape),
                                           _dev(sol, self.shape), C.byref(out)))
        return out.value

    # -- island machinery ------------------------------------------------------
    def streamfunction_init(self, opts: SolveOptions | None = None, **kw):
        """Solver part of streamfunction_init (streamfct_init.f90:269-303).
        Returns (psin (nisle, ny+4, nx+4) tensor, line_psin (nisle, nisle)
        numpy [n, isle], iterations list)."""
        opts = self._prepare(opts or SolveOptions(**kw))
        n = self.nisle
        psin = torch.zeros((max(n, 1),) + self.shape, dtype=torch.float64, device=self.device)
        line_psin = np.zeros((max(n, 1), max(n, 1)))
        its = np.zeros(max(n, 1), np.int32)
        self._sync_in()
        self._check(self.L.psi_streamfunction_init(self.ctx, C.byref(opts.as_c()),
                                                   C.c_void_p(psin.data_ptr()), _f64p(line_psin),
                                                   _i32p(its)))
        return psin[:n], line_psin[:n, :n], list(its[:n])

    def solve_streamfunction(self, state: "BarotropicState", fpx, fpy,
                             opts: SolveOptions | None = None, forc_out=None, **kw):
        """The 2-D body of solve_streamfunction (solve_streamfct.f90:39-94):
        curl forcing, warm start, solve, island constraint, AB step of psi."""
        opts = self._prepare(opts or SolveOptions(**kw))
        info = PsiInfo()
        dpsin = np.ascontiguousarray(state.dpsin)
        lp = np.ascontiguousarray(state.line_psin, np.float64) if self.nisle > 1 else np.zeros(1)
        psin_ptr = C.c_void_p(state.psin.data_ptr()) if state.psin is not None else None
        self._sync_in()
        rc = self.L.psi_solve_streamfunction(
            self.ctx, C.byref(opts.as_c()), _dev(fpx, self.shape), _dev(fpy, self.shape),
            C.c_void_p(state.psi.data_ptr()), C.c_void_p(state.dpsi.data_ptr()), _f64p(dpsin),
            psin_ptr, _f64p(lp), state.taum1, state.tau, state.taup1, state.dt_mom,
            state.AB_eps, _dev(forc_out, self.shape) if forc_out is not None else None,
            C.byref(info))
        self._check(rc, (_lib.PSI_OK, _lib.PSI_NOT_CONVERGED, _lib.PSI_DIVERGED))
        state.dpsin[:] = dpsin
        d = info.as_dict()
        return d["iterations"], bool(d["converged"]), d


    # -- the 3-D parts of solve_streamfunction --------------------------------
    def _levels(self, lev: "Levels"):
        if lev.maskT.shape[-2:] != self.shape:
            raise ValueError("levels do not match the basin")
        return C.byref(lev.as_c())

    def streamfunction_forcing(self, lev: "Levels", rho, du, dv, du_mix, dv_mix, p_hydro, fpx,
                               fpy, grav=GRAV_REFERENCE, rho_0=RHO_0_REFERENCE):
        """Hydrostatic pressure, its gradient into du/dv (in place) and the
        depth-integrated forcing fpx, fpy (solve_streamfct.f90:22-48).  rho,
        du, dv are the tau levels, (nz, ny+4, nx+4) each."""
        sh3 = (lev.nz,) + self.shape
        args = [_dev3(a, sh3) for a in (rho, du, dv, du_mix, dv_mix, p_hydro)]
        self._sync_in()
        self._check(self.L.psi_streamfunction_forcing(
            self.ctx, self._levels(lev), grav, rho_0, *args, _dev(fpx, self.shape),
            _dev(fpy, self.shape)))

    def velocity_update(self, lev: "Levels", dt_mom, u_tau, v_tau, du_tau, dv_tau, du_taum1,
                        dv_taum1, du_mix, dv_mix, psi_taup1, u_taup1, v_taup1,
                        AB_eps=AB_EPS_REFERENCE):
        """u, v at taup1 (solve_streamfct.f90:95-118 without open boundaries)."""
        sh3 = (lev.nz,) + self.shape
        a3 = [_dev3(a, sh3) for a in (u_tau, v_tau, du_tau, dv_tau, du_taum1, dv_taum1, du_mix,
                                      dv_mix)]
        self._sync_in()
        self._check(self.L.psi_velocity_update(
            self.ctx, self._levels(lev), dt_mom, AB_eps, *a3, _dev(psi_taup1, self.shape),
            _dev3(u_taup1, sh3), _dev3(v_taup1, sh3)))

    def solve_streamfunction_3d(self, state: "BarotropicState", lev: "Levels",
                                mom: "MomentumState", opts: SolveOptions | None = None,
                                grav=GRAV_REFERENCE, rho_0=RHO_0_REFERENCE, forc_out=None, **kw):
        """The whole solve_streamfunction (solve_streamfct.f90:10-120, serial,
        no open boundaries) on the model state: p_hydro, du/dv(tau) updated,
        psi/dpsi(taup1), dpsin(tau), u/v(taup1).  Returns
        (iterations, converged, info) like solve_streamfunction."""
        opts = self._prepare(opts or SolveOptions(**kw))
        if mom.nz != lev.nz:
            raise ValueError("momentum state and levels differ in nz")
        info = PsiInfo()
        dpsin = np.ascontiguousarray(state.dpsin)
        lp = np.ascontiguousarray(state.line_psin, np.float64) if self.nisle > 1 else np.zeros(1)
        psin_ptr = C.c_void_p(state.psin.data_ptr()) if state.psin is not None else None
        self._sync_in()
        rc = self.L.psi_solve_streamfunction_3d(
            self.ctx, C.byref(opts.as_c()), self._levels(lev), C.byref(mom.as_c(self.shape)),
            grav, rho_0, C.c_void_p(state.psi.data_ptr()), C.c_void_p(state.dpsi.data_ptr()),
            _f64p(dpsin), psin_ptr, _f64p(lp), state.taum1, state.tau, state.taup1,
            state.dt_mom, state.AB_eps,
            _dev(forc_out, self.shape) if forc_out is not None else None, C.byref(info))
        self._check(rc, (_lib.PSI_OK, _lib.PSI_NOT_CONVERGED, _lib.PSI_DIVERGED))
        state.dpsin[:] = dpsin
        d = info.as_dict()
        return d["iterations"], bool(d["converged"]), d


def _dev3(t: torch.Tensor, shape):
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError("expected a contiguous float64 CUDA tensor")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return C.c_void_p(t.data_ptr())


class Levels:
    """Vertical grid and 3-D masks (main_module.f90:96-98, 108): dzt, dzw on
    the host, maskT/U/V device tensors (nz, ny+4, nx+4)."""

    def __init__(self, dzt, dzw, maskT, maskU, maskV, device="cuda"):
        self.dzt = np.ascontiguousarray(dzt, np.float64)
        self.dzw = np.ascontiguousarray(dzw, np.float64)
        self.nz = int(self.dzt.shape[0])
        def t(a):
            if isinstance(a, torch.Tensor):
                return a.to(device=device, dtype=torch.float64).contiguous()
            return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                                   device=device).contiguous()
        self.maskT, self.maskU, self.maskV = t(maskT), t(maskU), t(maskV)
        if self.dzw.shape[0] != self.nz or any(m.shape[0] != self.nz for m in
                                               (self.maskT, self.maskU, self.maskV)):
            raise ValueError("dzt, dzw and the masks need nz levels")

    @classmethod
    def from_column(cls, col, device="cuda"):
        return cls(col.dzt, col.dzw, col.maskT, col.maskU, col.maskV, device)

    def as_c(self):
        return PsiLevels(self.nz, 0, self.dzt.ctypes.data, self.dzw.ctypes.data,
                         self.maskT.data_ptr(), self.maskU.data_ptr(), self.maskV.data_ptr())


@dataclasses.dataclass
class MomentumState:
    """The 3-D model state solve_streamfunction reads and writes
    (main_module.f90:127, 143-149, 154) as device tensors: rho, u, v, du, dv
    with 3 time levels (3, nz, ny+4, nx+4); du_mix, dv_mix, p_hydro (nz, ny+4, nx+4)."""
    rho: torch.Tensor
    u: torch.Tensor
    v: torch.Tensor
    du: torch.Tensor
    dv: torch.Tensor
    du_mix: torch.Tensor
    dv_mix: torch.Tensor
    p_hydro: torch.Tensor

    @property
    def nz(self):
        return int(self.du_mix.shape[0])

    @classmethod
    def from_numpy(cls, m, device="cuda"):
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                                      device=device).contiguous()
        return cls(t(m.rho), t(m.u), t(m.v), t(m.du), t(m.dv), t(m.du_mix), t(m.dv_mix),
                   torch.zeros(m.du_mix.shape, dtype=torch.float64, device=device))

    def as_c(self, shape):
        sh3 = (self.nz,) + tuple(shape)
        for a in (self.rho, self.u, self.v, self.du, self.dv):
            _dev3(a, (3,) + sh3)
        for a in (self.du_mix, self.dv_mix, self.p_hydro):
            _dev3(a, sh3)
        return PsiMomentum(*[a.data_ptr() for a in (self.rho, self.u, self.v, self.du, self.dv,
                                                    self.du_mix, self.dv_mix, self.p_hydro)])


@dataclasses.dataclass
class BarotropicState:
    """psi/dpsi time levels and island amplitudes (main_module.f90:155-159);
    level indices are 0-based (the reference's taum1=1, tau=2, taup1=3)."""
    psi: torch.Tensor            # (3, ny+4, nx+4)
    dpsi: torch.Tensor           # (3, ny+4, nx+4)
    dpsin: np.ndarray            # (nisle, 3)
    psin: torch.Tensor | None    # (nisle, ny+4, nx+4)
    line_psin: np.ndarray        # (nisle, nisle) [n, isle]
    dt_mom: float
    AB_eps: float = AB_EPS_REFERENCE
    taum1: int = 0
    tau: int = 1
    taup1: int = 2

    @classmethod
    def zeros(cls, solver: StreamfunctionSolver, dt_mom: float, psin=None, line_psin=None):
        n = max(solver.nisle, 1)
        sh = (3,) + solver.shape
        return cls(torch.zeros(sh, dtype=torch.float64, device=solver.device),
                   torch.zeros(sh, dtype=torch.float64, device=solver.device),
                   np.zeros((n, 3)), psin,
                   line_psin if line_psin is not None else np.zeros((n, n)), dt_mom)

    def time_goes_by(self):
        """Shift time levels by one (py_src/pyOM.py:199-207)."""
        self.taum1, self.tau, self.taup1 = self.tau, self.taup1, self.taum1
