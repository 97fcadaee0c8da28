"""Host side of the surface-pressure solver (for_src/external/solve_pressure.f90),
pyOM2's external mode when enable_streamfunction is off (momentum.f90:158-166),
over the C ABI of include/psisolve.h (psi_press_*).

The method names are the reference's: make_coeff_surf_press, apply_op,
congrad_surf_press, solve_pressure.  Non-convergence and divergence are
reported in the return value (the reference prints a warning and carries on);
a NaN estimate raises (the reference halts)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import PsiError, PsiGrid, PsiInfo, PsiPressFields
from .streamfunction import (AB_EPS_REFERENCE, GRAV_REFERENCE, RHO_0_REFERENCE, Levels,
                             MomentumState, SolveOptions, _dev, _f64p)


class SurfacePressureSolver:
    """Device-resident operator of one basin (one psi_press context).  hu, hv
    and maskT (= maskT(:,:,nz)) are (ny+4, nx+4) arrays; free_surface, grav and
    dt_mom enter the coefficients, which are built once, as the reference's
    SAVE'd cf is at its first call (solve_pressure.f90:155-160)."""

    def __init__(self, nx, ny, dxt, dxu, dyt, dyu, cost, cosu, hu, hv, maskT,
                 free_surface=True, dt_mom=3600.0, grav=GRAV_REFERENCE, cyclic_x=False,
                 device="cuda", stream: torch.cuda.Stream | None = None):
        self.L = _lib.lib()
        self.nx, self.ny, self.cyclic_x = int(nx), int(ny), bool(cyclic_x)
        self.shape = (self.ny + 4, self.nx + 4)
        self.device = torch.device(device)
        self.free_surface, self.dt_mom, self.grav = bool(free_surface), float(dt_mom), float(grav)
        self._metrics = [np.ascontiguousarray(a, np.float64) for a in
                         (dxt, dxu, dyt, dyu, cost, cosu)]
        if any(m.shape[0] != n for m, n in zip(self._metrics, [nx + 4] * 2 + [ny + 4] * 4)):
            raise ValueError("metric arrays need nx+4 / ny+4 entries")

        def t(a):
            if isinstance(a, torch.Tensor):
                a = a.to(device=self.device, dtype=torch.float64).contiguous()
            else:
                a = torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                                    device=self.device).contiguous()
            if tuple(a.shape) != self.shape:
                raise ValueError(f"expected shape {self.shape}, got {tuple(a.shape)}")
            return a
        self.hu, self.hv, self.maskT = t(hu), t(hv), t(maskT)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        g = PsiGrid(self.nx, self.ny, int(self.cyclic_x), *[_f64p(m) for m in self._metrics])
        f = PsiPressFields(self.hu.data_ptr(), self.hv.data_ptr(), self.maskT.data_ptr(),
                           int(self.free_surface), 0, self.grav, self.dt_mom)
        ctx = C.c_void_p()
        with torch.cuda.device(self.device):
            rc = self.L.psi_press_create(C.byref(ctx), C.byref(g), C.byref(f),
                                         C.c_void_p(self.stream.cuda_stream))
        if rc != _lib.PSI_OK:
            raise PsiError(rc, self.L.psi_press_last_error(None).decode())
        self.ctx = ctx

    @classmethod
    def from_basin(cls, b, maskT_top, free_surface=True, dt_mom=3600.0, **kw):
        """From a basins.Basin whose hu/hv hold the depths at u/v points."""
        return cls(b.nx, b.ny, b.dxt, b.dxu, b.dyt, b.dyu, b.cost, b.cosu, b.hu, b.hv,
                   maskT_top, free_surface, dt_mom, cyclic_x=b.cyclic_x, **kw)

    def close(self):
        if getattr(self, "ctx", None):
            self.L.psi_press_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, ok=(_lib.PSI_OK,)):
        if rc not in ok:
            raise PsiError(rc, self.L.psi_press_last_error(self.ctx).decode())
        return rc

    def _sync_in(self):
        cur = torch.cuda.current_stream(self.device)
        if cur != self.stream:
            self.stream.wait_stream(cur)

    def zeros(self):
        return torch.zeros(self.shape, dtype=torch.float64, device=self.device)

    @property
    def device_bytes(self):
        return int(self.L.psi_press_device_bytes(self.ctx))

    # -- building blocks -----------------------------------------------------
    def make_coeff_surf_press(self):
        """cf(:,:,3,3) of make_coeff_surf_press (solve_pressure.f90:87-152) as
        a (9, ny+4, nx+4) tensor, plane (ii+1) + 3 (jj+1)."""
        cf = torch.zeros((9,) + self.shape, dtype=torch.float64, device=self.device)
        self._sync_in()
        self._check(self.L.psi_make_coeff_surf_press(self.ctx, C.c_void_p(cf.data_ptr())))
        return cf

    def apply_op(self, p: torch.Tensor, res: torch.Tensor | None = None):
        """apply_op (solve_pressure.f90:254-274): res = A p."""
        res = self.zeros() if res is None else res
        self._sync_in()
        self._check(self.L.psi_press_apply_op(self.ctx, _dev(p, self.shape),
                                              _dev(res, self.shape)))
        return res

    def mg_setup(self):
        """Build the multigrid hierarchy of mode "mg" (free surface only; done
        on first use otherwise)."""
        self._sync_in()
        self._check(self.L.psi_press_mg_setup(self.ctx))

    # -- the solve ------------------------------------------------------------
    def congrad_surf_press(self, forc: torch.Tensor, psi: torch.Tensor,
                           opts: SolveOptions | None = None, **kw):
        """congrad_surf_press (solve_pressure.f90:157-249) on psi in place
        (first guess in, solution out).  mode "plain" with criterion
        "reference" runs the reference's algorithm; mode "mg" with criterion
        "relres" the multigrid-preconditioned solve to relres_target (free
        surface only).  Returns (iterations, converged, info)."""
        opts = opts or SolveOptions(**kw)
        info = PsiInfo()
        self._sync_in()
        rc = self.L.psi_congrad_surf_press(self.ctx, C.byref(opts.as_c()),
                                           _dev(forc, self.shape), _dev(psi, self.shape),
                                           C.byref(info))
        self._check(rc, (_lib.PSI_OK, _lib.PSI_NOT_CONVERGED, _lib.PSI_DIVERGED))
        d = info.as_dict()
        return d["iterations"], bool(d["converged"]), d

    def solve_pressure(self, psi: torch.Tensor, lev: Levels, mom: MomentumState,
                       taum1=0, tau=1, taup1=2, opts: SolveOptions | None = None,
                       dt_mom=None, AB_eps=AB_EPS_REFERENCE, grav=None, rho_0=RHO_0_REFERENCE,
                       forc_out=None, **kw):
        """The whole solve_pressure (solve_pressure.f90:9-84, serial) on the
        model state: p_hydro, du/dv(tau), u/v(taup1) and psi(taup1) (psi: the
        surface pressure, (3, ny+4, nx+4)).  Time levels are 0-based.
        Returns (iterations, converged, info)."""
        opts = opts or SolveOptions(**kw)
        if mom.nz != lev.nz:
            raise ValueError("momentum state and levels differ in nz")
        if lev.maskT.shape[-2:] != self.shape:
            raise ValueError("levels do not match the basin")
        if not (psi.is_cuda and psi.dtype == torch.float64 and psi.is_contiguous()
                and tuple(psi.shape) == (3,) + self.shape):
            raise ValueError("psi must be a contiguous float64 CUDA tensor (3, ny+4, nx+4)")
        info = PsiInfo()
        self._sync_in()
        rc = self.L.psi_solve_pressure(
            self.ctx, C.byref(opts.as_c()), C.byref(lev.as_c()), C.byref(mom.as_c(self.shape)),
            self.grav if grav is None else grav, rho_0, C.c_void_p(psi.data_ptr()), taum1, tau,
            taup1, self.dt_mom if dt_mom is None else dt_mom, AB_eps,
            _dev(forc_out, self.shape) if forc_out is not None else None, C.byref(info))
        self._check(rc, (_lib.PSI_OK, _lib.PSI_NOT_CONVERGED, _lib.PSI_DIVERGED))
        d = info.as_dict()
        return d["iterations"], bool(d["converged"]), d

    def mg_kernel_times(self, degree=0, reps=20):
        """Microseconds per launch of the multigrid mode's heaviest kernels:
        {post_level0, pre_level0, fcg_a, fcg_b} (HIP events, between solves)."""
        us = np.zeros(4)
        self._sync_in()
        self._check(self.L.psi_press_mg_kernel_times(self.ctx, degree, reps, _f64p(us), 4))
        return dict(zip(("post_level0", "pre_level0", "fcg_a", "fcg_b"), us.tolist()))

    def kernel_times(self, reps=20):
        """Microseconds per launch of the iteration's kernels:
        {dir_apply, update, fold_a, fold_b} (HIP events, between solves)."""
        us = np.zeros(4)
        self._sync_in()
        self._check(self.L.psi_press_kernel_times(self.ctx, reps, _f64p(us), 4))
        return dict(zip(("dir_apply", "update", "fold_a", "fold_b"), us.tolist()))
