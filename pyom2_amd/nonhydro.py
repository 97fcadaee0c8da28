"""Host side of the non-hydrostatic pressure solver
(for_src/non_hydrostatic/non_hydrostatic.f90), used when enable_hydrostatic is
off (momentum.f90:167), over the C ABI of include/psisolve.h (psi_nh_*).

The method names are the reference's: make_coeff_non_hydro, apply_op_3D,
congrad_non_hydro, solve_non_hydrostatic.  Non-convergence and divergence
are reported in the return value (the reference prints a warning and
carries on)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import PsiError, PsiGrid, PsiInfo
from .streamfunction import AB_EPS_REFERENCE, Levels, SolveOptions, _dev3, _f64p


class NonHydrostaticSolver:
    """Device-resident operator of one basin (one psi_nh context): the grid
    metrics, levels (dzt, dzw and maskT/U/V) and maskW, (nz, ny+4, nx+4)."""

    def __init__(self, nx, ny, dxt, dxu, dyt, dyu, cost, cosu, lev: Levels, maskW,
                 cyclic_x=False, device="cuda", stream: torch.cuda.Stream | None = None):
        self.L = _lib.lib()
        self.nx, self.ny, self.cyclic_x = int(nx), int(ny), bool(cyclic_x)
        self.shape = (self.ny + 4, self.nx + 4)
        self.device = torch.device(device)
        self.lev = lev
        self.nz = lev.nz
        self.shape3 = (self.nz,) + self.shape
        self._metrics = [np.ascontiguousarray(a, np.float64) for a in
                         (dxt, dxu, dyt, dyu, cost, cosu)]
        if any(m.shape[0] != n for m, n in zip(self._metrics, [nx + 4] * 2 + [ny + 4] * 4)):
            raise ValueError("metric arrays need nx+4 / ny+4 entries")
        if isinstance(maskW, torch.Tensor):
            maskW = maskW.to(device=self.device, dtype=torch.float64).contiguous()
        else:
            maskW = torch.as_tensor(np.ascontiguousarray(maskW), dtype=torch.float64,
                                    device=self.device).contiguous()
        if tuple(lev.maskT.shape) != self.shape3 or tuple(maskW.shape) != self.shape3:
            raise ValueError("levels and maskW must be (nz, ny+4, nx+4)")
        self.maskW = maskW
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        g = PsiGrid(self.nx, self.ny, int(self.cyclic_x), *[_f64p(m) for m in self._metrics])
        ctx = C.c_void_p()
        with torch.cuda.device(self.device):
            rc = self.L.psi_nh_create(C.byref(ctx), C.byref(g), C.byref(lev.as_c()),
                                      C.c_void_p(maskW.data_ptr()),
                                      C.c_void_p(self.stream.cuda_stream))
        if rc != _lib.PSI_OK:
            raise PsiError(rc, self.L.psi_nh_last_error(None).decode())
        self.ctx = ctx

    @classmethod
    def from_column(cls, b, col, maskW, device="cuda", **kw):
        """From a basins.Basin and basins.Column (numpy) and maskW."""
        return cls(b.nx, b.ny, b.dxt, b.dxu, b.dyt, b.dyu, b.cost, b.cosu,
                   Levels.from_column(col, device), maskW, b.cyclic_x, device, **kw)

    def close(self):
        if getattr(self, "ctx", None):
            self.L.psi_nh_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, ok=(_lib.PSI_OK,)):
        if rc not in ok:
            raise PsiError(rc, self.L.psi_nh_last_error(self.ctx).decode())
        return rc

    def _sync_in(self):
        cur = torch.cuda.current_stream(self.device)
        if cur != self.stream:
            self.stream.wait_stream(cur)

    def zeros(self):
        return torch.zeros(self.shape3, dtype=torch.float64, device=self.device)

    @property
    def device_bytes(self):
        return int(self.L.psi_nh_device_bytes(self.ctx))

    def make_coeff_non_hydro(self):
        """The 7 non-zero coefficient blocks of make_coeff_non_hydro
        (non_hydrostatic.f90:72-141) as (7, nz, ny+4, nx+4): below, south,
        west, centre, east, north, above."""
        cf = torch.zeros((7,) + self.shape3, dtype=torch.float64, device=self.device)
        self._sync_in()
        self._check(self.L.psi_make_coeff_non_hydro(self.ctx, C.c_void_p(cf.data_ptr())))
        return cf

    def apply_op_3d(self, p: torch.Tensor, res: torch.Tensor | None = None):
        """apply_op_3D (non_hydrostatic.f90:247-270): res = A p."""
        res = self.zeros() if res is None else res
        self._sync_in()
        self._check(self.L.psi_apply_op_3d(self.ctx, _dev3(p, self.shape3),
                                           _dev3(res, self.shape3)))
        return res

    def congrad_non_hydro(self, forc: torch.Tensor, p: torch.Tensor,
                          opts: SolveOptions | None = None, **kw):
        """congrad_non_hydro (non_hydrostatic.f90:145-242) on p in place
        (first guess in, solution out).  opts.congr_epsilon and
        opts.max_iterations play congr_epsilon_non_hydro and
        congr_max_itts_non_hydro.  Returns (iterations, converged, info)."""
        opts = opts or SolveOptions(**kw)
        info = PsiInfo()
        self._sync_in()
        rc = self.L.psi_congrad_non_hydro(self.ctx, C.byref(opts.as_c()),
                                          _dev3(forc, self.shape3), _dev3(p, self.shape3),
                                          C.byref(info))
        self._check(rc, (_lib.PSI_OK, _lib.PSI_NOT_CONVERGED, _lib.PSI_DIVERGED))
        d = info.as_dict()
        return d["iterations"], bool(d["converged"]), d

    def solve_non_hydrostatic(self, u, v, w, dw, dw_mix, p_non_hydro, dt_mom, taum1=0, tau=1,
                              taup1=2, itt=1, opts: SolveOptions | None = None,
                              AB_eps=AB_EPS_REFERENCE, forc_out=None, **kw):
        """The whole solve_non_hydrostatic (non_hydrostatic.f90:6-64, serial)
        on the model state: u, v, w, dw and p_non_hydro (3, nz, ny+4, nx+4),
        dw_mix (nz, ny+4, nx+4), updated in place (w and u, v at taup1, their
        cyclic halos, p_non_hydro at taup1 and, at itt = 0, at tau and taum1).
        Time levels are 0-based.  Returns (iterations, converged, info)."""
        opts = opts or SolveOptions(**kw)
        s4 = (3,) + self.shape3
        args = [_dev3(a, s4) for a in (u, v, w, dw)]
        info = PsiInfo()
        self._sync_in()
        rc = self.L.psi_solve_non_hydrostatic(
            self.ctx, C.byref(opts.as_c()), *args, _dev3(dw_mix, self.shape3),
            _dev3(p_non_hydro, s4), taum1, tau, taup1, dt_mom, AB_eps, itt,
            _dev3(forc_out, self.shape3) if forc_out is not None else None, C.byref(info))
        self._check(rc, (_lib.PSI_OK, _lib.PSI_NOT_CONVERGED, _lib.PSI_DIVERGED))
        d = info.as_dict()
        return d["iterations"], bool(d["converged"]), d

    def kernel_times(self, reps=20):
        """Microseconds per launch of the iteration's kernels: {dir_apply (the
        direction pair), update, fold_a, fold_b, dir (k_nh_p), apply
        (k_nh_ap)} (HIP events, between solves)."""
        us = np.zeros(6)
        self._sync_in()
        self._check(self.L.psi_nh_kernel_times(self.ctx, reps, _f64p(us), 6))
        return dict(zip(("dir_apply", "update", "fold_a", "fold_b", "dir", "apply"),
                        us.tolist()))
