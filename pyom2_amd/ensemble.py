"""Batched ensembles of independent streamfunction solves (include/psibatch.h).

Each member is one ``congrad_streamfunction`` solve
(for_src/external/solve_streamfct.f90:153-327) on its own basin; the whole
ensemble is advanced by one batched kernel per launch on the GPU, replayed
from a captured hipGraph.  Members run the mixed-precision path (FP32 inner
Jacobi-PCG, FP64 residual replacement) to a relative FP64 residual.

    ens = EnsembleSolver(basins)          # operators of every member
    ens.autotune()                        # pick tile shape / work split
    res = ens.solve(forcings, sols)       # per-member iterations, latency
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import torch

from . import _lib
from ._lib import PsiBatchMember, PsiBatchStats, PsiBatchTune, PsiError, PsiGrid, PsiInfo
from .streamfunction import Islands, SolveOptions, _f64p, islands_trace


@dataclasses.dataclass
class EnsembleResult:
    status: int
    infos: list            # psi_info as dicts, one per member
    latency_ms: np.ndarray  # completion time of each member from the batch start
    stats: dict

    @property
    def iterations(self):
        return np.array([i["iterations"] for i in self.infos])

    @property
    def relres_true(self):
        return np.array([i["relres_true"] for i in self.infos])


class EnsembleSolver:
    """Device-resident operators of many basins, solved together."""

    def __init__(self, basins, islands: list | None = None, device="cuda",
                 stream: torch.cuda.Stream | None = None):
        self.L = _lib.lib()
        self.device = torch.device(device)
        self.count = len(basins)
        if self.count < 1:
            raise ValueError("empty ensemble")
        self.shapes = [(b.ny + 4, b.nx + 4) for b in basins]
        # a stream of our own (default priority): the batch captures its
        # launches into a graph, which the null stream does not allow
        self.stream = stream if stream is not None else torch.cuda.Stream(self.device)
        self._keep = []
        members = (PsiBatchMember * self.count)()
        self.islands = []
        for k, b in enumerate(basins):
            isl = islands[k] if islands is not None else islands_trace(b.kbot, b.cyclic_x)
            self.islands.append(isl)
            metrics = [np.ascontiguousarray(a, np.float64) for a in
                       (b.dxt, b.dxu, b.dyt, b.dyu, b.cost, b.cosu)]
            hur = torch.as_tensor(np.ascontiguousarray(b.hur), dtype=torch.float64).to(self.device)
            hvr = torch.as_tensor(np.ascontiguousarray(b.hvr), dtype=torch.float64).to(self.device)
            self._keep += metrics + [hur, hvr]
            m = members[k]
            m.grid = PsiGrid(b.nx, b.ny, int(b.cyclic_x), *[_f64p(a) for a in metrics])
            m.hur, m.hvr = hur.data_ptr(), hvr.data_ptr()
            m.islands = isl.as_c() if isl.nisle > 0 else _lib.PsiIslands(0, None, None, None,
                                                                         None, None)
        batch = C.c_void_p()
        torch.cuda.current_stream(self.device).synchronize()
        with torch.cuda.device(self.device):
            rc = self.L.psi_batch_create(C.byref(batch), self.count, members,
                                         C.c_void_p(self.stream.cuda_stream))
        if rc != _lib.PSI_OK:
            raise PsiError(rc, self.L.psi_batch_last_error(None).decode())
        self.batch = batch
        self._keep = []     # the library copied what it needs
        self.active_points = int(self.L.psi_batch_active_points(self.batch))

    def close(self):
        if getattr(self, "batch", None):
            self.L.psi_batch_destroy(self.batch)
            self.batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, ok=(_lib.PSI_OK,)):
        if rc not in ok:
            raise PsiError(rc, self.L.psi_batch_last_error(self.batch).decode())
        return rc

    def zeros(self):
        return [torch.zeros(s, dtype=torch.float64, device=self.device) for s in self.shapes]

    # -- schedule -------------------------------------------------------------
    @property
    def tune(self) -> dict:
        t = PsiBatchTune()
        self._check(self.L.psi_batch_get_tune(self.batch, C.byref(t)))
        return t.as_dict()

    def set_tune(self, tile_x=64, tile_y=16, tiles_per_item=4, items_per_grab=2,
                 groups_per_cu=None, graph_launches=16):
        if groups_per_cu is None:
            groups_per_cu = self.L.psi_batch_max_groups_per_cu(self.batch, tile_x, tile_y)
        t = PsiBatchTune(tile_x, tile_y, tiles_per_item, items_per_grab, groups_per_cu,
                         graph_launches, 0.0)
        self._check(self.L.psi_batch_set_tune(self.batch, C.byref(t)))

    def autotune(self, launches: int = 24):
        """Time every candidate schedule on the members' operators and keep
        the fastest.  Returns (best, table of all candidates)."""
        cap = 128
        table = np.zeros((cap, 6), np.float64)
        n = C.c_int32()
        best = PsiBatchTune()
        torch.cuda.current_stream(self.device).synchronize()
        self._check(self.L.psi_batch_autotune(self.batch, launches, C.byref(best), _f64p(table),
                                              cap, C.byref(n)))
        rows = [dict(tile_x=int(r[0]), tile_y=int(r[1]), tiles_per_item=int(r[2]),
                     items_per_grab=int(r[3]), groups_per_cu=int(r[4]), ns_per_point=float(r[5]))
                for r in table[:min(n.value, cap)]]
        return best.as_dict(), rows

    # -- the solve ------------------------------------------------------------
    def solve(self, forcs, sols, opts: SolveOptions | None = None, **kw) -> EnsembleResult:
        """Solve A sol[k] = forc[k] for every member (sol[k]: first guess in,
        solution out).  Raises only on argument / HIP errors; per-member
        convergence is in the result."""
        opts = opts or SolveOptions(mode="mixed", criterion="relres", max_iterations=100000)
        for k, v in kw.items():
            setattr(opts, k, v)
        if len(forcs) != self.count or len(sols) != self.count:
            raise ValueError("one forcing and one solution per member")
        for k in range(self.count):
            for t in (forcs[k], sols[k]):
                if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
                        and tuple(t.shape) == self.shapes[k]):
                    raise ValueError(f"member {k}: expected contiguous float64 CUDA "
                                     f"{self.shapes[k]}")
        fp = (C.c_void_p * self.count)(*[t.data_ptr() for t in forcs])
        sp = (C.c_void_p * self.count)(*[t.data_ptr() for t in sols])
        infos = (PsiInfo * self.count)()
        lat = np.zeros(self.count, np.float64)
        stats = PsiBatchStats()
        cur = torch.cuda.current_stream(self.device)
        if cur != self.stream:
            self.stream.wait_stream(cur)
        o = opts.as_c()
        rc = self.L.psi_batch_congrad(self.batch, C.byref(o), fp, sp, infos, _f64p(lat),
                                      C.byref(stats))
        self._check(rc, ok=(_lib.PSI_OK, _lib.PSI_NOT_CONVERGED, _lib.PSI_DIVERGED,
                            _lib.PSI_NAN))
        return EnsembleResult(rc, [i.as_dict() for i in infos], lat, stats.as_dict())
