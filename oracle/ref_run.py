"""Run the reference's own Fortran routines through oracle/_ref/ref_harness
(TEST INFRASTRUCTURE ONLY; built by ``make -C oracle ref`` from the pyOM2
sources where they lie).  Used to write golden fixtures and to pin the C
oracle.  The harness binary exists only where the reference sources were
present at build time."""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "_ref", "ref_harness")
REF_SRC = "/root/reference/for_src"


def available() -> bool:
    return os.path.exists(HARNESS)


def build() -> bool:
    if not os.path.isdir(REF_SRC):
        return available()
    subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return available()


def run(b, forc, sol, eps=1e-12, max_iter=1000, uloc=None, vloc=None):
    """Run streamfunction_init + congrad_streamfunction + apply_op +
    line_integral of the reference on basin b.  Returns a dict."""
    nx, ny = b.nx, b.ny
    uloc = np.zeros(b.shape) if uloc is None else uloc
    vloc = np.zeros(b.shape) if vloc is None else vloc
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([nx, ny, int(b.cyclic_x), max_iter], np.int32).tofile(f)
            np.array([eps], np.float64).tofile(f)
            for a in (b.dxt, b.dxu, b.dyt, b.dyu, b.cost, b.cosu):
                np.ascontiguousarray(a, np.float64).tofile(f)
            np.ascontiguousarray(b.kbot, np.int32).tofile(f)
            for a in (b.hur, b.hvr, b.maskU, b.maskV, forc, sol, uloc, vloc):
                np.ascontiguousarray(a, np.float64).tofile(f)
        p = subprocess.run([HARNESS, fin, fout], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, timeout=3600)
        if not os.path.exists(fout):
            raise RuntimeError("reference harness failed:\n" + p.stdout.decode()[-2000:])
        raw = open(fout, "rb").read()
    off = 0

    def ints(n):
        nonlocal off
        v = np.frombuffer(raw, np.int32, n, off)
        off += 4 * n
        return v.copy()

    def dbl(n):
        nonlocal off
        v = np.frombuffer(raw, np.float64, n, off)
        off += 8 * n
        return v.copy()

    nisle = int(ints(1)[0])
    nrb = ints(nisle)
    bnd = []
    for k in range(nisle):
        n = int(nrb[k])
        bi, bj = ints(n), ints(n)
        di, dj = ints(n), ints(n)
        bnd.append((bi, bj, di, dj))
    iters, conv = ints(2)
    plane = (ny + 4) * (nx + 4)
    line_psin = dbl(nisle * nisle).reshape(nisle, nisle).T   # Fortran (n, isle)
    psin = dbl(plane * nisle).reshape(nisle, ny + 4, nx + 4)
    sol_out = dbl(plane).reshape(ny + 4, nx + 4)
    res = dbl(plane).reshape(ny + 4, nx + 4)
    lines = dbl(nisle)
    return dict(nisle=nisle, nr_boundary=nrb, boundary=bnd, iterations=int(iters),
                converged=bool(conv), line_psin=line_psin, psin=psin, sol=sol_out,
                res=res, lines=lines, log=p.stdout.decode(errors="replace"))


STEP = os.path.join(HERE, "_ref", "ref_step")


def step_available() -> bool:
    return os.path.exists(STEP)


def run_step(b, col, mom, eps=1e-12, max_iter=1000, dt_mom=3600.0):
    """Run the reference's streamfunction_init + solve_streamfunction on basin
    b with levels col and state mom (basins.Column / basins.Momentum, time
    levels taum1=0, tau=1, taup1=2).  The basin must have no islands.
    Returns a dict of the outputs."""
    nx, ny, nz = b.nx, b.ny, col.nz
    f64 = lambda a: np.ascontiguousarray(a, np.float64)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([nx, ny, nz, int(b.cyclic_x), max_iter], np.int32).tofile(f)
            np.array([eps, dt_mom], np.float64).tofile(f)
            for a in (b.dxt, b.dxu, b.dyt, b.dyu, b.cost, b.cosu, col.dzt, col.dzw):
                f64(a).tofile(f)
            np.ascontiguousarray(col.kbot, np.int32).tofile(f)
            for a in (b.hur, b.hvr, col.maskT, col.maskU, col.maskV, mom.rho[1], mom.du[1],
                      mom.du[0], mom.dv[1], mom.dv[0], mom.du_mix, mom.dv_mix, mom.u[1],
                      mom.v[1], mom.psi[1], mom.dpsi[1], mom.dpsi[0]):
                f64(a).tofile(f)
        p = subprocess.run([STEP, fin, fout], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, timeout=3600)
        if not os.path.exists(fout):
            raise RuntimeError("reference step failed:\n" + p.stdout.decode()[-2000:])
        raw = open(fout, "rb").read()
    plane = (ny + 4) * (nx + 4)
    nisle, itts = np.frombuffer(raw, np.int32, 2, 0)
    off = 8
    out = {}
    for name, n in (("p_hydro", nz), ("du", nz), ("dv", nz), ("u", nz), ("v", nz),
                    ("psi", 1), ("dpsi", 1)):
        a = np.frombuffer(raw, np.float64, n * plane, off).copy()
        off += 8 * n * plane
        out[name] = a.reshape((n, ny + 4, nx + 4) if name not in ("psi", "dpsi")
                              else (ny + 4, nx + 4))
    out.update(nisle=int(nisle), iterations=int(itts), log=p.stdout.decode(errors="replace"))
    return out


PRESS = os.path.join(HERE, "_ref", "ref_press")


def press_available() -> bool:
    return os.path.exists(PRESS)


def run_press(b, col, hu, hv, mom, free_surface, eps=1e-12, max_iter=1000, dt_mom=3600.0):
    """Run the reference's solve_pressure on basin b with levels col, depths
    hu/hv at u/v points and state mom (time levels taum1=0, tau=1, taup1=2;
    psi holds the surface pressure).  Returns a dict of the outputs."""
    nx, ny, nz = b.nx, b.ny, col.nz
    f64 = lambda a: np.ascontiguousarray(a, np.float64)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([nx, ny, nz, int(b.cyclic_x), max_iter, int(free_surface)],
                     np.int32).tofile(f)
            np.array([eps, dt_mom], np.float64).tofile(f)
            for a in (b.dxt, b.dxu, b.dyt, b.dyu, b.cost, b.cosu, col.dzt, col.dzw):
                f64(a).tofile(f)
            for a in (hu, hv, col.maskT, col.maskU, col.maskV, mom.rho[1], mom.du[1], mom.du[0],
                      mom.dv[1], mom.dv[0], mom.du_mix, mom.dv_mix, mom.u[1], mom.v[1],
                      mom.psi[1], mom.psi[0]):
                f64(a).tofile(f)
        p = subprocess.run([PRESS, fin, fout], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, timeout=3600)
        if not os.path.exists(fout):
            raise RuntimeError("reference solve_pressure failed:\n" + p.stdout.decode()[-2000:])
        raw = open(fout, "rb").read()
    plane = (ny + 4) * (nx + 4)
    itts = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    off = 4
    out = {}
    for name in ("p_hydro", "du", "dv", "u", "v", "psi"):
        n = 1 if name == "psi" else nz
        a = np.frombuffer(raw, np.float64, n * plane, off).copy()
        off += 8 * n * plane
        out[name] = a.reshape((ny + 4, nx + 4) if name == "psi" else (nz, ny + 4, nx + 4))
    out.update(iterations=itts, log=p.stdout.decode(errors="replace"))
    return out


NONHYDRO = os.path.join(HERE, "_ref", "ref_nonhydro")


def nonhydro_available() -> bool:
    return os.path.exists(NONHYDRO)


def run_nonhydro(b, col, maskW, nh, eps=1e-12, max_iter=1000, dt_mom=3600.0, itt=1):
    """Run the reference's solve_non_hydrostatic on basin b with levels col,
    maskW and state nh (basins.NonHydro; time levels taum1=0, tau=1,
    taup1=2).  Returns a dict of the outputs and make_coeff_non_hydro's cf as
    (27, nz, ny+4, nx+4), plane (ii+1) + 3 (jj+1) + 9 (kk+1)."""
    nx, ny, nz = b.nx, b.ny, col.nz
    f64 = lambda a: np.ascontiguousarray(a, np.float64)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            np.array([nx, ny, nz, int(b.cyclic_x), max_iter, itt], np.int32).tofile(f)
            np.array([eps, dt_mom], np.float64).tofile(f)
            for a in (b.dxt, b.dxu, b.dyt, b.dyu, b.cost, b.cosu, col.dzt, col.dzw):
                f64(a).tofile(f)
            for a in (col.maskT, col.maskU, col.maskV, maskW, nh.u[2], nh.v[2], nh.w[1],
                      nh.w[2], nh.dw[1], nh.dw[0], nh.dw_mix, nh.p[1], nh.p[0]):
                f64(a).tofile(f)
        p = subprocess.run([NONHYDRO, fin, fout], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, timeout=3600)
        if not os.path.exists(fout):
            raise RuntimeError("reference solve_non_hydrostatic failed:\n" +
                               p.stdout.decode()[-2000:])
        raw = open(fout, "rb").read()
    n3 = nz * (ny + 4) * (nx + 4)
    itts = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    off = 4
    out = {}
    for name in ("u", "v", "w", "p", "p_tau", "p_taum1"):
        out[name] = np.frombuffer(raw, np.float64, n3, off).copy().reshape(col.shape)
        off += 8 * n3
    cf = np.frombuffer(raw, np.float64, 27 * n3, off).copy()
    # Fortran cf(i,j,k,ii,jj,kk): i fastest, kk slowest
    out["cf"] = cf.reshape(27, nz, ny + 4, nx + 4)
    out.update(iterations=itts, log=p.stdout.decode(errors="replace"))
    return out


RESTART = os.path.join(HERE, "_ref", "ref_restart")
RESTART_FIELDS = ("rho", "Nsqr", "Hd", "int_drhodT", "int_drhodS", "dHd", "temp", "dtemp",
                  "salt", "dsalt", "u", "du", "v", "dv", "w")


def restart_available() -> bool:
    return os.path.exists(RESTART)


def _restart_header(nx, ny, nz, itt, nisle, sf):
    return np.array([nx, ny, nz, itt, nisle, int(sf)], np.int32).tobytes()


def run_restart_write(nx, ny, nz, itt, fields3, psi, dpsi, dpsin, sf=True):
    """The reference's write_restart on the given state (fields3[name] =
    (tau, taum1); psi = (taum1, tau); dpsi = (taum1, tau, taup1); dpsin =
    (tau, taum1)).  Returns the bytes of restart_PE_00000.dta."""
    nisle = 0 if dpsin is None else int(np.shape(dpsin[0])[0])
    f64 = lambda a: np.ascontiguousarray(a, np.float64).tobytes()
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        with open(fin, "wb") as f:
            f.write(_restart_header(nx, ny, nz, itt, nisle, sf))
            for name in RESTART_FIELDS:
                f.write(f64(fields3[name][0]) + f64(fields3[name][1]))
            for a in tuple(psi) + tuple(dpsi):
                f.write(f64(a))
            if nisle:
                f.write(f64(dpsin[0]) + f64(dpsin[1]))
        p = subprocess.run([RESTART, "write", fin], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, timeout=600, cwd=td)
        out = os.path.join(td, "restart_PE_00000.dta")
        if not os.path.exists(out):
            raise RuntimeError("reference write_restart failed:\n" + p.stdout.decode()[-2000:])
        return open(out, "rb").read()


def run_restart_read(data, nx, ny, nz, nisle, sf=True):
    """The reference's read_restart of the file bytes `data`; returns
    (itt, fields3, psi, dpsi, dpsin) as run_restart_write takes them."""
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        open(fin, "wb").write(_restart_header(nx, ny, nz, 0, nisle, sf))
        open(os.path.join(td, "restart_PE_00000.dta"), "wb").write(data)
        p = subprocess.run([RESTART, "read", fin, fout], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, timeout=600, cwd=td)
        if not os.path.exists(fout):
            raise RuntimeError("reference read_restart failed:\n" + p.stdout.decode()[-2000:])
        raw = open(fout, "rb").read()
    n3, n2 = nz * (ny + 4) * (nx + 4), (ny + 4) * (nx + 4)
    itt = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    off = 4
    take = lambda n: np.frombuffer(raw, np.float64, n, off).copy()
    fields = {}
    for name in RESTART_FIELDS:
        a = take(n3).reshape(nz, ny + 4, nx + 4)
        off += 8 * n3
        b = take(n3).reshape(nz, ny + 4, nx + 4)
        off += 8 * n3
        fields[name] = (a, b)
    two = []
    for _ in range(5):
        two.append(take(n2).reshape(ny + 4, nx + 4))
        off += 8 * n2
    dpsin = None
    if nisle:
        dpsin = (take(nisle), np.frombuffer(raw, np.float64, nisle, off + 8 * nisle).copy())
    return itt, fields, tuple(two[:2]), tuple(two[2:]), dpsin
