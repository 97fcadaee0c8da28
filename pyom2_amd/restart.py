"""pyOM2's restart file (for_src/main/restart.f90): read and write the records
of write_restart / read_restart, so a run on this library can restart from a
Fortran run's file and hand one back.

The file ``restart_PE_#####.dta`` is Fortran unformatted sequential: each
record is framed by 4-byte length markers, and the reference's builds write it
big-endian (``-fconvert=big-endian`` in every site_specific.mk).  Records
(restart.f90:23-50; read back at :124-165):

  1. nx, ny, nz, itt                               (int32)
  2. is, ie, js, je (the PE's array bounds)        (int32)
  3.-17. rho, Nsqr, Hd, int_drhodT, int_drhodS, dHd, temp, dtemp, salt, dsalt,
         u, du, v, dv, w: level tau then level taum1, each a(is:ie,js:je,1:nz)
 18. psi(taum1), psi(tau)
     with enable_streamfunction:
 19. dpsi(taum1), dpsi(tau)
 20. dpsi(taup1)
 21. dpsin(1:nisle, tau), dpsin(1:nisle, taum1)

followed by the records of the optional modules (eke, tke, rossmix2,
biharmonic thickness, idemix), which this module neither writes nor reads: it
writes the file a run with those modules off writes, and reads the records
above from any file.  A 3-D field in this package's layout, C-ordered
(nz, ny+4, nx+4), is byte for byte the Fortran a(-1:nx+2,-1:ny+2,1:nz).

Records over 2^31-9 bytes (a 2048^2 x 40 field pair is 2.7 GB) follow
gfortran's subrecord convention: the record is split into subrecords of at
most 2^31-9 bytes (gfortran's default -fmax-subrecord-length), each framed by markers holding its length, the leading
marker negated when another subrecord follows, the trailing one negated when
another precedes.  Files of that size were not written by the reference's
writer here, so that part is checked only against this module's own reader."""
from __future__ import annotations

import dataclasses
import os

import numpy as np

CORE_3D = ("rho", "Nsqr", "Hd", "int_drhodT", "int_drhodS", "dHd", "temp", "dtemp", "salt",
           "dsalt", "u", "du", "v", "dv", "w")
# bytes per subrecord: gfortran's default -fmax-subrecord-length, 2**31 - 9
MAX_SUB = 2147483639


def subrecord_sizes(n: int) -> list:
    """Payload bytes of each subrecord a record of n bytes is split into."""
    if n <= MAX_SUB:
        return [n]
    return [min(MAX_SUB, n - off) for off in range(0, n, MAX_SUB)]


def restart_filename(pe: int = 0) -> str:
    """'restart_PE_#####.dta' with the PE number zero-padded (restart.f90:18-19)."""
    return f"restart_PE_{pe:05d}.dta"


@dataclasses.dataclass
class RestartState:
    """The model state a restart file holds.  fields3[name] = (tau, taum1),
    each (nz, ny+4, nx+4); psi = (taum1, tau) and dpsi = (taum1, tau, taup1),
    each (ny+4, nx+4); dpsin = (tau, taum1), each (nisle,).  dpsi and dpsin are
    None when enable_streamfunction is off."""
    nx: int
    ny: int
    nz: int
    itt: int
    fields3: dict
    psi: tuple
    dpsi: tuple | None = None
    dpsin: tuple | None = None
    bounds: tuple = None     # (is, ie, js, je); the serial PE's by default

    def __post_init__(self):
        if self.bounds is None:
            self.bounds = (-1, self.nx + 2, -1, self.ny + 2)


class RestartError(RuntimeError):
    pass


# ---- records ---------------------------------------------------------------
def _write_record(f, payload: bytes, endian: str):
    n = len(payload)
    sizes = subrecord_sizes(n)
    if len(sizes) == 1:
        m = np.array([n], endian + "i4").tobytes()
        f.write(m + payload + m)
        return
    off = 0
    first = True
    for k in sizes:
        more = off + k < n
        lead = np.array([-k if more else k], endian + "i4").tobytes()
        trail = np.array([-k if not first else k], endian + "i4").tobytes()
        f.write(lead + payload[off:off + k] + trail)
        off += k
        first = False


def _read_record(f, endian: str) -> bytes:
    parts = []
    while True:
        h = f.read(4)
        if len(h) < 4:
            raise RestartError("unexpected end of the restart file")
        lead = int(np.frombuffer(h, endian + "i4")[0])
        k = abs(lead)
        data = f.read(k)
        t = f.read(4)
        if len(data) < k or len(t) < 4 or abs(int(np.frombuffer(t, endian + "i4")[0])) != k:
            raise RestartError("corrupt record markers in the restart file")
        parts.append(data)
        if lead >= 0:
            return b"".join(parts)


def _arrays(payload: bytes, endian: str, shapes):
    out, off = [], 0
    for shp in shapes:
        n = int(np.prod(shp))
        a = np.frombuffer(payload, endian + "f8", n, off).astype(np.float64).reshape(shp)
        out.append(a)
        off += 8 * n
    if off != len(payload):
        raise RestartError(f"record holds {len(payload)} bytes, expected {off}")
    return out


def _pack(arrays, endian):
    return b"".join(np.ascontiguousarray(a, np.float64).astype(endian + "f8").tobytes()
                    for a in arrays)


# ---- the file ------------------------------------------------------------------
def write_restart(path: str, st: RestartState, endian: str = ">"):
    """write_restart (restart.f90:1-92) for a run without the optional
    modules.  path: a file or a directory (then restart_PE_00000.dta in it)."""
    if os.path.isdir(path):
        path = os.path.join(path, restart_filename(0))
    sh3 = (st.nz, st.ny + 4, st.nx + 4)
    sh2 = (st.ny + 4, st.nx + 4)
    with open(path, "wb") as f:
        _write_record(f, np.array([st.nx, st.ny, st.nz, st.itt], endian + "i4").tobytes(),
                      endian)
        _write_record(f, np.array(st.bounds, endian + "i4").tobytes(), endian)
        for name in CORE_3D:
            tau, taum1 = st.fields3[name]
            if tuple(np.shape(tau)) != sh3 or tuple(np.shape(taum1)) != sh3:
                raise ValueError(f"{name}: expected {sh3}")
            _write_record(f, _pack((tau, taum1), endian), endian)
        if any(tuple(np.shape(a)) != sh2 for a in st.psi):
            raise ValueError(f"psi: expected {sh2}")
        _write_record(f, _pack(st.psi, endian), endian)
        if st.dpsi is not None:
            _write_record(f, _pack(st.dpsi[:2], endian), endian)
            _write_record(f, _pack(st.dpsi[2:], endian), endian)
            _write_record(f, _pack(st.dpsin, endian), endian)
    return path


def read_restart(path: str, nx: int, ny: int, nz: int, enable_streamfunction: bool = True,
                 nisle: int = 0, endian: str = ">",
                 bounds: tuple | None = None) -> RestartState:
    """read_restart (restart.f90:94-206): the dimensions and the PE bounds must
    match (:126-142; the reference halts, this raises RestartError).  Returns
    None when the file does not exist (the reference proceeds with initial
    conditions, :113-119)."""
    if os.path.isdir(path):
        path = os.path.join(path, restart_filename(0))
    if not os.path.exists(path):
        return None
    bounds = bounds or (-1, nx + 2, -1, ny + 2)
    sh3 = (nz, ny + 4, nx + 4)
    sh2 = (ny + 4, nx + 4)
    with open(path, "rb") as f:
        h = np.frombuffer(_read_record(f, endian), endian + "i4")
        if h.shape[0] != 4:
            raise RestartError("bad header record")
        nx_, ny_, nz_, itt = (int(x) for x in h)
        if (nx_, ny_, nz_) != (nx, ny, nz):
            raise RestartError(f"restart dimensions {nx_},{ny_},{nz_} do not match "
                               f"{nx},{ny},{nz}")
        b = tuple(int(x) for x in np.frombuffer(_read_record(f, endian), endian + "i4"))
        if b != tuple(bounds):
            raise RestartError(f"restart PE boundaries {b} do not match {tuple(bounds)}")
        fields = {}
        for name in CORE_3D:
            fields[name] = tuple(_arrays(_read_record(f, endian), endian, (sh3, sh3)))
        psi = tuple(_arrays(_read_record(f, endian), endian, (sh2, sh2)))
        dpsi = dpsin = None
        if enable_streamfunction:
            d01 = _arrays(_read_record(f, endian), endian, (sh2, sh2))
            d2 = _arrays(_read_record(f, endian), endian, (sh2,))
            dpsi = tuple(d01 + d2)
            dpsin = tuple(_arrays(_read_record(f, endian), endian, ((nisle,), (nisle,))))
    return RestartState(nx, ny, nz, itt, fields, psi, dpsi, dpsin, b)


# ---- the device state of this package -------------------------------------------
def barotropic_from_restart(st: RestartState, state, taum1: int = 0, tau: int = 1,
                            taup1: int = 2):
    """Fill a streamfunction.BarotropicState (psi, dpsi on the device,
    dpsin on the host; 0-based time levels) from a restart."""
    import torch
    if st.dpsi is None:
        raise RestartError("the restart holds no streamfunction records")
    dev = state.psi.device
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    state.psi[taum1] = T(st.psi[0])
    state.psi[tau] = T(st.psi[1])
    state.dpsi[taum1] = T(st.dpsi[0])
    state.dpsi[tau] = T(st.dpsi[1])
    state.dpsi[taup1] = T(st.dpsi[2])
    n = st.dpsin[0].shape[0]
    if n:
        state.dpsin[:n, tau] = st.dpsin[0]
        state.dpsin[:n, taum1] = st.dpsin[1]
    state.taum1, state.tau, state.taup1 = taum1, tau, taup1
    return state


def barotropic_records(state, nisle: int):
    """(psi, dpsi, dpsin) record tuples of a BarotropicState at its current
    time levels, for RestartState."""
    N = lambda t: t.detach().cpu().numpy()
    psi = (N(state.psi[state.taum1]), N(state.psi[state.tau]))
    dpsi = (N(state.dpsi[state.taum1]), N(state.dpsi[state.tau]), N(state.dpsi[state.taup1]))
    dpsin = (np.array(state.dpsin[:nisle, state.tau]), np.array(state.dpsin[:nisle, state.taum1]))
    return psi, dpsi, dpsin
