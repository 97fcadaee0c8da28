"""Seeded synthetic basins for the barotropic streamfunction solve.

A basin holds everything ``congrad_streamfunction`` and the island machinery
read from pyOM2's model state: grid metrics with a 2-point halo, the land map
``kbot``, inverse depths ``hur``/``hvr`` at u/v points, the surface-level
u/v masks and a vorticity forcing.  Fields are C-ordered ``(ny+4, nx+4)``
float64 arrays, byte-identical to pyOM2's Fortran ``a(-1:nx+2, -1:ny+2)``
with ``a(i, j)`` at ``[j+1, i+1]``.

Grid construction follows ``calc_grid`` (for_src/main/numerics.f90:34-168,
``u_centered_grid`` :1-28), masks and depths follow ``calc_topo``
(numerics.f90:187-261) in the continuous-depth form: the depth at a u (v)
point is the shallower of its two T cells, which is what summing
``maskU = min(maskT(i), maskT(i+1))`` over levels gives.  Constants follow
for_src/main/main_module.f90:11-13.

Workloads (sizes and distributions as the task configs state them):

* ``beta_plane_basin``  -- closed 128^2 Cartesian beta-plane basin, 20 km
  spacing, 4000 m minus 8 Gaussian seamounts, land = 1-cell wall rim + H<200 m,
  no islands (reseeded until the seamounts make no land), 16 Fourier forcing
  modes with k <= 8, amplitude N(0, 1e-10) s^-2, zero mean.
* ``spherical_basin``   -- 2048^2 (or any size) 0.05 deg spherical grid.
  dyt = dxt = 0.05 deg; the interior spans 102.4 deg of latitude centred on
  the equator (2048 * 0.05 deg; the config's "50S-50N" cannot hold 2048 rows
  at 0.05 deg, so the spacing is kept and the extent follows).  5000 m base
  minus 64 Gaussians; seamount depths are clipped at 150 m so that land
  (H < 100 m) comes only from the rim and the seeded elliptical islands
  (semi-axes 20-60 cells, i.e. 40-120 cells across).
* ``ensemble_basins``   -- uneven batch: sizes log-uniform in [64^2, 512^2]
  points, aspect ratio in [0.5, 2], 0-3 islands each, beta-plane
  construction.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
import scipy.ndimage

PI = 3.14159265358979323846264338327950588
RADIUS = 6370.0e3
DEGTOM = RADIUS / 180.0 * PI
ONX = 2


@dataclasses.dataclass
class Basin:
    nx: int
    ny: int
    cyclic_x: bool
    dxt: np.ndarray
    dxu: np.ndarray
    dyt: np.ndarray
    dyu: np.ndarray
    cost: np.ndarray
    cosu: np.ndarray
    kbot: np.ndarray          # int32 (ny+4, nx+4); 0 = land
    ht: np.ndarray            # depth at T points (0 on land)
    hu: np.ndarray
    hv: np.ndarray
    hur: np.ndarray
    hvr: np.ndarray
    maskT: np.ndarray
    maskU: np.ndarray
    maskV: np.ndarray
    forc: np.ndarray          # vorticity forcing at psi points
    name: str = ""
    seed: int = 0
    n_islands_planned: int = 0

    @property
    def shape(self):
        return (self.ny + 4, self.nx + 4)

    def area_z(self) -> np.ndarray:
        """area_z(i,j) = dxu(i)*cosu(j)*dyu(j) (numerics.f90:165)."""
        return (self.cosu * self.dyu)[:, None] * self.dxu[None, :]


def u_centered_grid(dyt: np.ndarray):
    """Restates u_centered_grid (numerics.f90:1-28): returns (dyu, yt, yu)."""
    n = dyt.shape[0]
    yu = np.zeros(n)
    yt = np.zeros(n)
    dyu = np.zeros(n)
    for i in range(1, n):
        yu[i] = yu[i - 1] + dyt[i]
    yt[0] = yu[0] - dyt[0] * 0.5
    for i in range(1, n):
        yt[i] = 2 * yu[i - 1] - yt[i - 1]
    for i in range(n - 1):
        dyu[i] = yt[i + 1] - yt[i]
    dyu[n - 1] = 2 * dyt[n - 1] - dyu[max(0, n - 2)]
    return dyu, yt, yu


def _axis(d_interior: np.ndarray, origin: float, cyclic: bool):
    """Halo-extended metric axis as in calc_grid (numerics.f90:46-110)."""
    n = d_interior.shape[0]
    d = np.empty(n + 2 * ONX)
    d[ONX:ONX + n] = d_interior
    for k in range(1, ONX + 1):
        if cyclic:
            d[ONX + n - 1 + k] = d_interior[k - 1]
            d[ONX - k] = d_interior[n - k]
        else:
            d[ONX + n - 1 + k] = d_interior[n - 1]
            d[ONX - k] = d_interior[0]
    du, t, u = u_centered_grid(d)
    t = t - u[0] + origin
    u = u - u[0] + origin
    if cyclic:
        for k in range(1, ONX + 1):
            t[ONX + n - 1 + k] = t[ONX + k - 1]
            t[ONX - k] = t[ONX + n - k]
            u[ONX + n - 1 + k] = t[ONX + k - 1]
            u[ONX - k] = u[ONX + n - k]
            du[ONX + n - 1 + k] = du[ONX + k - 1]
            du[ONX - k] = du[ONX + n - k]
    return d, du, t, u


def make_grid(nx, ny, dx, dy, *, degrees=False, x_origin=0.0, y_origin=0.0,
              cyclic_x=False):
    """Metric arrays (dxt, dxu, dyt, dyu, cost, cosu, xt, yt) with halos."""
    dxt, dxu, xt, xu = _axis(np.full(nx, float(dx)), x_origin, cyclic_x)
    dyt, dyu, yt, yu = _axis(np.full(ny, float(dy)), y_origin, False)
    if degrees:
        dxt, dxu, dyt, dyu = dxt * DEGTOM, dxu * DEGTOM, dyt * DEGTOM, dyu * DEGTOM
        cost = np.cos(yt / 180.0 * PI)
        cosu = np.cos(yu / 180.0 * PI)
    else:
        cost = np.ones(ny + 4)
        cosu = np.ones(ny + 4)
    return dxt, dxu, dyt, dyu, cost, cosu, xt, yt


def topo_fields(ht_int: np.ndarray, ocean_int: np.ndarray, cyclic_x: bool):
    """kbot, masks and depths from interior depth and ocean flags
    (calc_topo, numerics.f90:187-261, continuous-depth form)."""
    ny, nx = ht_int.shape
    kbot = np.zeros((ny + 4, nx + 4), np.int32)
    kbot[2:-2, 2:-2] = ocean_int.astype(np.int32)
    if cyclic_x:
        kbot[:, nx + 2:nx + 4] = kbot[:, 2:4]
        kbot[:, 0:2] = kbot[:, nx:nx + 2]
    maskT = (kbot != 0).astype(np.float64)
    ht = np.zeros((ny + 4, nx + 4))
    ht[2:-2, 2:-2] = np.where(ocean_int, ht_int, 0.0)
    if cyclic_x:
        ht[:, nx + 2:nx + 4] = ht[:, 2:4]
        ht[:, 0:2] = ht[:, nx:nx + 2]
    maskU = maskT.copy()
    maskU[:, :-1] = np.minimum(maskT[:, :-1], maskT[:, 1:])
    maskV = maskT.copy()
    maskV[:-1, :] = np.minimum(maskT[:-1, :], maskT[1:, :])
    hu = ht * maskU
    hu[:, :-1] = np.minimum(ht[:, :-1], ht[:, 1:]) * maskU[:, :-1]
    hv = ht * maskV
    hv[:-1, :] = np.minimum(ht[:-1, :], ht[1:, :]) * maskV[:-1, :]
    with np.errstate(divide="ignore"):
        hur = np.where(hu != 0.0, 1.0 / np.where(hu != 0.0, hu, 1.0), 0.0)
        hvr = np.where(hv != 0.0, 1.0 / np.where(hv != 0.0, hv, 1.0), 0.0)
    return kbot, ht, hu, hv, hur, hvr, maskT, maskU, maskV


def psi_ocean_mask(maskT: np.ndarray) -> np.ndarray:
    """1 where all four T cells around a psi point (i,j) are ocean."""
    m = np.zeros_like(maskT)
    m[:-1, :-1] = maskT[:-1, :-1] * maskT[:-1, 1:] * maskT[1:, :-1] * maskT[1:, 1:]
    return m


def fourier_forcing(rng, nx, ny, n_modes, kmax, amp_sigma, psimask):
    """Sum of n_modes seeded plane waves cos(2 pi (kx i/nx + ky j/ny) + phi)
    with integer |kx|,|ky| <= kmax (not both 0), amplitudes N(0, amp_sigma),
    restricted to interior ocean psi points and made zero-mean there."""
    forc = np.zeros((ny + 4, nx + 4))
    ii = np.arange(1, nx + 1, dtype=np.float64)
    jj = np.arange(1, ny + 1, dtype=np.float64)
    f = np.zeros((ny, nx))
    for _ in range(n_modes):
        while True:
            kx, ky = rng.integers(-kmax, kmax + 1, size=2)
            if kx != 0 or ky != 0:
                break
        a = rng.normal(0.0, amp_sigma)
        phi = rng.uniform(0.0, 2 * math.pi)
        ax = 2 * math.pi * kx * ii / nx
        ay = 2 * math.pi * ky * jj / ny + phi
        # cos(ax + ay) = cos ay cos ax - sin ay sin ax  (separable)
        f += a * (np.outer(np.cos(ay), np.cos(ax)) - np.outer(np.sin(ay), np.sin(ax)))
    m = psimask[2:-2, 2:-2]
    cnt = m.sum()
    if cnt > 0:
        f -= (f * m).sum() / cnt
    forc[2:-2, 2:-2] = f * m
    return forc


def _gaussians(rng, ht, xt, yt, count, amp_lo, amp_hi, r_lo, r_hi, xscale=None):
    """Subtract `count` Gaussian features from ht (interior, (ny, nx)).
    xt, yt: interior coordinates in metres (xt may vary with latitude via
    xscale(j) -> metres per x unit)."""
    ny, nx = ht.shape
    for _ in range(count):
        amp = rng.uniform(amp_lo, amp_hi)
        rad = rng.uniform(r_lo, r_hi)
        ci = rng.uniform(0, nx)
        cj = rng.uniform(0, ny)
        # bounding box of 4 radii, in cells
        dxc = xt[1] - xt[0] if xscale is None else xscale
        dyc = yt[1] - yt[0]
        wi = int(4 * rad / dxc) + 2
        wj = int(4 * rad / dyc) + 2
        i0, i1 = max(0, int(ci) - wi), min(nx, int(ci) + wi + 1)
        j0, j1 = max(0, int(cj) - wj), min(ny, int(cj) + wj + 1)
        if i0 >= i1 or j0 >= j1:
            continue
        x = (np.arange(i0, i1) - ci) * dxc
        y = (np.arange(j0, j1) - cj) * dyc
        g = np.exp(-(y[:, None] ** 2 + x[None, :] ** 2) / (2 * rad * rad))
        ht[j0:j1, i0:i1] -= amp * g


def _ellipses(rng, ocean, count, a_lo, a_hi, margin=6):
    """Carve `count` non-touching elliptical islands into the ocean map
    (interior (ny, nx) bool).  Returns the number placed."""
    ny, nx = ocean.shape
    placed = []
    tries = 0
    while len(placed) < count and tries < 1000:
        tries += 1
        a = rng.uniform(a_lo, a_hi)
        b = rng.uniform(a_lo, a_hi)
        th = rng.uniform(0, math.pi)
        R = max(a, b)
        lo_i, hi_i = R + margin + 2, nx - R - margin - 2
        lo_j, hi_j = R + margin + 2, ny - R - margin - 2
        if lo_i >= hi_i or lo_j >= hi_j:
            return len(placed)
        ci = rng.uniform(lo_i, hi_i)
        cj = rng.uniform(lo_j, hi_j)
        if any(math.hypot(ci - pi_, cj - pj_) < R + pr + margin for pi_, pj_, pr in placed):
            continue
        placed.append((ci, cj, R))
        i0, i1 = int(ci - R) - 1, int(ci + R) + 2
        j0, j1 = int(cj - R) - 1, int(cj + R) + 2
        x = np.arange(i0, i1) + 0.5 - ci
        y = np.arange(j0, j1) + 0.5 - cj
        X, Y = np.meshgrid(x, y)
        c, s = math.cos(th), math.sin(th)
        xr = c * X + s * Y
        yr = -s * X + c * Y
        inside = (xr / a) ** 2 + (yr / b) ** 2 <= 1.0
        ocean[j0:j1, i0:i1] &= ~inside
    return len(placed)


def _assemble(name, seed, nx, ny, grid, ht_int, ocean_int, rng, n_modes, kmax,
              amp_sigma, cyclic_x=False, n_isl=0):
    dxt, dxu, dyt, dyu, cost, cosu, xt, yt = grid
    kbot, ht, hu, hv, hur, hvr, maskT, maskU, maskV = topo_fields(ht_int, ocean_int, cyclic_x)
    pm = psi_ocean_mask(maskT)
    forc = fourier_forcing(rng, nx, ny, n_modes, kmax, amp_sigma, pm)
    return Basin(nx, ny, cyclic_x, dxt, dxu, dyt, dyu, cost, cosu, kbot, ht, hu, hv,
                 hur, hvr, maskT, maskU, maskV, forc, name, seed, n_isl)


def beta_plane_basin(seed: int = 0, nx: int = 128, ny: int = 128, dx: float = 20e3,
                     n_islands: int = 0, n_seamounts: int = 8, n_modes: int = 16,
                     kmax: int = 8, cyclic_x: bool = False) -> Basin:
    """Closed Cartesian beta-plane basin (f0=1e-4 s^-1 and beta=2e-11 do not
    enter the streamfunction operator).  Land: 1-cell wall rim (none across
    the x seam when cyclic), H < 200 m, optional elliptical islands."""
    rng = np.random.default_rng(seed)
    grid = make_grid(nx, ny, dx, dx, cyclic_x=cyclic_x)
    xt = np.arange(nx) * dx
    yt = np.arange(ny) * dx
    # seamounts must not create land (the basin has no islands unless asked):
    # redraw up to 20 times, then fill the shallows up to 200 m
    for _attempt in range(20):
        ht = np.full((ny, nx), 4000.0)
        _gaussians(rng, ht, xt, yt, n_seamounts, 500.0, 2500.0, 50e3, 200e3)
        if (ht >= 200.0).all():
            break
    ht = np.maximum(ht, 200.0)
    ocean = np.ones((ny, nx), bool)
    ocean[0, :] = False
    ocean[-1, :] = False
    if not cyclic_x:
        ocean[:, 0] = False
        ocean[:, -1] = False
    placed = 0
    if n_islands:
        a_lo = max(2.0, min(nx, ny) / 24)
        a_hi = max(a_lo + 1, min(nx, ny) / 10)
        placed = _ellipses(rng, ocean, n_islands, a_lo, a_hi,
                           margin=max(2, min(6, min(nx, ny) // 10)))
    return _assemble("beta_plane", seed, nx, ny, grid, ht, ocean, rng, n_modes, kmax,
                     1e-10, cyclic_x, placed)


def _land_below(ht: np.ndarray, h_min: float) -> np.ndarray:
    """Ocean map (interior, bool) of a closed basin: a 1-cell wall rim plus
    every cell shallower than h_min that is 8-connected to the rim.  Shallow
    patches not joined to the rim would each become an island with its own
    line-integral constraint; they are filled to h_min instead (ht is raised
    in place), so a basin's islands are only the ones placed on purpose."""
    land = ht < h_min
    land[0, :] = land[-1, :] = land[:, 0] = land[:, -1] = True
    lab, _ = scipy.ndimage.label(land, structure=np.ones((3, 3), bool))
    coast = lab == lab[0, 0]
    np.maximum(ht, np.where(coast, ht, h_min), out=ht)
    return ~coast


def spherical_basin(seed: int = 0, n: int = 2048, ny: int | None = None,
                    n_islands: int = 2, n_features: int = 64, n_modes: int = 16,
                    kmax: int = 32, dlon: float = 0.05) -> Basin:
    """Spherical basin on a dlon-degree grid centred on the equator: 5000 m
    base minus n_features Gaussians, land where H < 100 m joined to the wall
    rim (isolated shallows are filled to 100 m) plus n_islands seeded
    elliptical islands."""
    nx = n
    ny = n if ny is None else ny
    rng = np.random.default_rng(seed)
    y0 = -0.5 * ny * dlon - ONX * dlon     # origin at the first halo u point
    grid = make_grid(nx, ny, dlon, dlon, degrees=True, x_origin=0.0, y_origin=y0)
    dxm = dlon * DEGTOM
    xt = np.arange(nx) * dxm
    yt = np.arange(ny) * dxm
    ht = np.full((ny, nx), 5000.0)
    scale = n / 2048.0
    _gaussians(rng, ht, xt, yt, n_features, 1000.0, 4500.0,
               max(2.0, 20 * scale) * dxm, max(4.0, 120 * scale) * dxm)
    ocean = _land_below(ht, 100.0)
    placed = 0
    if n_islands:
        a_lo, a_hi = 20.0 * scale, 60.0 * scale
        a_lo, a_hi = max(a_lo, 2.0), max(a_hi, 3.0)
        placed = _ellipses(rng, ocean, n_islands, a_lo, a_hi)
    return _assemble("spherical", seed, nx, ny, grid, ht, ocean, rng, n_modes, kmax,
                     1e-10, False, placed)


def ar1_forcings(basin: Basin, steps: int, rho: float = 0.99, seed: int = 1,
                 n_modes: int = 16, kmax: int | None = None):
    """Slowly drifting forcings f_t = rho f_{t-1} + sqrt(1-rho^2) xi_t."""
    rng = np.random.default_rng(seed)
    kmax = kmax if kmax is not None else (32 if basin.nx >= 1024 else 8)
    pm = psi_ocean_mask(basin.maskT)
    f = basin.forc.copy()
    c = math.sqrt(1 - rho * rho)
    for _ in range(steps):
        xi = fourier_forcing(rng, basin.nx, basin.ny, n_modes, kmax, 1e-10, pm)
        f = rho * f + c * xi
        yield f


def ensemble_basins(seed: int = 0, count: int = 512, lo: int = 64, hi: int = 512):
    """Uneven batch: N = nx*ny log-uniform in [lo^2, hi^2], aspect in [0.5, 2],
    0-3 islands, beta-plane construction."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        npts = math.exp(rng.uniform(math.log(lo * lo), math.log(hi * hi)))
        asp = math.exp(rng.uniform(math.log(0.5), math.log(2.0)))
        nx = int(round(math.sqrt(npts * asp)))
        ny = int(round(npts / max(nx, 1)))
        nx = min(max(nx, 16), 2 * hi)
        ny = min(max(ny, 16), 2 * hi)
        nisl = int(rng.integers(0, 4))
        out.append(beta_plane_basin(seed=int(rng.integers(0, 2**31 - 1)), nx=nx, ny=ny,
                                    n_islands=nisl))
    return out


# ---------------------------------------------------------------------------
# Vertical levels around the barotropic solve
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Column:
    """Vertical grid and 3-D land masks of a basin (calc_grid / calc_topo,
    numerics.f90:132-136, 215-259).  Level k = 1..nz (index k-1 here) counts
    upwards: level nz is the surface.  3-D fields are C-ordered
    ``(nz, ny+4, nx+4)``, byte-identical to Fortran ``a(-1:nx+2, -1:ny+2, nz)``."""
    nz: int
    dzt: np.ndarray           # (nz,)
    dzw: np.ndarray           # (nz,)
    kbot: np.ndarray          # int32 (ny+4, nx+4): bottom level, 0 on land
    maskT: np.ndarray
    maskU: np.ndarray
    maskV: np.ndarray

    @property
    def shape(self):
        return self.maskT.shape


def vertical_grid(nz: int, depth: float = 5000.0, stretch: float = 1.08):
    """Layer thicknesses growing geometrically with depth (thinnest at the
    surface, level nz), summing to ``depth``; dzw from u_centered_grid as
    calc_grid does (numerics.f90:132)."""
    w = stretch ** np.arange(nz)[::-1]
    dzt = depth * w / w.sum()
    dzw, _, _ = u_centered_grid(dzt)
    return dzt, dzw


def _setcyclic(a: np.ndarray, nx: int):
    """setcyclic_xy on the last axis (parallel_none.f90:188-211)."""
    a[..., nx + 2:nx + 4] = a[..., 2:4]
    a[..., 0:2] = a[..., nx:nx + 2]


def column_levels(b: Basin, nz: int, depth: float | None = None, stretch: float = 1.08):
    """Layer thicknesses and bottom levels of basin b: each ocean column is
    wet from the level whose centre lies above its depth ht up to the surface
    (at least one level).  Returns (dzt, dzw, kbot) with kbot 0 on land."""
    depth = float(b.ht.max()) if depth is None else depth
    dzt, dzw = vertical_grid(nz, depth, stretch)
    ztop = np.concatenate([np.cumsum(dzt[::-1])[::-1][1:], [0.0]])   # depth of each level's top
    zc = ztop + 0.5 * dzt                                              # centre depth
    wet_levels = np.zeros(b.shape, np.int64)
    for z in zc:
        wet_levels += z < b.ht
    wet_levels = np.maximum(wet_levels, 1)
    kbot = np.where(b.kbot != 0, nz - wet_levels + 1, 0).astype(np.int32)
    return dzt, dzw, kbot


def column_basin(b: Basin, nz: int, depth: float | None = None, stretch: float = 1.08):
    """Levels for basin b (column_levels); masks follow calc_topo
    (numerics.f90:215-247) and the inverse depths hur = 1/sum_k maskU dzt,
    hvr likewise (:253-259).  Returns the basin with hu/hv/hur/hvr and the
    surface masks replaced by those of the levels (so the 2-D solve and the
    3-D fields agree) and the Column."""
    dzt, dzw, kbot = column_levels(b, nz, depth, stretch)
    k = np.arange(1, nz + 1)[:, None, None]
    maskT = ((kbot[None] != 0) & (kbot[None] <= k)).astype(np.float64)
    if b.cyclic_x:
        _setcyclic(maskT, b.nx)
    maskU = maskT.copy()
    maskU[:, :, :-1] = np.minimum(maskT[:, :, :-1], maskT[:, :, 1:])
    maskV = maskT.copy()
    maskV[:, :-1, :] = np.minimum(maskT[:, :-1, :], maskT[:, 1:, :])
    if b.cyclic_x:
        _setcyclic(maskU, b.nx)
        _setcyclic(maskV, b.nx)
    hu = np.zeros(b.shape)
    hv = np.zeros(b.shape)
    for kk in range(nz):
        hu = hu + maskU[kk] * dzt[kk]
        hv = hv + maskV[kk] * dzt[kk]
    with np.errstate(divide="ignore"):
        hur = np.where(hu != 0.0, 1.0 / np.where(hu != 0.0, hu, 1.0), 0.0)
        hvr = np.where(hv != 0.0, 1.0 / np.where(hv != 0.0, hv, 1.0), 0.0)
    b3 = dataclasses.replace(b, hu=hu, hv=hv, hur=hur, hvr=hvr, maskT=maskT[-1].copy(),
                             maskU=maskU[-1].copy(), maskV=maskV[-1].copy())
    return b3, Column(nz, dzt, dzw, kbot, maskT, maskU, maskV)


@dataclasses.dataclass
class Momentum:
    """The model state solve_streamfunction reads and writes
    (main_module.f90:147-158, 264-311): time levels are the leading axis
    (0..2 for the reference's 1..3)."""
    rho: np.ndarray        # (3, nz, ny+4, nx+4)
    u: np.ndarray
    v: np.ndarray
    du: np.ndarray
    dv: np.ndarray
    du_mix: np.ndarray     # (nz, ny+4, nx+4)
    dv_mix: np.ndarray
    psi: np.ndarray        # (3, ny+4, nx+4)
    dpsi: np.ndarray


def momentum_fields(col: Column, seed: int = 0) -> Momentum:
    """Seeded synthetic state of plausible magnitudes: density 1024-1028
    kg/m^3 stratified with horizontal structure, velocities ~0.1 m/s,
    tendencies ~1e-6 m/s^2, streamfunction ~1e6 m^3/s; all masked."""
    rng = np.random.default_rng(seed)
    nz, ny4, nx4 = col.shape
    mT, mU, mV = col.maskT, col.maskU, col.maskV
    strat = np.linspace(4.0, 0.0, nz)[:, None, None]
    jj, ii = np.meshgrid(np.linspace(0, 1, ny4), np.linspace(0, 1, nx4), indexing="ij")
    horiz = 0.5 * np.sin(3 * np.pi * ii) * np.cos(2 * np.pi * jj)
    rho = np.empty((3, nz, ny4, nx4))
    u = np.empty_like(rho)
    v = np.empty_like(rho)
    du = np.empty_like(rho)
    dv = np.empty_like(rho)
    for t in range(3):
        rho[t] = (1024.0 + strat + horiz[None] + 0.05 * rng.normal(size=(nz, ny4, nx4))) * mT
        u[t] = 0.1 * rng.normal(size=(nz, ny4, nx4)) * mU
        v[t] = 0.1 * rng.normal(size=(nz, ny4, nx4)) * mV
        du[t] = 1e-6 * rng.normal(size=(nz, ny4, nx4)) * mU
        dv[t] = 1e-6 * rng.normal(size=(nz, ny4, nx4)) * mV
    du_mix = 1e-7 * rng.normal(size=(nz, ny4, nx4)) * mU
    dv_mix = 1e-7 * rng.normal(size=(nz, ny4, nx4)) * mV
    psi = 1e6 * rng.normal(size=(3, ny4, nx4))
    dpsi = 1e4 * rng.normal(size=(3, ny4, nx4))
    return Momentum(rho, u, v, du, dv, du_mix, dv_mix, psi, dpsi)


def mask_w(maskT: np.ndarray) -> np.ndarray:
    """maskW of calc_topo (numerics.f90:246-249): min(maskT(k), maskT(k+1)) for
    k < nz, maskT(nz) at the top level."""
    w = np.array(maskT, dtype=np.float64, copy=True)
    w[:-1] = np.minimum(maskT[:-1], maskT[1:])
    return w


@dataclasses.dataclass
class NonHydro:
    """The state solve_non_hydrostatic reads and writes (main_module.f90:
    143-145, 216-220): u, v, w, dw and p_non_hydro with 3 time levels
    (3, nz, ny+4, nx+4), dw_mix (nz, ny+4, nx+4)."""
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    dw: np.ndarray
    dw_mix: np.ndarray
    p: np.ndarray


def nonhydro_fields(col: Column, seed: int = 0, cyclic_nx: int | None = None) -> NonHydro:
    """Seeded synthetic non-hydrostatic state: velocities ~0.1 m/s (w ~1e-4,
    zero at the surface), w tendencies ~1e-8 m/s^2, pressure ~1e-2 m^2/s^2;
    all masked.  With cyclic_nx the x-halos are made periodic, as a model
    state's are."""
    rng = np.random.default_rng(1000 + seed)
    nz, ny4, nx4 = col.shape
    mW = mask_w(col.maskT)
    sh = (3, nz, ny4, nx4)
    u = 0.1 * rng.normal(size=sh) * col.maskU
    v = 0.1 * rng.normal(size=sh) * col.maskV
    w = 1e-4 * rng.normal(size=sh) * mW
    dw = 1e-8 * rng.normal(size=sh) * mW
    dw_mix = 1e-9 * rng.normal(size=(nz, ny4, nx4)) * mW
    p = 1e-2 * rng.normal(size=sh) * col.maskT
    # w at the top face is zero (rigid lid), so the divergence forcing of the
    # closed basin sums to zero as the pressure equation needs
    w[:, -1] = 0.0
    dw[:, -1] = 0.0
    dw_mix[-1] = 0.0
    if cyclic_nx is not None:
        for a in (u, v, w, dw, dw_mix, p):
            _setcyclic(a, cyclic_nx)
    return NonHydro(u, v, w, dw, dw_mix, p)


def nonhydro_box(seed: int, nx: int, ny: int, nz: int, cyclic_x: bool = False,
                 dx: float = 100.0, depth: float = 1000.0, stretch: float = 1.0,
                 n_islands: int = 1):
    """A small-scale basin for the non-hydrostatic solver: the beta-plane
    basin's land and seamounts with the horizontal spacing scaled to dx
    metres and the depth to `depth` (so dx is comparable to the layer
    thickness, as in non-hydrostatic configurations), levels as in
    column_basin.  Returns (basin, Column)."""
    b = beta_plane_basin(seed=seed, nx=nx, ny=ny, n_islands=0 if cyclic_x else n_islands,
                         cyclic_x=cyclic_x)
    s = dx / 20e3
    b = dataclasses.replace(b, dxt=b.dxt * s, dxu=b.dxu * s, dyt=b.dyt * s, dyu=b.dyu * s,
                            ht=b.ht * (depth / b.ht.max()))
    return column_basin(b, nz, depth=depth, stretch=stretch)
