"""Seeded synthetic sea-ice states for the benchmark and test workloads.

No forcing data or host ocean model is available, so every workload builds its
grid, masks, ice and forcing fields here from
``numpy.random.default_rng(SeedSequence([config, member]))``:

* grid: spherical latitude-longitude (R = 6.37e6 m, so a 4-degree cell is
  ``gridcellWidth`` = 444709.89 m wide) for the global workloads, x-cyclic;
  a uniform 25 km Cartesian grid on a polar latitude band for regional members,
  closed on every side.  Metrics are derived by :func:`derive_grid` as the
  plugin derives them;
* land mask: Gaussian-smoothed white noise (sigma 3 cells, periodic in x for
  the global grids) thresholded at its 30th percentile, halo rows in y land;
* ice: ``hIceMean ~ U(0,3)`` m, ``Area ~ U(0,1)``, ``hSnowMean ~ U(0,0.5)`` m
  where |lat| > 55 degrees, zero elsewhere and on land;
* forcing: wind stress ``N(0,0.1)`` N/m2 inverted through the bulk law
  ``tau = rhoAir*airIceDrag*|U|*U`` into ``uWind``/``vWind``; ocean velocity
  ``N(0,0.1)`` m/s; SST from -1.8 to 28 degC by latitude; ``SWdown ~ U(0,300)``;
  ``LWdown = 250``; ``ATemp`` 250-275 K by latitude; ``aqh = 2e-3``;
  ``Qnet ~ N(0,50)``; ``Qsw = -SWdown``; ``surfPress = 1e5``; ``ocSalt = 34.7``;
  ``R_low = 4000`` m on ocean; precipitation, snowfall, evaporation, runoff 0;
* velocities and stresses start at zero.
"""

from __future__ import annotations

import math

import numpy as np

from .set_inits import derive_grid
from .state import VerisState, make_state_settings
from .variables import VARIABLE_NAMES, VariablesNamespace

R_EARTH = 6.37e6
OMEGA = 7.292e-5
OL = 2  # overlap width


def rng_for(config: int, member: int = 0) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence([config, member]))


def _gauss_smooth(field: np.ndarray, sigma: float, periodic_x: bool) -> np.ndarray:
    """Gaussian filter by FFT; non-periodic axes are zero-padded by 4 sigma."""
    pad = int(math.ceil(4 * sigma))
    px = 0 if periodic_x else pad
    f = np.pad(field, ((px, px), (pad, pad)), mode="wrap")
    nx, ny = f.shape
    kx = np.fft.fftfreq(nx)[:, None]
    ky = np.fft.rfftfreq(ny)[None, :]
    g = np.exp(-2.0 * (math.pi * sigma) ** 2 * (kx**2 + ky**2))
    out = np.fft.irfft2(np.fft.rfft2(f) * g, s=f.shape)
    return out[px: nx - px, pad: ny - pad]


def fill_x(a: np.ndarray) -> np.ndarray:
    """x-periodic overlap fill (what a single-process cyclic host does)."""
    a = a.copy()
    a[-OL:] = a[OL: 2 * OL]
    a[:OL] = a[-2 * OL: -OL]
    return a


def _pad_interior(interior: np.ndarray, cyclic: bool) -> np.ndarray:
    nx, ny = interior.shape
    out = np.zeros((nx + 2 * OL, ny + 2 * OL))
    out[OL:-OL, OL:-OL] = interior
    return fill_x(out) if cyclic else out


def _land_mask(rng, nx, ny, cyclic, land_frac=0.3, sigma=3.0):
    noise = rng.standard_normal((nx, ny))
    sm = _gauss_smooth(noise, sigma, periodic_x=cyclic)
    thr = np.quantile(sm, land_frac)
    return (sm > thr).astype(np.float64)  # 1 = ocean


def _ocean_grid(nx, ny, lat_t_pad, dx_pad, dy_pad, mask_int, cyclic, cos_lat=True):
    """Host-ocean grid arrays in the layout ``set_inits`` consumes."""
    NX, NY = nx + 2 * OL, ny + 2 * OL
    maskT = _pad_interior(mask_int, cyclic)
    maskU = maskT * np.roll(maskT, 1, 0)
    maskV = maskT * np.roll(maskT, 1, 1)
    if cyclic:
        maskU = fill_x(maskU)
        maskV = fill_x(maskV)
    else:
        maskU[:OL] = 0
        maskU[-OL:] = 0
        maskV[:OL] = 0
        maskV[-OL:] = 0
    lat = np.broadcast_to(lat_t_pad[None, :], (NX, NY))
    dphi_half = 0.5 * (dy_pad / R_EARTH) * 180 / math.pi
    lat_v = lat - dphi_half[None, :]
    ct = np.cos(np.deg2rad(lat)) if cos_lat else np.ones_like(lat)
    cv = np.cos(np.deg2rad(lat_v)) if cos_lat else np.ones_like(lat)
    dxt = dx_pad.copy()
    dyt = dy_pad.copy()
    grid = dict(
        maskT=maskT,
        maskU=maskU,
        maskV=maskV,
        dxt=dxt,
        dxu=dxt.copy(),
        dyt=dyt,
        dyu=dyt.copy(),
        area_t=dxt[:, None] * ct * dyt[None, :],
        area_u=dxt[:, None] * ct * dyt[None, :],
        area_v=dxt[:, None] * cv * dyt[None, :],
        coriolis_t=2 * OMEGA * np.sin(np.deg2rad(lat)),
        ht=4000.0 * maskT,
    )
    return grid, lat


def _forcing(rng, vs, lat, sett, ice_lat=55.0, h_noise=None):
    """Ice state and forcing fields on the padded grid (in-place on ``vs``)."""
    shape = vs.iceMask.shape
    M = vs.iceMask
    polar = (np.abs(lat) > ice_lat).astype(np.float64) * M
    if h_noise is None:
        h = rng.uniform(0.0, 3.0, shape)
    else:
        h = h_noise
    vs.hIceMean = h * polar
    vs.Area = rng.uniform(0.0, 1.0, shape) * polar
    vs.hSnowMean = rng.uniform(0.0, 0.5, shape) * polar
    # wind: draw tau ~ N(0, 0.1) and invert tau = rhoAir*Cd*|U|*U
    taux = rng.normal(0.0, 0.1, shape)
    tauy = rng.normal(0.0, 0.1, shape)
    cd = sett.rhoAir * np.where(vs.fCori < 0, sett.airIceDrag_south, sett.airIceDrag)
    tmag = np.sqrt(taux**2 + tauy**2)
    umag = np.sqrt(tmag / cd)
    with np.errstate(invalid="ignore", divide="ignore"):
        scale = np.where(umag > 0, 1.0 / (cd * umag), 0.0)
    vs.uWind = taux * scale
    vs.vWind = tauy * scale
    vs.wSpeed = np.sqrt(vs.uWind**2 + vs.vWind**2)
    vs.uOcean = rng.normal(0.0, 0.1, shape) * vs.iceMaskU
    vs.vOcean = rng.normal(0.0, 0.1, shape) * vs.iceMaskV
    c2 = np.cos(np.deg2rad(lat)) ** 2
    vs.theta = (-1.8 + 29.8 * c2) + 273.15
    vs.SWdown = rng.uniform(0.0, 300.0, shape)
    vs.Qnet = rng.normal(0.0, 50.0, shape)
    vs.LWdown = np.full(shape, 250.0)
    vs.ATemp = 250.0 + 25.0 * c2
    vs.aqh = np.full(shape, 2e-3)
    vs.Qsw = -vs.SWdown
    vs.surfPress = np.full(shape, 1e5)
    vs.ssh_an = np.zeros(shape)
    vs.ocSalt = np.full(shape, 34.7)
    for name in ("precip", "snowfall", "evap", "runoff"):
        vs[name] = np.zeros(shape)
    if sett.enable_cyclic_x:
        for name in ("hIceMean", "Area", "hSnowMean", "uWind", "vWind", "wSpeed",
                     "uOcean", "vOcean", "SWdown", "Qnet", "Qsw"):
            vs[name] = fill_x(vs[name])


def _zero_rest(vs, shape):
    for name in VARIABLE_NAMES:
        if name not in vs:
            vs[name] = np.zeros(shape)


def global_latlon(config: int, nx: int, ny: int, lat_south: float, lat_north: float,
                  member: int = 0, h_sigma: float | None = None, polar_land: float | None = None,
                  **setting_overrides) -> VerisState:
    """x-cyclic spherical latitude-longitude workload (4, 1 and 0.25 degree)."""
    rng = rng_for(config, member)
    sett = make_state_settings(nx, ny, enable_cyclic_x=True, **setting_overrides)
    NX, NY = nx + 2 * OL, ny + 2 * OL
    dlat = (lat_north - lat_south) / ny
    dlon = 360.0 / nx
    lat_int = lat_south + (np.arange(ny) + 0.5) * dlat
    lat_pad = np.concatenate([np.full(OL, lat_int[0]), lat_int, np.full(OL, lat_int[-1])])
    dx = np.full(NX, R_EARTH * math.radians(dlon))
    dy = np.full(NY, R_EARTH * math.radians(dlat))
    mask = _land_mask(rng, nx, ny, cyclic=True)
    if polar_land is not None:
        mask = mask * (np.abs(lat_int)[None, :] <= polar_land)
    grid, lat = _ocean_grid(nx, ny, lat_pad, dx, dy, mask, cyclic=True)
    vs = VariablesNamespace()
    derive_grid(vs, grid)
    h_noise = None
    if h_sigma is not None:
        z = _gauss_smooth(rng.standard_normal((nx, ny)), h_sigma, periodic_x=True)
        z = (z - z.mean()) / z.std()
        h_noise = _pad_interior(np.clip(1.5 + 0.75 * z, 0.0, 3.0), cyclic=True)
    _forcing(rng, vs, lat, sett, h_noise=h_noise)
    _zero_rest(vs, (NX, NY))
    return VerisState(sett, vs)


def regional(config: int, member: int, nx: int, ny: int, dx: float = 25e3,
             lat_south: float = 60.0, lat_north: float = 85.0, h_sigma: float | None = None,
             **setting_overrides) -> VerisState:
    """Closed regional domain on a uniform Cartesian grid over a polar band.

    ``h_sigma`` (cells) replaces the uniform 0-3 m thickness by smooth noise
    of that correlation length (mean 1.5 m, clipped to 0-3 m)."""
    rng = rng_for(config, member)
    sett = make_state_settings(nx, ny, enable_cyclic_x=False, **setting_overrides)
    NX, NY = nx + 2 * OL, ny + 2 * OL
    lat_int = lat_south + (np.arange(ny) + 0.5) * (lat_north - lat_south) / ny
    lat_pad = np.concatenate([np.full(OL, lat_int[0]), lat_int, np.full(OL, lat_int[-1])])
    mask = _land_mask(rng, nx, ny, cyclic=False)
    grid, lat = _ocean_grid(nx, ny, lat_pad, np.full(NX, dx), np.full(NY, dx), mask,
                            cyclic=False, cos_lat=False)
    vs = VariablesNamespace()
    derive_grid(vs, grid)
    h_noise = None
    if h_sigma is not None:
        z = _gauss_smooth(rng.standard_normal((nx, ny)), h_sigma, periodic_x=False)
        z = (z - z.mean()) / z.std()
        h_noise = _pad_interior(np.clip(1.5 + 0.75 * z, 0.0, 3.0), cyclic=False)
    _forcing(rng, vs, lat, sett, ice_lat=0.0, h_noise=h_noise)
    _zero_rest(vs, (NX, NY))
    return VerisState(sett, vs)


# -- the named workloads ----------------------------------------------------

def config0(**kw) -> VerisState:
    """4-degree global grid, 90 x 40 cells, latitudes -80..80."""
    return global_latlon(0, 90, 40, -80.0, 80.0, **kw)


def config1(**kw) -> VerisState:
    """1-degree global grid, 360 x 160 cells, latitudes -80..80."""
    return global_latlon(1, 360, 160, -80.0, 80.0, **kw)


def config3(**kw) -> VerisState:
    """1-degree grid of the adjoint workload (120 substeps by default)."""
    kw.setdefault("nEVPsteps", 120)
    return global_latlon(3, 360, 160, -80.0, 80.0, **kw)


def config4(**kw) -> VerisState:
    """0.25-degree global grid 1440 x 720, land beyond |lat| > 88, 10-cell thickness noise."""
    kw.setdefault("nEVPsteps", 800)
    return global_latlon(4, 1440, 720, -90.0, 90.0, h_sigma=10.0, polar_land=88.0, **kw)


def config3_adjoint_inputs(state: VerisState, member: int = 0):
    """Seeded inputs of the adjoint workload on ``state``'s grid: a spatial P*
    field (lognormal, median 27500 N/m^2, log-std 0.3, 5-cell correlation)
    and a target velocity field (N(0, 0.1) m/s smoothed over 3 cells, zero on
    land and in the overlap).  Returns (pStar, uTarget, vTarget)."""
    rng = rng_for(3, 20_000 + member)
    NX, NY = state.shape
    nx, ny = NX - 2 * OL, NY - 2 * OL
    cyc = bool(state.settings.enable_cyclic_x)

    def smooth_unit(sigma):
        z = _gauss_smooth(rng.standard_normal((nx, ny)), sigma, periodic_x=cyc)
        return (z - z.mean()) / z.std()

    pstar = np.pad(27500.0 * np.exp(0.3 * smooth_unit(5.0)), OL, mode="edge")
    pstar = fill_x(pstar) if cyc else pstar
    vs = state.variables
    ut = np.zeros((NX, NY))
    vt = np.zeros((NX, NY))
    ut[OL:-OL, OL:-OL] = 0.1 * smooth_unit(3.0)
    vt[OL:-OL, OL:-OL] = 0.1 * smooth_unit(3.0)
    return pstar, ut * np.asarray(vs.maskInU), vt * np.asarray(vs.maskInV)


def config4_adjoint_member(member: int, **kw) -> VerisState:
    """Member of the batched adjoint at 0.25 degree: 240 x 240 closed regional
    domain with 10-cell thickness noise."""
    kw.setdefault("nEVPsteps", 120)
    kw.setdefault("h_sigma", 10.0)
    return regional(4, 100 + member, 240, 240, **kw)


def config2_sizes(n_members: int = 48):
    """Member sizes (nx, ny) of the ragged ensemble, uniform in 64..512 x 64..384."""
    rng = rng_for(2, 10_000)
    nx = rng.integers(64, 513, n_members)
    ny = rng.integers(64, 385, n_members)
    return [(int(a), int(b)) for a, b in zip(nx, ny)]


def config2_member(member: int, size=None, **kw) -> VerisState:
    nx, ny = size if size is not None else config2_sizes()[member]
    return regional(2, member, nx, ny, **kw)


def tiny(nx: int = 8, ny: int = 6, cyclic: bool = True, config: int = 99, member: int = 0,
         **kw) -> VerisState:
    """Small state for brute-force checks (ice everywhere, 30% land)."""
    if cyclic:
        return global_latlon(config, nx, ny, 50.0, 80.0, member=member, **kw)
    return regional(config, member, nx, ny, **kw)
