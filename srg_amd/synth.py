"""Synthetic Sentinel-1-shaped bursts for tests and benchmarks.

A burst is built the way the reference pipeline sees one:

* an orbit given as state vectors every 10 s (like the 11-vector
  ``.orbtiming`` files), produced from an exact quadratic trajectory at 700 km
  altitude and 7.6 km/s with a small seeded curvature, so the 4-point Hermite
  interpolation of the reference reproduces it to rounding;
* per-pulse satellite positions from that Hermite interpolation (as
  sentinel/sentineltimingsub.f90:59-65 writes them), pulse times at the PRF;
* the range window ``rngstart = R_c - S/2 * dmrg``, ``rngend = rngstart +
  (S-1) dmrg`` (sentinel/backprojectgpusub.f90:222-224);
* a lat/lon DEM grid (north-up, ``deltalat < 0``) centred on the ground point
  seen at slant range ``R_c`` broadside to the right of the track at mid-burst;
  heights flat or seeded fractional-Brownian-motion terrain (int16 metres);
* range-compressed samples: circular Gaussian clutter band-limited along range
  by the chirp response, plus point-target echoes
  ``a * sinc((i - r)/res) * exp(-j 4 pi R / lambda)``.

Nothing here runs the backprojection; data synthesis for large bursts runs on
the GPU through the library's synth kernels, small ones in numpy.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import tdbp

C_LIGHT = 299792458.0
WVL_S1 = 0.05546576           # m, C band (the reference's literal)
FS_IW = 64.345e6              # Hz, IW range sampling rate
PRF_IW = 1717.0               # Hz
BW_IW = 56.5e6                # Hz, IW1 chirp bandwidth (sinc resolution fs/B)
WGS84_A = 6378137.0
WGS84_E2 = 0.0066943799901499996
RE_APERTURE = 6378137.0       # sentinel/backprojectgpusub.f90:149
HT_APERTURE = 700000.0        # sentinel/backprojectgpusub.f90:150


# --------------------------------------------------------------- geodesy


def llh_to_ecef(lat_deg, lon_deg, h):
    lat = np.radians(lat_deg)
    lon = np.radians(lon_deg)
    re = WGS84_A / np.sqrt(1.0 - WGS84_E2 * np.sin(lat) ** 2)
    return np.stack([(re + h) * np.cos(lat) * np.cos(lon),
                     (re + h) * np.cos(lat) * np.sin(lon),
                     (re * (1.0 - WGS84_E2) + h) * np.sin(lat)], axis=-1)


def ecef_to_llh(x):
    """Geodetic latitude/longitude (deg) and height by fixed-point iteration."""
    x = np.asarray(x, np.float64)
    p = np.hypot(x[..., 0], x[..., 1])
    lon = np.arctan2(x[..., 1], x[..., 0])
    lat = np.arctan2(x[..., 2], p * (1.0 - WGS84_E2))
    for _ in range(8):
        re = WGS84_A / np.sqrt(1.0 - WGS84_E2 * np.sin(lat) ** 2)
        h = p / np.cos(lat) - re
        lat = np.arctan2(x[..., 2], p * (1.0 - WGS84_E2 * re / (re + h)))
    re = WGS84_A / np.sqrt(1.0 - WGS84_E2 * np.sin(lat) ** 2)
    h = p / np.cos(lat) - re
    return np.degrees(lat), np.degrees(lon), h


def enu_basis(lat_deg, lon_deg):
    lat, lon = math.radians(lat_deg), math.radians(lon_deg)
    east = np.array([-math.sin(lon), math.cos(lon), 0.0])
    north = np.array([-math.sin(lat) * math.cos(lon), -math.sin(lat) * math.sin(lon),
                      math.cos(lat)])
    up = np.array([math.cos(lat) * math.cos(lon), math.cos(lat) * math.sin(lon),
                   math.sin(lat)])
    return east, north, up


# ----------------------------------------------------------------- orbit


@dataclass
class Orbit:
    """Quadratic trajectory s(t) = p0 + v0 (t-t0) + acc (t-t0)^2 / 2 (ECEF)."""

    t0: float
    p0: np.ndarray
    v0: np.ndarray
    acc: np.ndarray

    def pos(self, t):
        dt = np.asarray(t, np.float64)[..., None] - self.t0
        return self.p0 + self.v0 * dt + 0.5 * self.acc * dt * dt

    def vel(self, t):
        dt = np.asarray(t, np.float64)[..., None] - self.t0
        return self.v0 + self.acc * dt

    def state_vectors(self, tcentre: float, n: int = 11, spacing: float = 10.0):
        """n state vectors every `spacing` s, starting on a whole second."""
        tfirst = math.floor(tcentre - spacing * (n - 1) / 2.0)
        t = tfirst + spacing * np.arange(n, dtype=np.float64)
        return t, self.pos(t), self.vel(t)

    def shifted(self, offset: np.ndarray) -> "Orbit":
        return Orbit(self.t0, self.p0 + offset, self.v0.copy(), self.acc.copy())


def make_orbit(lat_n: float, lon_n: float, t0: float, heading_deg: float = 0.0,
               alt: float = 700e3, speed: float = 7600.0, seed: int = 0) -> Orbit:
    """Orbit through (lat_n, lon_n, alt) at t0 with the given ground heading.

    The acceleration is the centripetal one of a circle at that radius with a
    seeded 1 % perturbation plus a small seeded cross-track component: a
    straight line with a small seeded curvature.
    """
    rng = np.random.default_rng(seed)
    p0 = llh_to_ecef(lat_n, lon_n, alt)
    east, north, up = enu_basis(lat_n, lon_n)
    hd = math.radians(heading_deg)
    v0 = speed * (math.cos(hd) * north + math.sin(hd) * east)
    g = speed * speed / np.linalg.norm(p0)
    acc = -g * (1.0 + 0.01 * rng.standard_normal()) * up
    acc = acc + 0.02 * rng.standard_normal() * np.cross(v0 / speed, up)
    return Orbit(t0, p0, v0, acc)


def broadside_ground_point(orbit: Orbit, t: float, slant_range: float,
                           h: float = 0.0, right: bool = True):
    """Point at height h, range slant_range, in the zero-Doppler plane at t."""
    s = orbit.pos(t)
    v = orbit.vel(t)
    vh = v / np.linalg.norm(v)
    up = s / np.linalg.norm(s)
    up = up - vh * np.dot(up, vh)
    up /= np.linalg.norm(up)
    side = np.cross(vh, up)
    if not right:
        side = -side

    def height(theta):
        x = s + slant_range * (-math.cos(theta) * up + math.sin(theta) * side)
        return ecef_to_llh(x)[2] - h, x

    lo, hi = 0.0, math.radians(85.0)
    for _ in range(100):
        mid = 0.5 * (lo + hi)
        if height(mid)[0] > 0.0:
            hi = mid
        else:
            lo = mid
    return height(0.5 * (lo + hi))[1]


def zero_doppler_time(orbit: Orbit, x, t: float) -> float:
    """Time where (x - s(t)).v(t) = 0 on the analytic trajectory (Newton)."""
    for _ in range(20):
        d = x - orbit.pos(t)
        v = orbit.vel(t)
        f = float(np.dot(d, v))
        fp = float(np.dot(d, orbit.acc) - np.dot(v, v))
        step = f / fp
        t -= step
        if abs(step) < 1e-9:
            break
    return t


# ------------------------------------------------------------------- DEM


def fbm_surface(shape, hurst: float = 0.8, hmin: float = 0.0, hmax: float = 3000.0,
                seed: int = 0) -> np.ndarray:
    """Periodic fractional-Brownian-motion heights (int16 metres).

    Spectral synthesis: white Gaussian noise shaped by |k|^-(H+1) (power
    |k|^-(2H+2) for a 2-D fBm surface), rescaled to [hmin, hmax].
    """
    rng = np.random.default_rng(seed)
    ny, nx = shape
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    spec = (rng.standard_normal(amp.shape) + 1j * rng.standard_normal(amp.shape)) * amp
    z = np.fft.irfft2(spec, s=shape)
    z = (z - z.min()) / (z.max() - z.min())
    return np.round(hmin + (hmax - hmin) * z).astype(np.int16)


# --------------------------------------------------------------- bursts


@dataclass
class BurstScene:
    geom: tdbp.BurstGeometry
    satloc: np.ndarray                 # [L][3]
    times: np.ndarray                  # [L]
    dem: np.ndarray                    # int16 [nlines][W] window
    demlength: int                     # rows of the full grid (drop-in files)
    latlons: np.ndarray                # (min_lat, max_lat, min_lon, max_lon)
    targets_xyz: np.ndarray            # [n][3]
    targets_amp: np.ndarray            # complex [n]
    targets_px: np.ndarray             # [n][2] (window row, column)
    clutter_sigma: float
    seed: int
    res: float = FS_IW / BW_IW
    extra: dict = field(default_factory=dict)

    @property
    def L(self):
        return self.geom.L

    @property
    def S(self):
        return self.geom.S


def aperture_from_caller(rngstart: float, wvl: float, prf: float, vel_mid_burst,
                         angc1: float = 0.0) -> int:
    """Integration length of sentinel/backprojectgpusub.f90:193-209."""
    vmag = float(np.linalg.norm(vel_mid_burst)) * (RE_APERTURE / (RE_APERTURE + HT_APERTURE))
    aperturetime = (rngstart * wvl / 2 / vmag / vmag * prf) / (
        1.0 + rngstart * angc1 * math.pi / 180.0 / vmag)
    return int(aperturetime * prf * 0.5)


def make_burst(L: int, S: int, W: int, nlines: int, *, centre_range: float = 850e3,
               dlat: Optional[float] = None, dlon: float = 0.0002,
               aperture: Optional[int] = None, dem: str = "flat",
               dem_source: Optional[np.ndarray] = None, dem_offset=(0, 0),
               ntargets: int = 5, clutter_sigma: float = 0.1, seed: int = 0,
               t0: float = 43200.0, lat_n: float = 35.0, lon_n: float = -118.0,
               orbit: Optional[Orbit] = None, wvl: float = WVL_S1,
               fs: float = FS_IW, prf: float = PRF_IW, angc0: float = 0.0,
               angc1: float = 0.0, fit_last_line: int = 0,
               valid_fill: float = 0.75) -> BurstScene:
    """One burst with a DEM window of nlines x W pixels centred on the scene.

    dlat (deg, negative) defaults to the posting that makes the band of rows
    with a valid aperture centre cover `valid_fill` of the window.
    """
    rng = np.random.default_rng(seed)
    if orbit is None:
        orbit = make_orbit(lat_n, lon_n, t0, seed=seed)
    dmrg = C_LIGHT / 2.0 / fs
    dtaz = 1.0 / prf
    tstart = t0 - 0.5 * (L - 1) * dtaz
    tend = tstart + (L - 1) * dtaz
    tmid = 0.5 * (tstart + tend)
    timeorbit, xx, vv = orbit.state_vectors(tmid)
    times = tstart + dtaz * np.arange(L, dtype=np.float64)
    satloc, satvel = tdbp.orbit_interp(timeorbit, xx, vv, times)
    xyz_mid, vel_mid = tdbp.orbit_interp(timeorbit, xx, vv, [tmid])
    xyz_mid, vel_mid = xyz_mid[0], vel_mid[0]

    rngstart = centre_range - 0.5 * S * dmrg
    rngend = rngstart + (S - 1) * dmrg
    if aperture is None:
        aperture = aperture_from_caller(rngstart, wvl, prf, satvel[L // 2], angc1)

    # Scene centre: broadside point at the centre range, at mid burst.
    g0 = broadside_ground_point(orbit, tmid, centre_range)
    lat_c, lon_c, _ = ecef_to_llh(g0)
    lat_c, lon_c = float(lat_c), float(lon_c)
    if dlat is None:
        ground_step = np.linalg.norm(satvel[L // 2]) * dtaz * (
            np.linalg.norm(g0) / np.linalg.norm(xyz_mid))
        band_m = max(L - aperture, 1) * ground_step
        dlat = -(band_m / (valid_fill * nlines)) / 111320.0
    dlat = -abs(dlat)
    firstlat = lat_c - 0.5 * nlines * dlat
    firstlon = lon_c - 0.5 * W * dlon
    demlength = nlines

    if dem == "flat":
        hgt = np.zeros((nlines, W), np.int16)
    else:
        src = dem_source if dem_source is not None else fbm_surface((1024, 1024), seed=seed)
        hgt = dem_window(src, dem_offset, nlines, W)

    return _burst_on_window(
        L=L, S=S, aperture=int(aperture), rngstart=rngstart, rngend=rngend, dmrg=dmrg,
        wvl=wvl, prf=prf, angc0=angc0, angc1=angc1, tstart=tstart, tmid=tmid,
        timeorbit=timeorbit, xx=xx, vv=vv, satloc=satloc, times=times, xyz_mid=xyz_mid,
        vel_mid=vel_mid, orbit=orbit, firstlat=firstlat, firstlon=firstlon, dlat=dlat,
        dlon=dlon, firstline=0, hgt=hgt, demlength=demlength, ntargets=ntargets,
        clutter_sigma=clutter_sigma, seed=seed, target_rng=rng, fs=fs,
        fit_last_line=fit_last_line)


def dem_window(src: np.ndarray, offset, nlines: int, W: int) -> np.ndarray:
    """nlines x W heights cut from a periodic source at (row, col) offset."""
    r0, c0 = offset
    rr = (np.arange(nlines) + r0) % src.shape[0]
    cc = (np.arange(W) + c0) % src.shape[1]
    return src[np.ix_(rr, cc)]


def _burst_on_window(*, L, S, aperture, rngstart, rngend, dmrg, wvl, prf, angc0, angc1,
                     tstart, tmid, timeorbit, xx, vv, satloc, times, xyz_mid, vel_mid,
                     orbit, firstlat, firstlon, dlat, dlon, firstline, hgt, demlength,
                     ntargets, clutter_sigma, seed, target_rng, fs,
                     fit_last_line) -> BurstScene:
    """Geometry, latlons and point targets of a burst whose DEM window is rows
    [firstline, firstline + nlines) of a grid starting at (firstlat, firstlon)."""
    nlines, W = hgt.shape
    lat0 = firstlat + firstline * dlat
    # latlons that reproduce exactly this window (rows firstline.., cols 0..W-1)
    latlons = np.array([lat0 + (nlines - 0.5) * dlat, lat0 + 0.5 * dlat,
                        firstlon + 0.5 * dlon, firstlon + (W - 0.5) * dlon])
    geom = tdbp.BurstGeometry(
        L=L, S=S, A=int(aperture), rngstart=rngstart, rngend=rngend, dmrg=dmrg, wvl=wvl,
        prf=prf, angc0=angc0, angc1=angc1, tmid=tmid, xyz_mid=xyz_mid, vel_mid=vel_mid,
        timeorbit=timeorbit, xx=xx, vv=vv, demwidth=W, firstline=firstline, nlines=nlines,
        firstpix=0, lastpix=W - 1, firstlat=firstlat, firstlon=firstlon, deltalat=dlat,
        deltalon=dlon, fit_last_line=fit_last_line)

    # Point targets on DEM pixels whose zero-Doppler pulse lies well inside
    # the burst (the aperture centre is then valid), drawn near the centre.
    rng = target_rng
    rows, cols = [], []
    margin = aperture // 2 + 8
    for _ in range(2000):
        if len(rows) == ntargets:
            break
        r = int(rng.integers(1, max(2, geom.nfit - 1)))
        c = int(rng.integers(W // 8, W - W // 8))
        x = llh_to_ecef(lat0 + r * dlat, firstlon + c * dlon, float(hgt[r, c]))
        p = (zero_doppler_time(orbit, x, tmid) - tstart) * prf
        if margin <= p <= L - 1 - margin:
            rows.append(r)
            cols.append(c)
    rows = np.array(rows, np.int64)
    cols = np.array(cols, np.int64)
    ntargets = len(rows)
    lat_t = lat0 + rows * dlat
    lon_t = firstlon + cols * dlon
    txyz = llh_to_ecef(lat_t, lon_t, hgt[rows, cols].astype(np.float64)).reshape(-1, 3)
    tamp = np.exp(1j * rng.uniform(-np.pi, np.pi, ntargets))
    return BurstScene(geom=geom, satloc=satloc, times=times, dem=hgt, demlength=demlength,
                      latlons=latlons, targets_xyz=txyz, targets_amp=tamp,
                      targets_px=np.stack([rows, cols], 1), clutter_sigma=clutter_sigma,
                      seed=seed, res=fs / BW_IW)


def range_taps(res: float, ntaps: int = 9) -> np.ndarray:
    """Real low-pass FIR (Hann-windowed sinc, cutoff 1/res cycles/sample) with
    unit energy: white noise through it keeps its variance."""
    k = np.arange(ntaps) - (ntaps - 1) / 2.0
    h = np.sinc(k / res) * np.hanning(ntaps + 2)[1:-1]
    return (h / np.sqrt(np.sum(h * h))).astype(np.float32)


def host_samples(sc: BurstScene, ntaps: int = 9, halfwidth: int = 16) -> np.ndarray:
    """Range-compressed burst in numpy (complex64 [L][S]); small bursts only."""
    rng = np.random.default_rng(sc.seed + 1000)
    L, S = sc.L, sc.S
    g = sc.geom
    d = np.zeros((L, S), np.complex128)
    if sc.clutter_sigma > 0:
        w = (rng.standard_normal((L, S)) + 1j * rng.standard_normal((L, S))) * (
            sc.clutter_sigma / math.sqrt(2.0))
        h = range_taps(sc.res, ntaps).astype(np.float64)
        half = ntaps // 2
        for t in range(ntaps):
            sh = t - half
            if sh >= 0:
                d[:, sh:] += h[t] * w[:, :S - sh]
            else:
                d[:, :S + sh] += h[t] * w[:, -sh:]
    j = np.arange(-halfwidth, halfwidth + 1)
    for x, a in zip(sc.targets_xyz, sc.targets_amp):
        R = np.linalg.norm(x[None, :] - sc.satloc, axis=1)
        r = (R - g.rngstart) / g.dmrg
        idx = np.rint(r).astype(np.int64)[:, None] + j[None, :]
        ok = (idx >= 0) & (idx < S)
        cyc = R * (2.0 / g.wvl)
        ph = np.exp(-2j * np.pi * (cyc - np.rint(cyc)))
        val = a * np.sinc((idx - r[:, None]) / sc.res) * ph[:, None]
        p = np.broadcast_to(np.arange(L)[:, None], idx.shape)
        np.add.at(d, (p[ok], idx[ok]), val[ok])
    return d.astype(np.complex64)


def device_samples(sc: BurstScene, d_out: "tdbp.DeviceArray", d_satloc, ntaps: int = 9,
                   halfwidth: int = 16, stream=None):
    """Same model as host_samples on the GPU (different noise realisation)."""
    tdbp.synth_clutter(d_out, sc.L, sc.S, sc.seed + 1000, sc.clutter_sigma,
                       range_taps(sc.res, ntaps), stream)
    if len(sc.targets_xyz):
        tdbp.synth_targets(d_out, sc.geom, d_satloc, sc.targets_xyz, sc.targets_amp,
                           res=sc.res, halfwidth=halfwidth, stream=stream)


# ---------------------------------------------------------- configurations


def small_burst(seed: int = 0) -> BurstScene:
    """One synthetic burst on the reference CPU path: 512 pulses x 4096
    samples, 256 x 256 grid at 0.0002 deg on a flat DEM, A = 256, five unit
    point targets plus clutter sigma 0.1."""
    return make_burst(512, 4096, 256, 256, dlat=-0.0002, dlon=0.0002, aperture=256,
                      dem="flat", ntargets=5, clutter_sigma=0.1, seed=seed)


def iw_burst(seed: int = 0, dem_source=None) -> BurstScene:
    """One full IW burst: 1500 pulses x 21000 samples, 4000 x 1500 grid on an
    fBm DEM (0-3000 m, Hurst 0.8), A from the caller's formula."""
    src = dem_source if dem_source is not None else fbm_surface((2048, 2048), seed=seed + 7)
    return make_burst(1500, 21000, 4000, 1500, dem="fbm", dem_source=src, ntargets=5,
                      clutter_sigma=0.1, seed=seed)


def iw_scene(seed: int = 1, nsub: int = 3, nburst: int = 9) -> List[BurstScene]:
    """IW scene: 3 subswaths x 9 bursts of uneven size.  Pulses U[1350,1550],
    samples U[18000,25000], output tiles U[1000,4500] x U[1200,1600] on the
    fBm DEM; subswath centre ranges 825/875/925 km."""
    rng = np.random.default_rng(seed)
    src = fbm_surface((2048, 2048), seed=seed + 7)
    orbit = make_orbit(35.0, -118.0, 43200.0, seed=seed)
    scenes = []
    burst_dt = 2.758  # s between bursts of one subswath (IW cycle time)
    for sw in range(nsub):
        for k in range(nburst):
            L = int(rng.integers(1350, 1551))
            S = int(rng.integers(18000, 25001))
            W = int(rng.integers(1000, 4501))
            nl = int(rng.integers(1200, 1601))
            t0 = 43200.0 + (k - (nburst - 1) / 2.0) * burst_dt + 0.9 * sw
            sc = make_burst(L, S, W, nl, centre_range=825e3 + 50e3 * sw, dem="fbm",
                            dem_source=src,
                            dem_offset=(int(rng.integers(0, 2048)), int(rng.integers(0, 2048))),
                            ntargets=3, clutter_sigma=0.1, seed=seed * 1000 + 100 * sw + k,
                            t0=t0, orbit=orbit)
            scenes.append(sc)
    return scenes


@dataclass
class StackGrid:
    """A common lat/lon DEM grid that every burst of a stack maps onto."""

    firstlat: float
    firstlon: float
    dlat: float
    dlon: float
    dem: np.ndarray                    # int16 [lines][W]

    @property
    def lines(self):
        return self.dem.shape[0]

    @property
    def W(self):
        return self.dem.shape[1]


def stack_burst(grid: StackGrid, orbit: Orbit, firstline: int, nlines: int, L: int,
                S: int, tcentre: float, centre_range: float, *, seed: int,
                target_seed: int, ntargets: int = 3, clutter_sigma: float = 0.1,
                wvl: float = WVL_S1, fs: float = FS_IW, prf: float = PRF_IW,
                aperture: Optional[int] = None) -> BurstScene:
    """A burst of L pulses centred at tcentre whose DEM window is grid rows
    [firstline, firstline + nlines), all W columns.  Targets are drawn from
    target_seed, so bursts of several acquisitions over one window share them."""
    dmrg = C_LIGHT / 2.0 / fs
    dtaz = 1.0 / prf
    tstart = tcentre - 0.5 * (L - 1) * dtaz
    tend = tstart + (L - 1) * dtaz
    tmid = 0.5 * (tstart + tend)
    timeorbit, xx, vv = orbit.state_vectors(tmid)
    times = tstart + dtaz * np.arange(L, dtype=np.float64)
    satloc, satvel = tdbp.orbit_interp(timeorbit, xx, vv, times)
    xyz_mid, vel_mid = tdbp.orbit_interp(timeorbit, xx, vv, [tmid])
    rngstart = centre_range - 0.5 * S * dmrg
    rngend = rngstart + (S - 1) * dmrg
    if aperture is None:
        aperture = aperture_from_caller(rngstart, wvl, prf, satvel[L // 2])
    return _burst_on_window(
        L=L, S=S, aperture=aperture, rngstart=rngstart, rngend=rngend, dmrg=dmrg, wvl=wvl,
        prf=prf, angc0=0.0, angc1=0.0, tstart=tstart, tmid=tmid, timeorbit=timeorbit,
        xx=xx, vv=vv, satloc=satloc, times=times, xyz_mid=xyz_mid[0], vel_mid=vel_mid[0],
        orbit=orbit, firstlat=grid.firstlat, firstlon=grid.firstlon, dlat=grid.dlat,
        dlon=grid.dlon, firstline=firstline, hgt=grid.dem[firstline:firstline + nlines],
        demlength=grid.lines, ntargets=ntargets, clutter_sigma=clutter_sigma, seed=seed,
        target_rng=np.random.default_rng(target_seed), fs=fs, fit_last_line=0)


@dataclass
class Stack:
    """Acquisitions x bursts over one grid; bursts[a][k] is burst k of
    acquisition a, baselines[a] its perpendicular orbit offset (ECEF, m)."""

    grid: StackGrid
    bursts: List[List[BurstScene]]
    baselines: np.ndarray

    def items(self):
        """(acquisition, burst index, scene) in processing order."""
        for a, row in enumerate(self.bursts):
            for k, sc in enumerate(row):
                yield a, k, sc


def ts_stack(seed: int = 1, nacq: int = 30, nburst: int = 9, lines: int = 13500,
            W: int = 4000, max_baseline: float = 200.0, pulses=(1350, 1550),
            samples=(18000, 25000), rows=(1200, 1600), aperture: Optional[int] = None,
            ntargets: int = 3) -> Stack:
    """Time-series stack: nacq acquisitions x nburst bursts over one common
    lines x W grid on an fBm DEM.

    Bursts are sized as in the IW scene (pulses U[1350,1550], samples
    U[18000,25000], window rows U[1200,1600]); burst k's window is centred on
    grid row (k + 1/2) lines/nburst, so neighbouring windows overlap or leave
    gaps and the .geo merge order decides the overlap.  The row posting makes
    the band of valid aperture centres of a 1450-pulse burst span lines/nburst
    rows (the ranges are parameters; `aperture` fixes A instead of the
    caller's formula, for bursts shorter than that formula's A).  Every
    acquisition flies the reference orbit shifted by a seeded
    perpendicular baseline of magnitude U[0, max_baseline] m at a seeded angle
    about the along-track axis; the pulses of burst k are centred on the
    reference orbit's zero-Doppler time of the window centre.  This is a
    stripmap-like synthetic stack (no TOPS steering): what it fixes is the
    work per burst, the grid and the merge, not the IW timeline.
    """
    rng = np.random.default_rng(seed)
    centre_range = 875e3
    t0 = 43200.0
    orbit0 = make_orbit(35.0, -118.0, t0, seed=seed)
    g0 = broadside_ground_point(orbit0, t0, centre_range)
    lat_c, lon_c, _ = ecef_to_llh(g0)
    lat_c, lon_c = float(lat_c), float(lon_c)
    s0, v0 = orbit0.pos(t0), orbit0.vel(t0)
    ground_step = np.linalg.norm(v0) / PRF_IW * (np.linalg.norm(g0) / np.linalg.norm(s0))
    dmrg = C_LIGHT / 2.0 / FS_IW
    l_mid = 0.5 * (pulses[0] + pulses[1])
    a_ref = aperture if aperture is not None else aperture_from_caller(
        centre_range - 0.25 * (samples[0] + samples[1]) * dmrg, WVL_S1, PRF_IW, v0)
    spacing = lines / nburst
    dlat = -((l_mid - a_ref) * ground_step / spacing) / 111320.0
    dlon = 0.0002
    src = fbm_surface((2048, 2048), seed=seed + 7)
    grid = StackGrid(firstlat=lat_c - 0.5 * lines * dlat, firstlon=lon_c - 0.5 * W * dlon,
                     dlat=dlat, dlon=dlon,
                     dem=dem_window(src, (int(rng.integers(0, 2048)),
                                          int(rng.integers(0, 2048))), lines, W))

    # Burst windows and pulse centres, common to every acquisition.
    windows, tcs = [], []
    for k in range(nburst):
        nl = int(rng.integers(rows[0], rows[1] + 1))
        crow = int(round((k + 0.5) * spacing))
        fl = min(max(crow - nl // 2, 0), lines - nl)
        x = llh_to_ecef(grid.firstlat + crow * dlat, grid.firstlon + 0.5 * W * dlon,
                        float(grid.dem[crow, W // 2]))
        windows.append((fl, nl))
        tcs.append(zero_doppler_time(orbit0, x, t0))

    vh = v0 / np.linalg.norm(v0)
    up = s0 / np.linalg.norm(s0)
    up = up - vh * np.dot(up, vh)
    up /= np.linalg.norm(up)
    side = np.cross(vh, up)
    bursts, baselines = [], []
    for a in range(nacq):
        bmag = rng.uniform(0.0, max_baseline)
        ang = rng.uniform(-math.pi, math.pi)
        off = bmag * (math.cos(ang) * up + math.sin(ang) * side)
        baselines.append(off)
        orbit = orbit0.shifted(off)
        row = []
        for k, ((fl, nl), tc) in enumerate(zip(windows, tcs)):
            L = int(rng.integers(pulses[0], pulses[1] + 1))
            S = int(rng.integers(samples[0], samples[1] + 1))
            row.append(stack_burst(grid, orbit, fl, nl, L, S, tc, centre_range,
                                   seed=seed * 100000 + 100 * a + k,
                                   target_seed=seed * 1000 + k, aperture=aperture,
                                   ntargets=ntargets))
        bursts.append(row)
    return Stack(grid=grid, bursts=bursts, baselines=np.array(baselines))
