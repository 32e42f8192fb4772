"""ctypes wrappers of the checker libraries.

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product (srg_amd/).

* ``Oracle``: oracle/libtdbp_oracle.so, the C restatement of the reference
  algorithm (tdbp_oracle.c; each function cites the reference lines).
* ``RefCPU``: oracle/_ref/libref_tdbp_cpu.so, the reference's own CPU path
  (sentinel/azimuth_compress_cpu.c + orbit Fortran) compiled from its sources
  by oracle/Makefile.  Present where it was built; absent is not an error.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libtdbp_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_tdbp_cpu.so")
REF_O0_SO = os.path.join(HERE, "_ref", "libref_tdbp_cpu_O0.so")
BOUNDS_SO = os.path.join(HERE, "_ref", "libref_bounds.so")

CPU_SEMANTICS, GPU_SEMANTICS = 0, 1

_D = C.POINTER(C.c_double)
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int)
_S16 = C.POINTER(C.c_int16)
_I64 = C.POINTER(C.c_int64)


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        L = C.CDLL(path)
        L.orc_orbithermite.argtypes = [_D, _D, _D, C.c_double, _D, _D]
        L.orc_intp_orbit.argtypes = [_D, _D, _D, C.c_int, C.c_double, _D, _D]
        L.orc_orbitrangetime.argtypes = [_D, _D, _D, _D, C.c_int, C.c_double, _D, _D, _D, _D]
        L.orc_setxyz.argtypes = [_S16, _D, _D, C.c_int, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_fitcoef.argtypes = [_D, C.c_int, _D, _D, _D, C.c_int, C.c_double, _D, _D,
                                  C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _D]
        L.orc_setazoff.argtypes = [_D, _D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_pixelint.argtypes = [_F, _D, _D, _D, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_double, C.c_double, C.c_double, C.c_double, _F,
                                   C.c_int, C.c_int, _D, _I64, C.c_int]
        L.orc_window.argtypes = [_D, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_int, C.c_int, _I, _I, _I, _I]
        L.orc_backproject.restype = C.c_int64
        L.orc_azimuth_compress.argtypes = [
            _F, _D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
            C.c_double, C.c_double, C.c_double, _D, _D, _D, _D, C.c_int, C.c_double,
            C.c_double, C.c_double, _D, _D, C.c_double, C.c_double, C.c_int, C.c_double,
            C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_merge_geo.argtypes = [_F, _F, C.c_long, C.c_int]
        self.lib = L

    # -- orbit
    def orbithermite(self, x4, v4, t4, time):
        x4, v4, t4 = _c(x4, np.float64), _c(v4, np.float64), _c(t4, np.float64)
        pos, vel = np.empty(3), np.empty(3)
        self.lib.orc_orbithermite(_p(x4, _D), _p(v4, _D), _p(t4, _D), time, _p(pos, _D),
                                  _p(vel, _D))
        return pos, vel

    def intp_orbit(self, t, x, v, time):
        t, x, v = _c(t, np.float64), _c(x, np.float64), _c(v, np.float64)
        pos, vel = np.empty(3), np.empty(3)
        self.lib.orc_intp_orbit(_p(t, _D), _p(x, _D), _p(v, _D), len(t), time, _p(pos, _D),
                                _p(vel, _D))
        return pos, vel

    def orbitrangetime(self, xyz, t, x, v, tline0, satx0, satv0):
        xyz = _c(xyz, np.float64)
        t, x, v = _c(t, np.float64), _c(x, np.float64), _c(v, np.float64)
        s0, v0 = _c(satx0, np.float64), _c(satv0, np.float64)
        tl, rg = C.c_double(), C.c_double()
        self.lib.orc_orbitrangetime(_p(xyz, _D), _p(t, _D), _p(x, _D), _p(v, _D), len(t),
                                    tline0, _p(s0, _D), _p(v0, _D), C.byref(tl), C.byref(rg))
        return tl.value, rg.value

    # -- per-burst stages on a geometry object (srg_amd.tdbp.BurstGeometry-like)
    def setxyz(self, g, dem):
        dem = _c(dem, np.int16)
        xyz = np.zeros((g.nlines, g.demwidth, 3))
        xyzfit = np.zeros((g.nlines, 9))
        self.lib.orc_setxyz(_p(dem, _S16), _p(xyz, _D), _p(xyzfit, _D), g.demwidth,
                            g.firstlat, g.deltalat, g.firstlon, g.deltalon, g.firstline,
                            g.nlines, g.firstpix, g.lastpix)
        return xyz, xyzfit

    def fitcoef(self, g, xyzfit, nfit=None):
        nfit = g.nfit if nfit is None else nfit
        xyzfit = _c(xyzfit, np.float64)
        t, x, v = (_c(g.timeorbit, np.float64), _c(g.xx, np.float64), _c(g.vv, np.float64))
        xm, vm = _c(g.xyz_mid, np.float64), _c(g.vel_mid, np.float64)
        coef = np.zeros((g.nlines, 3))
        self.lib.orc_fitcoef(_p(xyzfit, _D), nfit, _p(t, _D), _p(x, _D), _p(v, _D), len(t),
                             g.tmid, _p(xm, _D), _p(vm, _D), g.wvl, g.angc0, g.angc1, g.prf,
                             g.L, _p(coef, _D))
        return coef

    def setazoff(self, g, coef):
        coef = _c(coef, np.float64)
        azoff = np.empty((g.nlines, g.demwidth))
        self.lib.orc_setazoff(_p(coef, _D), _p(azoff, _D), g.demwidth, g.nlines, g.firstpix,
                              g.lastpix, g.A, g.L)
        return azoff

    def pixelint(self, g, burst, satloc, xyz, azoff, semantics=GPU_SEMANTICS,
                 nthreads=0, rows: Optional[slice] = None):
        """Returns (cf32 out, complex128 accumulators, pairs).  `rows` limits
        the computation to a slice of window lines (xyz/azoff sliced too)."""
        burst = _c(burst, np.complex64)
        satloc = _c(satloc, np.float64)
        if rows is not None:
            xyz, azoff = xyz[rows], azoff[rows]
        xyz, azoff = _c(xyz, np.float64), _c(azoff, np.float64)
        nl = azoff.shape[0]
        out = np.zeros((nl, g.demwidth), np.complex64)
        acc = np.zeros((nl, g.demwidth), np.complex128)
        pairs = C.c_int64()
        self.lib.orc_pixelint(_p(burst.view(np.float32), _F), _p(satloc, _D), _p(xyz, _D),
                              _p(azoff, _D), g.demwidth, g.A, g.L, g.S, g.rngstart,
                              g.rngend, g.dmrg, g.wvl, _p(out.view(np.float32), _F), nl,
                              semantics, _p(acc.view(np.float64), _D), C.byref(pairs),
                              nthreads)
        return out, acc, pairs.value

    def backproject(self, g, burst, satloc, dem, semantics=GPU_SEMANTICS, nthreads=0):
        """Full in-memory burst: (cf32 out, complex128 acc, pairs, coef, azoff)."""
        xyz, xyzfit = self.setxyz(g, dem)
        coef = self.fitcoef(g, xyzfit)
        azoff = self.setazoff(g, coef)
        out, acc, pairs = self.pixelint(g, burst, satloc, xyz, azoff, semantics, nthreads)
        return out, acc, pairs, coef, azoff

    def window(self, latlons, firstlat, firstlon, deltalat, deltalon, W, demlength):
        ll = _c(latlons, np.float64)
        o = [C.c_int() for _ in range(4)]
        self.lib.orc_window(_p(ll, _D), firstlat, firstlon, deltalat, deltalon, W, demlength,
                            *[C.byref(x) for x in o])
        return tuple(x.value for x in o)

    def merge_geo(self, grid, out, firstline, semantics=GPU_SEMANTICS):
        """Merge a burst window out [nlines][W] into grid rows [firstline, ...)
        in place (grid: C-contiguous complex64)."""
        assert grid.dtype == np.complex64 and grid.flags.c_contiguous
        out = _c(out, np.complex64)
        nl, W = out.shape
        assert grid.shape[1] == W and firstline + nl <= grid.shape[0]
        win = grid[firstline:firstline + nl]
        self.lib.orc_merge_geo(_p(win.view(np.float32), _F), _p(out.view(np.float32), _F),
                               nl * W, semantics)

    def azimuth_compress(self, burst, satloc, L, S, W, demlength, fdout, fddem, dlat, dlon,
                         firstlat, firstlon, latlons, t, x, v, rngstart, rngend, tmid,
                         xyz_mid, vel_mid, dmrg, wvl, A, angc0, angc1, prf,
                         semantics=GPU_SEMANTICS, nfit_last=0):
        burst = _c(burst, np.complex64)
        arr = [_c(a, np.float64) for a in (satloc, latlons, t, x, v, xyz_mid, vel_mid)]
        satloc, latlons, t, x, v, xyz_mid, vel_mid = arr
        self.lib.orc_azimuth_compress(
            _p(burst.view(np.float32), _F), _p(satloc, _D), L, S, W, demlength, fdout, fddem,
            dlat, dlon, firstlat, firstlon, _p(latlons, _D), _p(t, _D), _p(x, _D), _p(v, _D),
            len(t), rngstart, rngend, tmid, _p(xyz_mid, _D), _p(vel_mid, _D), dmrg, wvl, A,
            angc0, angc1, prf, semantics, nfit_last)


class RefCPU:
    """The reference CPU path compiled from its own sources (oracle/_ref)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        L = C.CDLL(path)
        # sentinel/azimuth_compress_cpu.c:40, :89, :155, :178; Fortran by-ref
        L.pixelint.argtypes = [_F, _D, _D, _D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_double, C.c_double, C.c_double, C.c_double, _F, C.c_int]
        L.setxyz.argtypes = [_S16, _D, _D, _D, C.c_int, C.c_double, C.c_double, C.c_double,
                             C.c_double, C.c_int, C.c_int, C.c_int, C.c_int]
        L.setazoff.argtypes = [_D, _D, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orbithermite_.argtypes = [_D, _D, _D, _D, _D, _D]
        L.intp_orbit_.argtypes = [_D, _D, _D, _I, _D, _D, _D]
        L.intp_orbit_.restype = C.c_int
        L.orbitrangetime_.argtypes = [_D, _D, _D, _D, _I, _D, _D, _D, _D, _D]
        L.azimuth_compress_cpu_.argtypes = [C.c_void_p] + [C.c_void_p] * 32
        self.lib = L

    def orbithermite(self, x4, v4, t4, time):
        x4, v4, t4 = _c(x4, np.float64), _c(v4, np.float64), _c(t4, np.float64)
        tm = C.c_double(time)
        pos, vel = np.empty(3), np.empty(3)
        self.lib.orbithermite_(_p(x4, _D), _p(v4, _D), _p(t4, _D), C.byref(tm), _p(pos, _D),
                               _p(vel, _D))
        return pos, vel

    def intp_orbit(self, t, x, v, time):
        t, x, v = _c(t, np.float64), _c(x, np.float64), _c(v, np.float64)
        n, tm = C.c_int(len(t)), C.c_double(time)
        pos, vel = np.empty(3), np.empty(3)
        self.lib.intp_orbit_(_p(t, _D), _p(x, _D), _p(v, _D), C.byref(n), C.byref(tm),
                             _p(pos, _D), _p(vel, _D))
        return pos, vel

    def orbitrangetime(self, xyz, t, x, v, tline0, satx0, satv0):
        xyz = _c(xyz, np.float64)
        t, x, v = _c(t, np.float64), _c(x, np.float64), _c(v, np.float64)
        s0, v0 = _c(satx0, np.float64).copy(), _c(satv0, np.float64).copy()
        n, t0 = C.c_int(len(t)), C.c_double(tline0)
        tl, rg = C.c_double(), C.c_double()
        self.lib.orbitrangetime_(_p(xyz, _D), _p(t, _D), _p(x, _D), _p(v, _D), C.byref(n),
                                 C.byref(t0), _p(s0, _D), _p(v0, _D), C.byref(tl), C.byref(rg))
        return tl.value, rg.value

    def setxyz(self, g, dem):
        dem = _c(dem, np.int16)
        xyz = np.zeros((g.nlines, g.demwidth, 3))
        xyzfit = np.zeros((g.nlines, 9))
        azoff = np.zeros((g.nlines, g.demwidth))
        self.lib.setxyz(_p(dem, _S16), _p(xyz, _D), _p(xyzfit, _D), _p(azoff, _D), g.demwidth,
                        g.firstlat, g.deltalat, g.firstlon, g.deltalon, g.firstline,
                        g.nlines, g.firstpix, g.lastpix)
        return xyz, xyzfit

    def setazoff(self, g, coef):
        coef = _c(coef, np.float64)
        azoff = np.zeros((g.nlines, g.demwidth))  # the reference leaves outside pixels
        self.lib.setazoff(_p(coef, _D), _p(azoff, _D), g.demwidth, g.nlines, g.firstpix,
                          g.lastpix, g.A, g.L)
        return azoff

    def pixelint(self, g, burst, satloc, xyz, azoff, rows: Optional[slice] = None):
        burst = _c(burst, np.complex64)
        satloc = _c(satloc, np.float64)
        if rows is not None:
            xyz, azoff = xyz[rows], azoff[rows]
        xyz, azoff = _c(xyz, np.float64), _c(azoff, np.float64)
        nl = azoff.shape[0]
        out = np.zeros((nl, g.demwidth), np.complex64)
        self.lib.pixelint(_p(burst.view(np.float32), _F), _p(satloc, _D), _p(xyz, _D),
                          _p(azoff, _D), g.demwidth, g.A, 0, g.L, g.S, g.rngstart, g.rngend,
                          g.dmrg, g.wvl, _p(out.view(np.float32), _F), nl)
        return out

    def azimuth_compress_cpu(self, burst, satloc, L, S, W, demlength, fdout, fddem, dlat,
                             dlon, firstlat, firstlon, latlons, t, x, v, rngstart, rngend,
                             tmid, xyz_mid, vel_mid, dmrg, wvl, A, angc0, angc1, prf):
        """The reference CPU entry with its by-reference Fortran ABI."""
        burst = _c(burst, np.complex64)
        sat = np.zeros((100000, 3))
        sat[:L] = satloc
        tt = np.zeros(1000)
        tt[:len(t)] = t
        xx = np.zeros((1000, 3))
        xx[:len(t)] = x
        vv = np.zeros((1000, 3))
        vv[:len(t)] = v
        ib = lambda v_: C.byref(C.c_int(int(v_)))  # noqa: E731
        db = lambda v_: C.byref(C.c_double(float(v_)))  # noqa: E731
        keep = [sat, tt, xx, vv, _c(latlons, np.float64), _c(xyz_mid, np.float64),
                _c(vel_mid, np.float64), np.zeros(100000)]
        args = [burst.ctypes.data, sat.ctypes.data, ib(L), ib(S), ib(W), ib(demlength),
                ib(fdout), ib(fddem), db(dlat), db(dlon), db(firstlat), db(firstlon),
                keep[4].ctypes.data, tt.ctypes.data, xx.ctypes.data, vv.ctypes.data,
                ib(len(t)), db(rngstart), db(rngend), db(0.0), db(0.0), db(tmid),
                keep[5].ctypes.data, keep[6].ctypes.data, keep[7].ctypes.data,
                db(1.0 / prf), db(dmrg), db(wvl), ib(A), ib(0), db(angc0), db(angc1),
                db(prf)]
        self.lib.azimuth_compress_cpu_(*args)
        return keep


class RefBounds:
    """bounds_ and latlon_ of the reference compiled from its own Fortran
    sources (oracle/_ref/libref_bounds.so)."""

    def __init__(self, path: str = BOUNDS_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        self.lib = C.CDLL(path)

    def bounds(self, startime, endtime, startr, endr, timeorbit, xx, vv):
        t = _c(timeorbit, np.float64)
        x = _c(xx, np.float64)
        v = _c(vv, np.float64)
        out = np.zeros(4)
        d = lambda a: C.byref(C.c_double(float(a)))  # noqa: E731
        self.lib.bounds_(d(startime), d(endtime), d(startr), d(endr), _p(t, _D), _p(x, _D),
                         _p(v, _D), C.byref(C.c_int(len(t))), _p(out, _D))
        return out


def bounds_available(path: str = None) -> bool:
    return os.path.exists(path or BOUNDS_SO)


def ref_available(path: str = REF_SO) -> bool:
    return os.path.exists(path)
