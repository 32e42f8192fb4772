"""Write golden vectors from the reference's own CPU path.

Runs in the build container, where oracle/_ref/libref_tdbp_cpu.so is compiled
from the reference sources (sentinel/azimuth_compress_cpu.c, orbitrangetime.f90,
intp_orbit.f90, orbithermite.f) by oracle/Makefile.  The reference holds no
fixtures of its own, so these are its outputs on inputs made by this repo's
synthetic generator (seeded).  Each fixture stores its inputs (or, for the
config-1 burst, a checksum of inputs regenerated from the seed) and the
reference outputs:

  hermite.npz   orbithermite_ on random 4-vector sets
  orbit.npz     intp_orbit_ and orbitrangetime_ on a synthetic orbit
  tiny.npz      setxyz / setazoff / pixelint and the full entry
                azimuth_compress_cpu_ (DEM and .geo through file descriptors)
                on a 64-pulse x 256-sample burst, 16 x 24 grid
  small_burst.npz   pixelint on the small burst (512 x 4096, 256 x 256, A = 256): the
                computed rows only, with the input checksum

Usage: python tests/golden/make_golden.py
"""
import ctypes
import hashlib
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402
from srg_amd import synth  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def tiny_scene(seed=3):
    return synth.make_burst(64, 256, 24, 16, centre_range=846e3, dlat=-0.00002,
                            dlon=0.0002, aperture=32, dem="fbm",
                            dem_source=synth.fbm_surface((64, 64), hmax=800.0, seed=seed),
                            ntargets=2, clutter_sigma=0.3, seed=seed)


def input_digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def geom_dict(g):
    keys = ["L", "S", "A", "rngstart", "rngend", "dmrg", "wvl", "prf", "angc0", "angc1",
            "tmid", "demwidth", "firstline", "nlines", "firstpix", "lastpix", "firstlat",
            "firstlon", "deltalat", "deltalon", "fit_last_line"]
    d = {f"g_{k}": np.array(getattr(g, k)) for k in keys}
    d.update(g_xyz_mid=np.asarray(g.xyz_mid), g_vel_mid=np.asarray(g.vel_mid),
             g_timeorbit=np.asarray(g.timeorbit), g_xx=np.asarray(g.xx),
             g_vv=np.asarray(g.vv))
    return d


def main():
    ref = pyoracle.RefCPU()
    orc = pyoracle.Oracle()
    libc = ctypes.CDLL(None)
    rng = np.random.default_rng(20261202)

    # 1. orbithermite on random vector sets (positions/velocities of a smooth
    #    trajectory plus noise, uneven times).
    cases = []
    for _ in range(32):
        t4 = np.sort(rng.uniform(0, 40, 4)) + 43000.0
        x4 = rng.normal(0, 7e6, (4, 3))
        v4 = rng.normal(0, 7e3, (4, 3))
        tm = rng.uniform(t4[0] - 5, t4[3] + 5)
        p, v = ref.orbithermite(x4, v4, t4, tm)
        cases.append((x4, v4, t4, tm, p, v))
    np.savez_compressed(os.path.join(OUT, "hermite.npz"),
                        x4=np.array([c[0] for c in cases]), v4=np.array([c[1] for c in cases]),
                        t4=np.array([c[2] for c in cases]), tm=np.array([c[3] for c in cases]),
                        pos=np.array([c[4] for c in cases]), vel=np.array([c[5] for c in cases]))

    # 2. intp_orbit and orbitrangetime on the tiny scene's orbit.
    sc = tiny_scene()
    g = sc.geom
    times = np.concatenate([np.linspace(g.timeorbit[0] - 3, g.timeorbit[-1] + 3, 41),
                            sc.times[::7]])
    ip = np.array([np.concatenate(ref.intp_orbit(g.timeorbit, g.xx, g.vv, t)) for t in times])
    xyz, xyzfit = ref.setxyz(g, sc.dem)
    pts = np.concatenate([xyzfit.reshape(-1, 3), sc.targets_xyz])
    ort = np.array([ref.orbitrangetime(p, g.timeorbit, g.xx, g.vv, g.tmid, g.xyz_mid,
                                       g.vel_mid) for p in pts])
    np.savez_compressed(os.path.join(OUT, "orbit.npz"), times=times, intp=ip, points=pts,
                        ort=ort, **geom_dict(g))

    # 3. tiny burst: stages and the full entry through files.
    burst = synth.host_samples(sc)
    coef = orc.fitcoef(g, xyzfit)          # any coefficients; setazoff is pinned on them
    azoff = ref.setazoff(g, coef)
    out = ref.pixelint(g, burst, sc.satloc, xyz, azoff)
    with tempfile.TemporaryDirectory() as td:
        demf, geof = os.path.join(td, "dem"), os.path.join(td, "geo")
        # a DEM two lines taller than the window, window starting at line 1
        full = np.concatenate([sc.dem[:1], sc.dem, sc.dem[-1:]]).astype(np.int16)
        full.tofile(demf)
        np.zeros((full.shape[0], g.demwidth), np.complex64).tofile(geof)
        fl = g.firstlat - g.deltalat       # grid starts one line north
        latlons = np.array([fl + (g.nlines + 0.5) * g.deltalat, fl + 1.5 * g.deltalat,
                            g.firstlon + 0.5 * g.deltalon,
                            g.firstlon + (g.demwidth - 0.5) * g.deltalon])
        fdd = os.open(demf, os.O_RDWR)
        fdo = os.open(geof, os.O_RDWR)
        args = (burst, sc.satloc, g.L, g.S, g.demwidth, full.shape[0], fdo, fdd, g.deltalat,
                g.deltalon, fl, g.firstlon, latlons, g.timeorbit, g.xx, g.vv, g.rngstart,
                g.rngend, g.tmid, g.xyz_mid, g.vel_mid, g.dmrg, g.wvl, g.A, g.angc0,
                g.angc1, g.prf)
        ref.azimuth_compress_cpu(*args)
        # the reference leaves its FILE* unflushed until exit
        libc.fflush(None)
        os.close(fdd)
        os.close(fdo)
        geo = np.fromfile(geof, np.complex64).reshape(full.shape[0], g.demwidth)
    np.savez_compressed(os.path.join(OUT, "tiny.npz"), burst=burst, satloc=sc.satloc,
                        dem=sc.dem, dem_full=full, xyz=xyz, xyzfit=xyzfit, coef=coef,
                        azoff=azoff, out=out, geo=geo, latlons=latlons, firstlat_full=fl,
                        **geom_dict(g))

    # 4. the small burst (inputs regenerated from the seed; checksum stored).
    sc1 = synth.small_burst()
    b1 = synth.host_samples(sc1)
    g1 = sc1.geom
    xyz1, xyzfit1 = ref.setxyz(g1, sc1.dem)
    coef1 = orc.fitcoef(g1, xyzfit1)
    az1 = ref.setazoff(g1, coef1)
    out1 = ref.pixelint(g1, b1, sc1.satloc, xyz1, az1)
    rows = np.nonzero((az1 > 0).any(1))[0]
    np.savez_compressed(os.path.join(OUT, "small_burst.npz"), rows=rows, out=out1[rows],
                        azoff=az1[rows], coef=coef1,
                        digest=np.array(input_digest(b1, sc1.satloc, sc1.dem)))
    print("golden vectors written to", OUT)


if __name__ == "__main__":
    main()
