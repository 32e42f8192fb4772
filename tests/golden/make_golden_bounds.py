"""Write tests/golden/bounds.npz: the reference's burst footprint bounds_
(sentinel/bounds.f90 and the geodesy it calls, compiled from the reference
sources into oracle/_ref/libref_bounds.so by oracle/Makefile) on seeded
synthetic orbits: several headings, latitudes, burst times and range
windows.  Stores the inputs and the four latlons per case.

Usage: python tests/golden/make_golden_bounds.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import pyoracle  # noqa: E402
from srg_amd import synth  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def cases(n=32, seed=11):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        lat = float(rng.uniform(-75, 75))
        lon = float(rng.uniform(-180, 180))
        hdg = float(rng.uniform(-180, 180))
        t0 = 43200.0 + float(rng.uniform(-50, 50))
        orbit = synth.make_orbit(lat, lon, 43200.0, heading_deg=hdg, seed=int(rng.integers(1e6)))
        t, x, v = orbit.state_vectors(t0)
        r0 = float(rng.uniform(780e3, 950e3))
        out.append((t0 - 0.45 - 2.0, t0 + 0.45 + 2.0, r0, r0 + float(rng.uniform(3e4, 6e4)),
                    t, x, v))
    return out


def main():
    rb = pyoracle.RefBounds()
    cs = cases()
    ll = np.array([rb.bounds(*c) for c in cs])
    np.savez_compressed(os.path.join(OUT, "bounds.npz"),
                        scalars=np.array([c[:4] for c in cs]),
                        timeorbit=np.array([c[4] for c in cs]),
                        xx=np.array([c[5] for c in cs]), vv=np.array([c[6] for c in cs]),
                        latlons=ll)
    print("wrote bounds.npz", ll.shape)


if __name__ == "__main__":
    main()
