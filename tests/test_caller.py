"""Host-side tests of the burst set-up (the Fortran driver's parameter
derivation, sentinel/backprojectgpusub.f90:140-247): the footprint against the
reference's own compiled bounds.f90 (golden vectors, and live where
oracle/_ref is built), the header decoding, steering fit and integration
length against the oracle restatement bit for bit and against closed forms."""
import math
import os

import numpy as np
import pytest

from oracle import caller_oracle as CO
from oracle import pyoracle
from srg_amd import raw, synth
from srg_amd import tdbp as T

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_bounds_golden_bitwise():
    d = np.load(os.path.join(GOLD, "bounds.npz"))
    for k in range(len(d["latlons"])):
        s = d["scalars"][k]
        got = T.bounds(s[0], s[1], s[2], s[3], d["timeorbit"][k], d["xx"][k], d["vv"][k])
        assert np.array_equal(got, d["latlons"][k]), (k, got - d["latlons"][k])


@pytest.mark.skipif(not pyoracle.bounds_available(), reason="oracle/_ref not built here")
def test_bounds_live_reference():
    rb = pyoracle.RefBounds()
    rng = np.random.default_rng(4)
    for _ in range(40):
        o = synth.make_orbit(float(rng.uniform(-80, 80)), float(rng.uniform(-180, 180)),
                             43200.0, heading_deg=float(rng.uniform(-180, 180)),
                             seed=int(rng.integers(1e6)))
        t, x, v = o.state_vectors(43200.0)
        r0 = float(rng.uniform(7e5, 1e6))
        args = (43198.0, 43203.5, r0, r0 + 5e4, t, x, v)
        assert np.array_equal(T.bounds(*args), rb.bounds(*args))


def test_bounds_contains_the_scene():
    """The footprint of the scene's burst holds its DEM window's centre."""
    sc = synth.iw_burst()
    g = sc.geom
    L = g.L
    t0 = g.tmid - (L - 1) / 2.0 / g.prf
    ll = T.bounds(t0 - 2.0, t0 + (L - 1) / g.prf + 2.0, g.rngstart, g.rngend, g.timeorbit,
                  g.xx, g.vv)
    lat_c = g.firstlat + 0.5 * g.nlines * g.deltalat
    lon_c = g.firstlon + 0.5 * g.demwidth * g.deltalon
    assert ll[0] < lat_c < ll[1] and ll[2] < lon_c < ll[3]


def iw_header(**kw):
    return raw.header_bytes(raw.HeaderCodes(**kw))


def test_burst_decode_matches_oracle_and_published_values():
    h = iw_header(starttime=43200.123456789, swst=5000)
    p = T.burst_decode(h).as_dict()
    o = CO.burstparams(h)
    for k, v in o.items():
        assert p[k] == v, k
    assert abs(p["prf"] - 1717.13) < 0.01                        # IW PRF
    assert abs(p["samplefreq"] - 64.345238e6) < 1.0
    assert p["starttime"] == 43200.123456789
    assert p["wvl"] == float(np.float32(0.05546576)) != 0.05546576      # Q5
    # range0 = c/2 (rank PRI + SWST) + 60 m
    c = 299792458.0
    assert abs(p["range0"] - (c / 2 * (9 * p["pri"] + p["swst"]) + 60.0)) < 1e-6
    rng = np.random.default_rng(2)
    for _ in range(200):
        h = iw_header(pri=int(rng.integers(1, 1 << 24)), swst=int(rng.integers(0, 1 << 24)),
                      rank=int(rng.integers(0, 32)), starttime=float(rng.uniform(0, 86400)))
        p = T.burst_decode(h).as_dict()
        assert p == CO.burstparams(h)


def ramp_headers(lines, a0, k, pri=21859):
    hs = np.zeros((lines, 80), np.uint8)
    for i in range(lines):
        hs[i] = np.frombuffer(iw_header(pri=pri, azbeam=int(round(a0 + k * i)) & 1023), np.uint8)
    return hs


def test_steering_fit_matches_oracle_and_recovers_a_ramp():
    hs = ramp_headers(1500, 100, 0.5)
    c0, c1 = T.steering_fit(hs)
    assert (c0, c1) == CO.steering_fit(hs)
    # beam angle ramps by 0.5 address / line = 0.5 * 1.8 / 1024 deg per line;
    # per second that is times the PRF; c0 is the angle at the burst centre
    prf = 1717.13
    assert abs(c1 - 0.5 * 1.8 / 1024 * prf) < 2e-3 * abs(c1)
    assert abs(c0 - ((100 + 0.5 * 750) / 1024 * 1.8 - 0.9)) < 2e-3
    rng = np.random.default_rng(7)
    for _ in range(20):
        hs = ramp_headers(int(rng.integers(60, 400)), float(rng.uniform(0, 500)),
                          float(rng.uniform(-1, 1)), pri=int(rng.integers(15000, 30000)))
        assert T.steering_fit(hs) == CO.steering_fit(hs)


def test_steering_fit_needs_lines():
    with pytest.raises(T.TdbpError):
        T.steering_fit(ramp_headers(41, 0, 1))


def test_aperture_matches_caller_formula():
    rng = np.random.default_rng(3)
    for _ in range(100):
        v = rng.normal(0, 1, 3)
        v *= 7600 / np.linalg.norm(v)
        r0, prf, ac = float(rng.uniform(7e5, 9.5e5)), float(rng.uniform(1500, 1800)), float(
            rng.uniform(-0.5, 0.5))
        wvl = float(np.float32(0.05546576))
        a = T.aperture(v, r0, wvl, prf, ac)
        assert a == CO.aperture(v, r0, wvl, prf, ac)
        if ac == 0.0:
            assert a == synth.aperture_from_caller(r0, wvl, prf, v)
    # the scene generator's aperture is the same formula
    v = np.array([7000.0, 2000.0, 2500.0])
    assert T.aperture(v, 875e3, 0.0554657, 1717.0, 0.0) == synth.aperture_from_caller(
        875e3, 0.0554657, 1717.0, v)


def test_text_round_trip_rule():
    """The device's quantisation rule (integer part scaled exactly, fraction
    rounded) equals writing with f16.7 / f14.8 and reading back."""
    rng = np.random.default_rng(8)

    def q(x, s):
        xi = np.trunc(x)
        return (xi * s + np.rint((x - xi) * s)) / s

    for x in rng.uniform(-7.1e6, 7.1e6, 20000):
        assert q(x, 1e7) == CO.text_round_trip(x, "%16.7f")
    for x in rng.uniform(-8000, 8000, 20000):
        assert q(x, 1e8) == CO.text_round_trip(x, "%14.8f")
