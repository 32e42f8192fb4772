"""The C-ABI boundary on the CPU: the library loads, exports every symbol the
header declares, its host-side functions agree with the oracle, and a Fortran
caller links against it through the interface module.  No GPU compute here:
device entry points must fail loudly (an error code), never fall back."""
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from srg_amd import synth, tdbp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "tdbp.h")


def header_functions():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", txt)
    skip = {"if", "sizeof", "defined"}
    return sorted({n for n in names if n not in skip and (n.startswith("tdbp_") or
                                                          n in ("azimuth_compress_", "azimuth_compress_cpu_"))})


def test_library_exports_every_declared_symbol():
    lib = tdbp.lib()
    fns = header_functions()
    assert "azimuth_compress_" in fns and len(fns) > 30
    missing = [f for f in fns if not hasattr(lib, f)]
    assert not missing, missing
    # the Python binding declares the same set
    assert set(tdbp.exported_symbols()) == set(fns)


def test_exported_dynamic_symbols_with_nm():
    nm = shutil.which("nm")
    if not nm:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", tdbp.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for f in header_functions():
        assert f in syms, f


def test_window_matches_oracle(oracle):
    rng = np.random.default_rng(0)
    for _ in range(200):
        W, Lg = int(rng.integers(50, 5000)), int(rng.integers(50, 5000))
        fl, fo = 40.0, -120.0
        dl, do = -abs(rng.uniform(1e-5, 3e-4)), abs(rng.uniform(1e-5, 3e-4))
        lat = sorted(rng.uniform(fl + 1.2 * Lg * dl, fl - 0.2 * Lg * dl, 2))
        lon = sorted(rng.uniform(fo - 0.2 * W * do, fo + 1.2 * W * do, 2))
        ll = np.array([lat[0], lat[1], lon[0], lon[1]])
        assert tdbp.window(ll, fl, fo, dl, do, W, Lg) == oracle.window(ll, fl, fo, dl, do, W, Lg)


def test_orbit_interp_bit_exact_with_oracle(oracle):
    sc = synth.small_burst()
    g = sc.geom
    times = np.linspace(g.timeorbit[0] - 2, g.timeorbit[-1] + 2, 97)
    pos, vel = tdbp.orbit_interp(g.timeorbit, g.xx, g.vv, times)
    for t, p, v in zip(times, pos, vel):
        po, vo = oracle.intp_orbit(g.timeorbit, g.xx, g.vv, t)
        np.testing.assert_array_equal(p, po)
        np.testing.assert_array_equal(v, vo)


def test_geom_struct_layout_matches_header():
    import ctypes as C
    # offsets the C compiler gives tdbp_geom (natural alignment)
    g = tdbp.Geom
    assert g.rngstart.offset == 16
    assert g.xyz_mid.offset == 80
    assert g.timeorbit.offset == 128
    assert g.demwidth.offset == 152
    assert g.firstlat.offset == 176
    assert g.fit_last_line.offset == 208
    assert C.sizeof(g) == 216


def test_device_entry_points_fail_loudly_without_gpu():
    lib = tdbp.lib()
    if lib.tdbp_device_count() > 0:
        pytest.skip("a GPU is visible; covered by the gpu tests")
    assert lib.tdbp_init(0) == tdbp.ERR_NODEV
    assert lib.tdbp_dev_alloc(1024) is None
    with pytest.raises(tdbp.TdbpError):
        tdbp.init(0)


def test_fortran_interface_links():
    fc = shutil.which("amdflang") or (os.path.exists("/opt/rocm/bin/amdflang") and
                                      "/opt/rocm/bin/amdflang")
    if not fc:
        pytest.skip("no Fortran compiler")
    src = """
program call_tdbp
  use tdbp_iface
  implicit none
  complex(4) :: burst(16)
  real(8) :: satloc(3, 4), latlons(4), to(4), xx(3, 4), vv(3, 4), xm(3), vm(3), t(4)
  integer :: n
  n = 0
  if (n > 0) then
     call azimuth_compress(burst, satloc, n, n, n, n, n, n, 0d0, 0d0, 0d0, 0d0, latlons, &
          to, xx, vv, n, 0d0, 0d0, 0d0, 0d0, 0d0, xm, vm, t, 0d0, 0d0, 0d0, n, n, &
          0d0, 0d0, 0d0)
     call azimuth_compress_cpu(burst, satloc, n, n, n, n, n, n, 0d0, 0d0, 0d0, 0d0, &
          latlons, to, xx, vv, n, 0d0, 0d0, 0d0, 0d0, 0d0, xm, vm, t, 0d0, 0d0, 0d0, n, &
          n, 0d0, 0d0, 0d0)
  end if
  print *, tdbp_dropin_status()
end program call_tdbp
"""
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "caller.f90")
        open(f, "w").write(src)
        exe = os.path.join(td, "caller")
        r = subprocess.run([fc, "-o", exe, os.path.join(ROOT, "include", "tdbp_iface.f90"), f,
                            "-L" + os.path.dirname(tdbp.LIB_PATH), "-lsrg_tdbp",
                            "-Wl,-rpath," + os.path.dirname(tdbp.LIB_PATH)],
                           cwd=td, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0 and r.stdout.strip() == "0"


def test_howmanygpus_prints_one_integer():
    """sentinel_gpu.py:98-100 reads the first word of its stdout as the GPU
    count; without a GPU the query fails and the tool prints 0."""
    exe = os.path.join(ROOT, "srg_amd", "bin", "howmanygpus")
    assert os.access(exe, os.X_OK), "build first (make -C srg_amd/csrc)"
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0
    words = out.stdout.split()
    assert len(words) == 1 and int(words[0]) >= 0
