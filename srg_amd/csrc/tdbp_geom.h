// tdbp_geom.h -- FP64 burst geometry shared by host and device code.
//
// The per-burst prologue of the TDBP path: 4-point Hermite orbit
// interpolation, the zero-Doppler Newton solve, WGS-84 geodetic -> ECEF of a
// DEM sample, the TOPS aperture-centre node, the parabola per DEM line and
// the pulse window of a pixel.  Everything here is FP64 and written with FP
// contraction off so that the device results equal a non-FMA CPU evaluation
// in the same order (the integer aperture centres then agree exactly).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tdbp {

constexpr double kWgs84A = 6378137.0;               // azimuth_compress.cu:104
constexpr double kWgs84E2 = 0.0066943799901499996;  // azimuth_compress.cu:105
// pi as the reference forms it (4*atan2(1,1)); the literal is that value.
constexpr double kPi = 3.141592653589793115997963468544185161590576171875;
constexpr double kDeg2Rad = kPi / 180.0;

#define TDBP_HD __host__ __device__ __forceinline__

// Hermite interpolation through 4 samples of position (x4) and velocity
// (v4) at times t4; follows sentinel/orbithermite.f:25-72 (Lagrange basis h,
// its derivative hd, and the f0/f1/g0/g1 weights).
TDBP_HD void hermite4(const double* x4, const double* v4, const double* t4,
                      double time, double* pos, double* vel) {
#pragma clang fp contract(off)
  double f0[4], f1[4], g0[4], g1[4], h[4], hd[4], sinv[4];
  for (int i = 0; i < 4; ++i) {
    double s = 0.0;
    for (int j = 0; j < 4; ++j)
      if (j != i) s = s + 1.0 / (t4[i] - t4[j]);
    sinv[i] = s;
    f1[i] = time - t4[i];
    f0[i] = 1.0 - 2.0 * (time - t4[i]) * s;
  }
  for (int i = 0; i < 4; ++i) {
    double p = 1.0;
    for (int k = 0; k < 4; ++k)
      if (k != i) p = p * (time - t4[k]) / (t4[i] - t4[k]);
    h[i] = p;
    double s = 0.0;
    for (int j = 0; j < 4; ++j) {
      if (j == i) continue;
      double q = 1.0;
      for (int k = 0; k < 4; ++k)
        if (k != i && k != j) q = q * (time - t4[k]) / (t4[i] - t4[k]);
      s = s + 1.0 / (t4[i] - t4[j]) * q;
    }
    hd[i] = s;
  }
  for (int i = 0; i < 4; ++i) {
    g1[i] = h[i] + 2.0 * (time - t4[i]) * hd[i];
    g0[i] = 2.0 * (f0[i] * hd[i] - h[i] * sinv[i]);
  }
  for (int k = 0; k < 3; ++k) {
    double sp = 0.0, sv = 0.0;
    for (int i = 0; i < 4; ++i)
      sp = sp + (x4[3 * i + k] * f0[i] + v4[3 * i + k] * f1[i]) * h[i] * h[i];
    for (int i = 0; i < 4; ++i)
      sv = sv + (x4[3 * i + k] * g0[i] + v4[3 * i + k] * g1[i]) * h[i];
    pos[k] = sp;
    vel[k] = sv;
  }
}

// State-vector selection of sentinel/intp_orbit.f90:6-20: 1-based index
// int((t - t1)/dt) clamped to [2, n-2], Hermite over the 4 vectors starting
// one before it.  t/x/v are the orbit arrays (x, v as [n][3]).
TDBP_HD void orbit_at(const double* t, const double* x, const double* v, int n,
                      double time, double* pos, double* vel) {
#pragma clang fp contract(off)
  int iloc = (int)((time - t[0]) / (t[1] - t[0]));
  iloc = iloc < 2 ? 2 : iloc;
  iloc = iloc > n - 2 ? n - 2 : iloc;
  const int first = iloc - 2;
  hermite4(x + 3 * first, v + 3 * first, t + first, time, pos, vel);
}

TDBP_HD double dot3(const double* a, const double* b) {
#pragma clang fp contract(off)
  double s = 0.0;
  for (int i = 0; i < 3; ++i) s = s + a[i] * b[i];
  return s;
}

// Zero-Doppler range of a ground point: Newton on (x - s(t)).v(t) with the
// derivative taken as -|v|^2, at most 51 steps, stop at |dt| < 5e-9 s
// (sentinel/orbitrangetime.f90:23-69).  s0/v0 are the starting state and are
// not modified.
TDBP_HD double zero_doppler_range(const double* xyz, const double* t,
                                  const double* x, const double* v, int n,
                                  double t0, const double* s0,
                                  const double* v0) {
#pragma clang fp contract(off)
  double tl = t0, sx[3] = {s0[0], s0[1], s0[2]}, sv[3] = {v0[0], v0[1], v0[2]};
  double dr[3];
  for (int k = 0; k < 51; ++k) {
    const double tprev = tl;
    for (int i = 0; i < 3; ++i) dr[i] = xyz[i] - sx[i];
    const double fn = dot3(dr, sv);
    const double fnprime = -dot3(sv, sv) + 0.0 * fn;
    tl = tl - fn / fnprime;
    orbit_at(t, x, v, n, tl, sx, sv);
    if (fabs(tl - tprev) < 5.0e-9) break;
  }
  for (int i = 0; i < 3; ++i) dr[i] = xyz[i] - sx[i];
  return sqrt(dot3(dr, dr));
}

// WGS-84 geodetic -> ECEF of DEM sample (window line `line`, column `pixel`)
// as in sentinel/azimuth_compress.cu:117-131.
// sqrt x from an estimate y0 of 1/sqrt x (relative error up to ~1e-7) for
// normal x > 0: Goldschmidt's iteration (g -> sqrt x, h -> 1/(2 sqrt x))
// and two residual corrections; correctly rounded on the host checks.
TDBP_HD double goldschmidt_sqrt(double x, double y0) {
  double g = x * y0, h = 0.5 * y0;
  const double r = fma(-h, g, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  double d = fma(-g, g, x);
  g = fma(d, h, g);
  d = fma(-g, g, x);
  return fma(d, h, g);
}

// sin and cos of 2 pi f for |f| <= 1/2 in FP64: t = 4f = n + d exactly
// (n = rint t), th = d pi/2 in [-pi/4, pi/4], Taylor series to th^13 (sin)
// and th^12 (cos), truncation below 1e-11; the quadrant n mod 4 swaps and
// negates.
TDBP_HD void sincos_2pi(double f, double* sn, double* cs) {
  const double t = 4.0 * f;
  const double n = rint(t);
  const double th = (t - n) * 1.5707963267948966;
  const double z = th * th;
  double ps = -1.0 / 39916800.0;
  ps = fma(z, ps, 1.0 / 362880.0);
  ps = fma(z, ps, -1.0 / 5040.0);
  ps = fma(z, ps, 1.0 / 120.0);
  ps = fma(z, ps, -1.0 / 6.0);
  const double s0 = fma(th * z, ps, th);
  double pc = 1.0 / 479001600.0;
  pc = fma(z, pc, -1.0 / 3628800.0);
  pc = fma(z, pc, 1.0 / 40320.0);
  pc = fma(z, pc, -1.0 / 720.0);
  pc = fma(z, pc, 1.0 / 24.0);
  pc = fma(z, pc, -0.5);
  const double c0 = fma(z, pc, 1.0);
  const int q = (int)n & 3;   // angle = q pi/2 + th
  const double sa = (q & 1) ? c0 : s0, ca = (q & 1) ? s0 : c0;
  *sn = (q & 2) ? -sa : sa;
  *cs = ((q + 1) & 2) ? -ca : ca;
}

TDBP_HD void dem_to_ecef(int line, int firstline, int pixel, double firstlat,
                         double deltalat, double firstlon, double deltalon,
                         double hgt, double* o) {
#pragma clang fp contract(off)
  const double lat = firstlat + (line + firstline) * deltalat;
  const double phi = lat * kDeg2Rad;
  const double lam = (firstlon + pixel * deltalon) * kDeg2Rad;
  // sincos: the library's sin and cos values from one shared argument
  // reduction (the compiler does not merge separate sin and cos calls).
  double sphi, cphi, slam, clam;
  sincos(phi, &sphi, &cphi);
  sincos(lam, &slam, &clam);
  const double re = kWgs84A / sqrt(1.0 - kWgs84E2 * sphi * sphi);
  o[0] = (re + hgt) * cphi * clam;
  o[1] = (re + hgt) * cphi * slam;
  o[2] = (re - (re * kWgs84E2) + hgt) * sphi;
}

// Column step between the three fit nodes of a line
// (sentinel/azimuth_compress.cu:138, :441).
TDBP_HD int fit_node_step(int firstpix, int lastpix) {
  return (int)((lastpix - firstpix) / 2 - 1);
}

struct FitParams {
  double tmid, wvl, angc0, angc1, prf;
  double xyz_mid[3], vel_mid[3];
  int rawdatalines;
};

// Pulse index of the aperture centre of ground point p
// (sentinel/azimuth_compress.cu:451-465): zero-Doppler range from the Newton
// solve started at mid-burst, Doppler of p seen from the mid-burst state,
// TOPS time offset td, n = int(td*prf + L/2).
TDBP_HD int aperture_centre(const double* p, const double* t, const double* x,
                            const double* v, int n, const FitParams& f) {
#pragma clang fp contract(off)
  const double rngpix = zero_doppler_range(p, t, x, v, n, f.tmid, f.xyz_mid,
                                           f.vel_mid);
  double u[3];
  for (int i = 0; i < 3; ++i) u[i] = p[i] - f.xyz_mid[i];
  const double umag = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  for (int i = 0; i < 3; ++i) u[i] = u[i] / umag;
  const double udotv =
      f.vel_mid[0] * u[0] + f.vel_mid[1] * u[1] + f.vel_mid[2] * u[2];
  const double fd = (2.0 / f.wvl) * udotv;
  const double veff = sqrt(f.vel_mid[0] * f.vel_mid[0] +
                           f.vel_mid[1] * f.vel_mid[1] +
                           f.vel_mid[2] * f.vel_mid[2]);
  const double td = (rngpix * f.wvl / 2. / veff / veff * fd -
                     rngpix * f.angc0 * kPi / 180. / veff) /
                    (1. + rngpix * f.angc1 * kPi / 180. / veff);
  return (int)(td * f.prf + f.rawdatalines / 2);
}

// Parabola through (-1, y1), (0, y2), (1, y3) (azimuth_compress.cu:472-474).
TDBP_HD void parabola(int y1, int y2, int y3, double* c) {
#pragma clang fp contract(off)
  c[0] = 0.5 * (y1 + y3 - 2. * y2);
  c[1] = 0.5 * (y3 - y1);
  c[2] = y2;
}

// Aperture centre of a pixel from its line parabola, -1 outside the valid
// band (A/2, L - A/2); the abscissa is a float division as in
// sentinel/azimuth_compress.cu:170-175.
TDBP_HD double pixel_azoff(const double* c, int pixel, int firstpix,
                           int lastpix, int aperture, int rawdatalines) {
#pragma clang fp contract(off)
  const double arg =
      (float)(pixel - firstpix) / (float)(lastpix - firstpix) * 2. - 1.;
  const double fit = c[0] * arg * arg + c[1] * arg + c[2];
  return (fit > aperture / 2 && fit < rawdatalines - aperture / 2) ? fit : -1.0;
}

// Pulse window [azs, aze) of a pixel with azoff > 0
// (sentinel/azimuth_compress.cu:55-58).
TDBP_HD void pulse_window(double azoff, int aperture, int rawdatalines,
                          int* azs, int* aze) {
  int s = (int)round(azoff - aperture / 2);
  int e = (int)round(azoff + aperture / 2);
  *azs = s < 0 ? 0 : s;
  *aze = e > rawdatalines - 1 ? rawdatalines - 1 : e;
}

}  // namespace tdbp
