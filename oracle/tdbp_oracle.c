/*
 * tdbp_oracle.c -- CPU restatement of the reference TDBP path (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library; the product library never links or calls it.
 *
 * Each function names the reference lines it restates.  Evaluation order of
 * every floating-point expression follows the reference source so that a
 * gcc -O2 build (no FMA contraction on baseline x86-64) reproduces the
 * compiled reference CPU path bit for bit.  Paths are relative to the
 * reference repository root.
 */
#include "tdbp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* WGS-84 as used by setxyz (sentinel/azimuth_compress_cpu.c:100-101). */
static const double ORC_WGS84_A = 6378137.0;
static const double ORC_WGS84_E2 = 0.0066943799901499996;

static double orc_pi(void) { return 4.0 * atan2(1.0, 1.0); }

/* Four-point Hermite interpolation of positions and velocities.
 * Restates sentinel/orbithermite.f:25-72 (NEC-memo form: f0, f1, g0, g1 built
 * from the Lagrange basis h and its derivative hdot). */
void orc_orbithermite(const double *x4, const double *v4, const double *t4,
                      double time, double *pos, double *vel) {
  double f0[4], f1[4], g0[4], g1[4], h[4], hd[4];
  for (int i = 0; i < 4; ++i) {
    double s = 0.0;
    f1[i] = time - t4[i];
    for (int j = 0; j < 4; ++j)
      if (j != i) s = s + 1.0 / (t4[i] - t4[j]);
    f0[i] = 1.0 - 2.0 * (time - t4[i]) * s;
  }
  for (int i = 0; i < 4; ++i) {
    double p = 1.0;
    for (int k = 0; k < 4; ++k)
      if (k != i) p = p * (time - t4[k]) / (t4[i] - t4[k]);
    h[i] = p;
    double s = 0.0;
    for (int j = 0; j < 4; ++j) {
      double q = 1.0;
      for (int k = 0; k < 4; ++k)
        if (k != i && k != j) q = q * (time - t4[k]) / (t4[i] - t4[k]);
      if (j != i) s = s + 1.0 / (t4[i] - t4[j]) * q;
    }
    hd[i] = s;
  }
  for (int i = 0; i < 4; ++i) {
    double s = 0.0;
    g1[i] = h[i] + 2.0 * (time - t4[i]) * hd[i];
    for (int j = 0; j < 4; ++j)
      if (i != j) s = s + 1.0 / (t4[i] - t4[j]);
    g0[i] = 2.0 * (f0[i] * hd[i] - h[i] * s);
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

/* Pick the 4 state vectors around `time` and interpolate.
 * Restates sentinel/intp_orbit.f90:6-20: ilocation = int((t-t1)/dt) clamped
 * to [2, n-2] (1-based), Hermite on vectors ilocation-1 .. ilocation+2. */
int orc_intp_orbit(const double *timeorbit, const double *xx, const double *vv,
                   int numstatevec, double time, double *pos, double *vel) {
  int iloc = (int)((time - timeorbit[0]) / (timeorbit[1] - timeorbit[0]));
  if (iloc < 2) iloc = 2;
  if (iloc > numstatevec - 2) iloc = numstatevec - 2;
  int first = iloc - 2; /* 0-based index of Fortran element ilocation-1 */
  orc_orbithermite(xx + 3 * first, vv + 3 * first, timeorbit + first, time,
                   pos, vel);
  return 0;
}

static double orc_dot3(const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < 3; ++i) s = s + a[i] * b[i];
  return s;
}

/* Zero-Doppler time and range of a ground point by Newton iteration on
 * f(t) = (x - s(t)).v(t) with f' ~ -|v|^2.  Restates
 * sentinel/orbitrangetime.f90:23-69 (at most 51 steps, stop at 5e-9 s). */
void orc_orbitrangetime(const double *xyz, const double *timeorbit,
                        const double *xx, const double *vv, int numstatevec,
                        double tline0, const double *satx0,
                        const double *satv0, double *tline, double *range) {
  double t = tline0, sx[3], sv[3], dr[3];
  memcpy(sx, satx0, sizeof sx);
  memcpy(sv, satv0, sizeof sv);
  for (int k = 1; k <= 51; ++k) {
    double tprev = t;
    for (int i = 0; i < 3; ++i) dr[i] = xyz[i] - sx[i];
    double fn = orc_dot3(dr, sv);
    double c1 = -orc_dot3(sv, sv);
    double c2 = 0.0;
    double fnprime = c1 + c2 * fn;
    t = t - fn / fnprime;
    orc_intp_orbit(timeorbit, xx, vv, numstatevec, t, sx, sv);
    if (fabs(t - tprev) < 5.0e-9) break;
  }
  for (int i = 0; i < 3; ++i) dr[i] = xyz[i] - sx[i];
  *tline = t;
  *range = sqrt(orc_dot3(dr, dr));
}

/* Geodetic (DEM grid) -> ECEF for pixels [firstpix, lastpix] of each line,
 * zero elsewhere, plus the three fit nodes per line.
 * Restates sentinel/azimuth_compress_cpu.c:89-152. */
void orc_setxyz(const int16_t *dem, double *xyz, double *xyzfit, int demwidth,
                double firstlat, double deltalat, double firstlon,
                double deltalon, int firstline, int nlines, int firstpix,
                int lastpix) {
  const double deg2rad = orc_pi() / 180.0;
  const int hstep = (int)((lastpix - firstpix) / 2 - 1);
  long total = (long)demwidth * (long)nlines;
#pragma omp parallel for schedule(static)
  for (long k = 0; k < total; ++k) {
    int line = (int)(k / demwidth);
    int pixel = (int)(k - (long)line * demwidth);
    double *o = xyz + 3 * k;
    o[0] = 0.0;
    o[1] = 0.0;
    o[2] = 0.0;
    if (pixel < firstpix || pixel > lastpix) continue;
    double lat = firstlat + (line + firstline) * deltalat;
    double phi = lat * deg2rad;
    double lam = (firstlon + pixel * deltalon) * deg2rad;
    double hgt = dem[k];
    double sphi = sin(phi);
    double re = ORC_WGS84_A / sqrt(1.0 - ORC_WGS84_E2 * sphi * sphi);
    o[0] = (re + hgt) * cos(phi) * cos(lam);
    o[1] = (re + hgt) * cos(phi) * sin(lam);
    o[2] = (re - (re * ORC_WGS84_E2) + hgt) * sphi;
    for (int node = 0; node < 3; ++node)
      if (pixel == firstpix + node * hstep)
        memcpy(xyzfit + 9 * line + 3 * node, o, 3 * sizeof(double));
  }
}

/* TOPS aperture-centre parabola per DEM line: for 3 nodes, the zero-Doppler
 * range from the Newton solve, Doppler seen from mid-burst, time offset td,
 * pulse index n = int(td*prf + L/2); coefficients of the parabola through
 * (-1,n1),(0,n2),(1,n3).  Restates sentinel/azimuth_compress_cpu.c:321-359
 * (the node look vector uses the unchanged mid-burst position and velocity,
 * because orbitrangetime does not write back satx0/satv0). */
void orc_fitcoef(const double *xyzfit, int nfit, const double *timeorbit,
                 const double *xx, const double *vv, int numstatevec,
                 double tmid, const double *xyz_mid, const double *vel_mid,
                 double wvl, double angc0, double angc1, double prf,
                 int rawdatalines, double *coef) {
  const double pi = orc_pi();
  for (int line = 0; line < nfit; ++line) {
    int y[3];
    for (int node = 0; node < 3; ++node) {
      const double *p = xyzfit + 9 * line + 3 * node;
      double tline, rngpix, u[3];
      orc_orbitrangetime(p, timeorbit, xx, vv, numstatevec, tmid, xyz_mid,
                         vel_mid, &tline, &rngpix);
      for (int i = 0; i < 3; ++i) u[i] = p[i] - xyz_mid[i];
      double umag = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      for (int i = 0; i < 3; ++i) u[i] = u[i] / umag;
      double udotv =
          vel_mid[0] * u[0] + vel_mid[1] * u[1] + vel_mid[2] * u[2];
      double fd = (2.0 / wvl) * udotv;
      double veff = sqrt(vel_mid[0] * vel_mid[0] + vel_mid[1] * vel_mid[1] +
                         vel_mid[2] * vel_mid[2]);
      double td = (rngpix * wvl / 2. / veff / veff * fd -
                   rngpix * angc0 * pi / 180. / veff) /
                  (1. + rngpix * angc1 * pi / 180. / veff);
      y[node] = (int)(td * prf + rawdatalines / 2);
    }
    coef[3 * line + 0] = 0.5 * (y[0] + y[2] - 2. * y[1]);
    coef[3 * line + 1] = 0.5 * (y[2] - y[0]);
    coef[3 * line + 2] = y[1];
  }
}

/* Aperture centre per pixel from the line parabola; -1 where it leaves the
 * valid band (A/2, L - A/2).  The abscissa is a float division, as in the
 * reference.  Restates sentinel/azimuth_compress_cpu.c:155-175; pixels
 * outside [firstpix, lastpix] are set to -1 (the reference leaves them). */
void orc_setazoff(const double *coef, double *azoff, int demwidth, int nlines,
                  int firstpix, int lastpix, int aperture, int rawdatalines) {
  for (int line = 0; line < nlines; ++line) {
    const double *c = coef + 3 * line;
    for (int pixel = 0; pixel < demwidth; ++pixel) {
      double *a = azoff + (long)line * demwidth + pixel;
      *a = -1.0;
      if (pixel < firstpix || pixel > lastpix) continue;
      double arg = (float)(pixel - firstpix) / (float)(lastpix - firstpix) *
                       2. - 1.;
      double fit = c[0] * arg * arg + c[1] * arg + c[2];
      if (fit > aperture / 2 && fit < rawdatalines - aperture / 2) *a = fit;
    }
  }
}

/* The coherent pulse sum per pixel.  Restates the CPU twin
 * sentinel/azimuth_compress_cpu.c:40-86 (FP64 lerp rounded to cf32, no 1/N)
 * or the GPU kernel sentinel/azimuth_compress.cu:33-82 (FP32 lerp weights and
 * products, output scaled by 1/(azend-azstart)).  Window:
 * azstart = round(azoff - A/2) >= 0, azend = min(round(azoff + A/2), L-1);
 * range gate rngstart <= R <= rngend - dmrg. */
void orc_pixelint(const float *burst, const double *satloc, const double *xyz,
                  const double *azoff, int demwidth, int aperture,
                  int rawdatalines, int samplesPerBurst, double rngstart,
                  double rngend, double dmrg, double wvl, float *out,
                  int nlines, int semantics, double *acc, int64_t *npairs,
                  int nthreads) {
  const double pi = orc_pi();
  const long total = (long)nlines * (long)demwidth;
  int64_t pairs = 0;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) reduction(+ : pairs) \
    num_threads(nthreads)
#endif
  for (long k = 0; k < total; ++k) {
    double ar = 0.0, ai = 0.0;
    int azs = 0, aze = 0;
    if (azoff[k] > 0.) {
      azs = (int)round(azoff[k] - aperture / 2);
      aze = (int)round(azoff[k] + aperture / 2);
      if (azs < 0) azs = 0;
      if (aze > rawdatalines - 1) aze = rawdatalines - 1;
      if (aze > azs) pairs += aze - azs;
      const double *x = xyz + 3 * k;
      for (int p = azs; p < aze; ++p) {
        double vx = x[0] - satloc[3 * p + 0];
        double vy = x[1] - satloc[3 * p + 1];
        double vz = x[2] - satloc[3 * p + 2];
        double range = sqrt(vx * vx + vy * vy + vz * vz);
        if (!(range >= rngstart && range <= rngend - dmrg)) continue;
        double r = (range - rngstart) / dmrg;
        int ir = (int)floor(r);
        double fr = r - ir;
        const float *b0 = burst + 2 * ((long)p * samplesPerBurst + ir);
        const float *b1 = b0 + 2;
        float cr, ci;
        if (semantics == ORC_GPU_SEMANTICS) {
          float w0 = (float)(1 - fr), w1 = (float)fr;
          cr = b0[0] * w0 + b1[0] * w1;
          ci = b0[1] * w0 + b1[1] * w1;
        } else {
          cr = (float)((double)b0[0] * (1 - fr) + (double)b1[0] * fr);
          ci = (float)((double)b0[1] * (1 - fr) + (double)b1[1] * fr);
        }
        double phase = 4. * pi / wvl * range;
        double pc = cos(phase), ps = sin(phase);
        double tr = (double)cr * pc - (double)ci * ps;
        double ti = (double)cr * ps + (double)ci * pc;
        ar = ar + tr;
        ai = ai + ti;
      }
    }
    if (acc) {
      acc[2 * k] = ar;
      acc[2 * k + 1] = ai;
    }
    float orr = (float)ar, oi = (float)ai;
    if (semantics == ORC_GPU_SEMANTICS) {
      if (azoff[k] > 0.) {
        float s = (float)(1. / (aze - azs));
        orr = orr * s;
        oi = oi * s;
      } else {
        orr = 0.f;
        oi = 0.f;
      }
    }
    out[2 * k] = orr;
    out[2 * k + 1] = oi;
  }
  if (npairs) *npairs = pairs;
}

/* Line and pixel window of a burst.  Restates
 * sentinel/azimuth_compress_cpu.c:246-251 and :283-284. */
void orc_window(const double *latlons, double firstlat, double firstlon,
                double deltalat, double deltalon, int demwidth, int demlength,
                int *firstline, int *lastline, int *firstpix, int *lastpix) {
  int ll = (int)((latlons[0] - firstlat) / deltalat);
  int fl = (int)((latlons[1] - firstlat) / deltalat);
  if (fl < 0) fl = 0;
  if (ll < 0) ll = 0;
  if (fl > demlength - 1) fl = demlength - 1;
  if (ll > demlength - 1) ll = demlength - 1;
  int fp = (int)((latlons[2] - firstlon) / deltalon);
  if (fp < 0) fp = 0;
  int lp = (int)((latlons[3] - firstlon) / deltalon);
  if (lp > demwidth) lp = demwidth;
  *firstline = fl;
  *lastline = ll;
  *firstpix = fp;
  *lastpix = lp;
}

int64_t orc_backproject(const float *burst, const double *satloc,
                        const int16_t *dem, int demwidth, int firstline,
                        int nlines, int firstpix, int lastpix,
                        double firstlat, double firstlon, double deltalat,
                        double deltalon, const double *timeorbit,
                        const double *xx, const double *vv, int numstatevec,
                        double rngstart, double rngend, double tmid,
                        const double *xyz_mid, const double *vel_mid,
                        double dmrg, double wvl, int aperture, double angc0,
                        double angc1, double prf, int rawdatalines,
                        int samplesPerBurst, int nfit, int semantics,
                        float *out, double *acc, int nthreads) {
  long npx = (long)nlines * demwidth;
  double *xyz = (double *)calloc(3 * npx, sizeof(double));
  double *xyzfit = (double *)calloc(9 * (size_t)nlines, sizeof(double));
  double *coef = (double *)calloc(3 * (size_t)nlines, sizeof(double));
  double *azoff = (double *)malloc(npx * sizeof(double));
  int64_t pairs = 0;
  orc_setxyz(dem, xyz, xyzfit, demwidth, firstlat, deltalat, firstlon,
             deltalon, firstline, nlines, firstpix, lastpix);
  orc_fitcoef(xyzfit, nfit, timeorbit, xx, vv, numstatevec, tmid, xyz_mid,
              vel_mid, wvl, angc0, angc1, prf, rawdatalines, coef);
  orc_setazoff(coef, azoff, demwidth, nlines, firstpix, lastpix, aperture,
               rawdatalines);
  orc_pixelint(burst, satloc, xyz, azoff, demwidth, aperture, rawdatalines,
               samplesPerBurst, rngstart, rngend, dmrg, wvl, out, nlines,
               semantics, acc, &pairs, nthreads);
  free(xyz);
  free(xyzfit);
  free(coef);
  free(azoff);
  return pairs;
}

/* The .geo merge of one burst window: grid pixel k takes out[k] where the
 * merge test on Re(out[k]) passes.  GPU entry: fabs((double)re) > 1e-18
 * (sentinel/azimuth_compress.cu:527-531).  CPU twin: abs((int)re) > 1e-18
 * (sentinel/azimuth_compress_cpu.c:387), whose conversion is undefined in C
 * beyond the int range; the reference's own build (gcc without -O,
 * build_proc:10) runs x86's truncating conversion, which gives INT_MIN for
 * |re| >= 2^31 and NaN, and abs(INT_MIN) stays negative.  Written out here
 * without the undefined conversion: 1 <= |re| < 2^31 (the live reference
 * builds pin both ends; an optimising gcc build takes |re| >= 2^31 too). */
void orc_merge_geo(float *grid, const float *out, long npx, int semantics) {
  for (long k = 0; k < npx; ++k) {
    float re = out[2 * k];
    float a = fabsf(re);
    int take = semantics == ORC_CPU_SEMANTICS ? (a >= 1.f && a < 2147483648.f)
                                              : (fabs((double)re) > 1.e-18);
    if (take) {
      grid[2 * k] = out[2 * k];
      grid[2 * k + 1] = out[2 * k + 1];
    }
  }
}

/* Entry point with the reference file semantics: DEM rows read from fddem at
 * firstline*W*2, .geo rows read and written back at firstline*W*8, a pixel
 * overwritten where the merge test passes.  Restates
 * sentinel/azimuth_compress_cpu.c:244-392 (CPU: skip when nlines == 1,
 * int-abs merge) and sentinel/azimuth_compress.cu:295-537 (GPU: fabs > 1e-18
 * merge).  The fit covers lines [0, nlines-1) like the reference loop; with
 * nfit_last the last line is fitted too, otherwise its coefficients are 0. */
void orc_azimuth_compress(const float *burst, const double *satloc,
                          int rawdatalines, int samplesPerBurst, int demwidth,
                          int demlength, int fdout, int fddem, double deltalat,
                          double deltalon, double firstlat, double firstlon,
                          const double *latlons, const double *timeorbit,
                          const double *xx, const double *vv, int numstatevec,
                          double rngstart, double rngend, double tmid,
                          const double *xyz_mid, const double *vel_mid,
                          double dmrg, double wvl, int aperture, double angc0,
                          double angc1, double prf, int semantics,
                          int nfit_last) {
  int firstline, lastline, firstpix, lastpix;
  orc_window(latlons, firstlat, firstlon, deltalat, deltalon, demwidth,
             demlength, &firstline, &lastline, &firstpix, &lastpix);
  int nlines = lastline - firstline + 1;
  if (semantics == ORC_CPU_SEMANTICS && nlines == 1) return;
  long npx = (long)nlines * demwidth;
  int16_t *dem = (int16_t *)calloc(npx, sizeof(int16_t));
  float *out = (float *)calloc(2 * npx, sizeof(float));
  float *old = (float *)calloc(2 * npx, sizeof(float));
  if (pread(fddem, dem, npx * 2, (off_t)firstline * demwidth * 2) < 0)
    fprintf(stderr, "orc: dem read error\n");
  int nfit = nfit_last ? nlines : nlines - 1;
  orc_backproject(burst, satloc, dem, demwidth, firstline, nlines, firstpix,
                  lastpix, firstlat, firstlon, deltalat, deltalon, timeorbit,
                  xx, vv, numstatevec, rngstart, rngend, tmid, xyz_mid,
                  vel_mid, dmrg, wvl, aperture, angc0, angc1, prf,
                  rawdatalines, samplesPerBurst, nfit, semantics, out, NULL,
                  0);
  off_t off = (off_t)firstline * demwidth * 8;
  ssize_t got = pread(fdout, old, npx * 8, off);
  (void)got;
  orc_merge_geo(old, out, npx, semantics);
  if (pwrite(fdout, old, npx * 8, off) != (ssize_t)(npx * 8))
    fprintf(stderr, "orc: geo write error\n");
  free(dem);
  free(out);
  free(old);
}
