// tdbp_caller.hip -- the per-burst parameter derivation of the reference's
// Fortran burst driver, so that a burst can be set up from its decoded lines
// and the orbit state vectors without the Fortran side:
//   * burst parameters from the first line's header (burstparams,
//     sentinel/backprojectgpusub.f90:406-439);
//   * TOPS steering coefficients from every line's azimuth beam address
//     (fdopcoefs + linfit, :262-300 and :355-383);
//   * the integration length (:193-209) and the range window (:220-224);
//   * the ground footprint latlons (sentinel/bounds.f90:1-165 with geo_hdg
//     :167-318, latlon.f90, radar_to_xyz.f90 / curvature.f90 rdir,
//     tcnbasis.f90, unitvec.f90, cross.f90);
//   * per-pulse satellite positions from the header times
//     (sentineltimingsub.f90:44-69), on the device.
// Host code keeps the reference's evaluation order and storage types
// (FP32 where the Fortran variables are default reals); contraction is off.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "tdbp_api_internal.h"
#include "tdbp_geom.h"
#include "tdbp_internal.h"

#pragma clang fp contract(off)

using namespace tdbp;

namespace {

constexpr double kFrefD = 37.53472224;           // burstparams: 37.53472224d0
constexpr double kSol = 299792458.0;

int be16(const uint8_t* b) { return (int)b[1] + 256 * (int)b[0]; }
int be24(const uint8_t* b) { return (int)b[2] + 256 * (int)b[1] + 65536 * (int)b[0]; }

double d8(const uint8_t* b) {
  double d;
  std::memcpy(&d, b, 8);
  return d;
}

// processsub.f90:181-212 (the same table as the range compression's).
double samplefrequency_mhz(int code) {
  float r;
  switch (code) {
    case 0: r = 3.f / 4.f; break;
    case 1: r = 2.f / 3.f; break;
    case 3: r = 5.f / 9.f; break;
    case 4: r = 4.f / 9.f; break;
    case 5: r = 3.f / 8.f; break;
    case 6: r = 1.f / 3.f; break;
    case 7: r = 1.f / 6.f; break;
    case 8: r = 3.f / 7.f; break;
    case 9: r = 5.f / 16.f; break;
    case 10: r = 3.f / 26.f; break;
    case 11: r = 4.f / 11.f; break;
    default: return 0.0;
  }
  const float r4 = r * 4.f;
  return (double)r4 * kFrefD;
}

// ---------------------------------------------------------------- geodesy

// bounds.f90:44-47 and tcnbasis.f90:29-30 set the ellipsoid from default-real
// literals, so e2 is the FP32 rounding of 0.0066943799901499996.
struct Elp {
  double a = (double)6378137.0f, e2 = (double)0.0066943799901499996f;
};

void cross3(const double* u, const double* v, double* w) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}

void unitvec(const double* v, double* u) {
  const double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (n != 0.0) {
    u[0] = v[0] / n;
    u[1] = v[1] / n;
    u[2] = v[2] / n;
  }
}

// latlon.f90 i_type 2: ECEF -> (lat, lon, h), radians.
void xyz_to_llh(const Elp& e, const double* v, double* llh) {
  const double q2 = 1.0 / (1.0 - e.e2);
  const double q = std::sqrt(q2);
  const double q3 = q2 - 1.0;
  const double b = e.a * std::sqrt(1.0 - e.e2);
  llh[1] = std::atan2(v[1], v[0]);
  const double p = std::sqrt(v[0] * v[0] + v[1] * v[1]);
  double tant = (v[2] / p) * q;
  const double theta = std::atan(tant);
  const double st = std::sin(theta), ct = std::cos(theta);
  tant = (v[2] + q3 * b * (st * st * st)) / (p - e.e2 * e.a * (ct * ct * ct));
  llh[0] = std::atan(tant);
  const double sl = std::sin(llh[0]);
  const double re = e.a / std::sqrt(1.0 - e.e2 * (sl * sl));
  llh[2] = p / std::cos(llh[0]) - re;
}

// curvature.f90 rdir: radius of curvature along heading hdg at latitude lat.
double rdir(double a, double e2, double hdg, double lat) {
  const double sl = std::sin(lat);
  const double reast = a / std::sqrt(1.0 - e2 * (sl * sl));
  const double rnorth = (a * (1.0 - e2)) / std::pow(1.0 - e2 * (sl * sl), 1.5);
  const double ch = std::cos(hdg), sh = std::sin(hdg);
  return (reast * rnorth) / (reast * (ch * ch) + rnorth * (sh * sh));
}

// tcnbasis.f90: track / cross / normal unit vectors at pos with velocity vel.
void tcnbasis(const double* pos, const double* vel, double* t, double* c, double* n) {
  Elp e;
  double llh[3], tmp[3];
  xyz_to_llh(e, pos, llh);
  n[0] = -std::cos(llh[0]) * std::cos(llh[1]);
  n[1] = -std::cos(llh[0]) * std::sin(llh[1]);
  n[2] = -std::sin(llh[0]);
  cross3(n, vel, tmp);
  unitvec(tmp, c);
  cross3(c, n, tmp);
  unitvec(tmp, t);
}

double fsign(double a, double b) { return b >= 0.0 ? std::fabs(a) : -std::fabs(a); }

// bounds.f90 geo_hdg, ellipsoid branch (e2 != 0): geodesic heading from
// (lati, loni) to (latf, lonf), radians.
double geo_hdg(double a, double e2, double lati, double loni, double latf, double lonf) {
  const double pi = 3.141592653589793;
  const double e = std::sqrt(e2);
  const double ome2 = 1.0 - e2;
  const double sqrtome2 = std::sqrt(ome2);
  const double f = 1.0 - sqrtome2;
  const double ep = e * f / (e2 - f);
  const double n = f / e2;
  const double k1 = (16.0 * e2 * (n * n) + ep * ep) / (ep * ep);
  const double k2 = (16.0 * e2 * (n * n)) / (16.0 * e2 * (n * n) + ep * ep);
  const double k3 = (16.0 * e2 * (n * n)) / (ep * ep);
  const double k4 = (16.0 * n - ep * ep) / (16.0 * e2 * (n * n) + ep * ep);
  const double k5 = 16.0 / (e2 * (16.0 * e2 * (n * n) + ep * ep));
  const double tanlati = std::tan(lati), tanlatf = std::tan(latf);
  const double l = std::fabs(lonf - loni);
  double lsign = lonf - loni;
  if (std::fabs(lsign) > pi) lsign = -(2.0 * pi - l) * fsign(1.0, -lsign);
  double sinlon = std::sin(l);
  const double coslon = std::cos(l);
  const double tanbetai = sqrtome2 * tanlati;
  const double tanbetaf = sqrtome2 * tanlatf;
  const double cosbetai = 1.0 / std::sqrt(1.0 + tanbetai * tanbetai);
  const double cosbetaf = 1.0 / std::sqrt(1.0 + tanbetaf * tanbetaf);
  const double sinbetai = tanbetai * cosbetai;
  const double sinbetaf = tanbetaf * cosbetaf;
  const double ac = sinbetai * sinbetaf;
  const double bc = cosbetai * cosbetaf;
  const double cosphi = ac + bc * coslon;
  const double sinphi = fsign(1.0, sinlon) * std::sqrt(1.0 - std::fmin(cosphi * cosphi, 1.0));
  const double phi = std::fabs(std::atan2(sinphi, cosphi));
  if (!(a * std::fabs(phi) > 1.0e-6)) return 0.0;
  const double ca = (bc * sinlon) / sinphi;
  const double cb = ca * ca;
  const double cc = (cosphi * (1.0 - cb)) / k1;
  const double cd = (-2.0 * ac) / k1;
  const double ce = -ac * k2;
  const double cf = k3 * cc;
  const double cg = (phi * phi) / sinphi;
  const double x = ((phi * (k4 + cb) + sinphi * (cc + cd) + cg * (cf + ce)) * ca) / k5;
  const double lambda = l + x;
  const double sinlam = std::sin(lambda), coslam = std::cos(lambda);
  const double t2 = (tanbetaf * cosbetai - coslam * sinbetai);
  sinlon = sinlam * fsign(1.0, lsign);
  return std::atan2(sinlon, t2);
}

}  // namespace

// ------------------------------------------------------------------ kernel

namespace {

struct Svs {
  double t[kMaxStateVec], x[kMaxStateVec][3], v[kMaxStateVec][3];
};

// sentineltimingsub.f90:55-66: line time from header bytes 68-75, state
// vector index int((t - t1)/dt), Hermite on the four vectors from index - 1
// (1-based); no clamp in the reference, so indices outside [2, n-2] are
// counted as errors (the line is still written, clamped).  quantize: the
// write with f16.7 / f14.8 and read back of the position file.
__global__ void k_pulse_orbit(const uint32_t* __restrict__ lines, int nlines, int stride_words,
                              const Svs* __restrict__ sv, int nsv, int quantize,
                              double* __restrict__ sat, double* __restrict__ vel,
                              double* __restrict__ times, int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlines) return;
  const uint32_t* h = lines + (size_t)i * stride_words;
  const uint64_t bits = (uint64_t)h[17] | ((uint64_t)h[18] << 32);   // bytes 68..75
  const double t = __longlong_as_double((long long)bits);
  const double dt = sv->t[1] - sv->t[0];
  const double q = (t - sv->t[0]) / dt;
  int il = (q > -2147483648.0 && q < 2147483647.0) ? (int)q : -1;   // Fortran INT
  if (il < 2 || il > nsv - 2) {
    atomicAdd(bad, 1);
    il = il < 2 ? 2 : nsv - 2;
  }
  double x[3], v[3];
  const int k = il - 2;   // 0-based first of the four vectors
  hermite4(&sv->x[k][0], &sv->v[k][0], &sv->t[k], t, x, v);
  if (quantize) {
    // Decimal rounding of the exact binary value: the integer part scaled is
    // exact, only the fraction's product is rounded (to ~1e-9 of a unit).
    for (int c = 0; c < 3; ++c) {
      const double xi = trunc(x[c]), vi = trunc(v[c]);
      x[c] = (xi * 1.0e7 + rint((x[c] - xi) * 1.0e7)) / 1.0e7;
      v[c] = (vi * 1.0e8 + rint((v[c] - vi) * 1.0e8)) / 1.0e8;
    }
  }
  for (int c = 0; c < 3; ++c) {
    sat[(size_t)i * 3 + c] = x[c];
    vel[(size_t)i * 3 + c] = v[c];
  }
  times[i] = t;
}

}  // namespace

extern "C" {

int tdbp_burst_decode(const uint8_t* h, tdbp_burst_params* p) {
  if (!h || !p) return set_error(TDBP_ERR_ARG, "burst params: null pointer");
  std::memset(p, 0, sizeof *p);
  p->starttime = d8(h + 68);
  const int pri_code = be24(h + 50);
  if (pri_code == 0) return set_error(TDBP_ERR_ARG, "burst params: PRI code 0");
  p->prf = kFrefD / pri_code * 1.e6;
  p->rangedecimation = h[40];
  const double fs = samplefrequency_mhz(p->rangedecimation);
  if (fs == 0.0)
    return set_error(TDBP_ERR_ARG, "range decimation code %d has no sampling rate",
                     p->rangedecimation);
  p->samplefreq = fs * 1.e6;
  // in3(...)/fref*1.e-6: 1.e-6 is a default-real literal (FP32-rounded).
  const double swst = be24(h + 53) / kFrefD * (double)1.e-6f;
  p->rank = h[49] & 31;
  p->pri = pri_code / kFrefD * (double)1.e-6f;
  // c/2.*(irank*pri+swst)+60.: 2. and 60. are default reals, exact.
  p->range0 = kSol / 2.0 * (p->rank * p->pri + swst) + 60.0;
  p->swst = swst;
  p->wvl = (double)0.05546576f;   // wvl=0.05546576, a default-real literal
  return TDBP_OK;
}

int tdbp_steering_fit(const uint8_t* headers, size_t stride_bytes, int lines, double* angc0,
                      double* angc1) {
  if (!headers || !angc0 || !angc1 || stride_bytes < 80 || lines < 42)
    return set_error(TDBP_ERR_ARG, "steering fit: %d lines (needs > 41)", lines);
  // fdopcoefs: fref=37.53472224 is a default-real literal assigned to real*8.
  const double fref = (double)37.53472224f;
  const int n = lines - 40;
  double sumx = 0.0, sumx2 = 0.0, sumxy = 0.0, sumy = 0.0;
  for (int i = 21; i <= lines - 20; ++i) {            // linfit(t(21), ang(21), lines-40)
    const uint8_t* b = headers + (size_t)(i - 1) * stride_bytes;
    const int word = be16(b + 60);
    const int azb = word & 1023;
    const float ang = (float)azb / 1024.f * 1.8f - 0.9f;
    const int pc = be24(b + 50);
    if (pc == 0) return set_error(TDBP_ERR_ARG, "steering fit: PRI code 0 on line %d", i);
    const float prf = (float)(fref / pc * 1.e6);
    const float t = ((float)i - (float)lines / 2.f) / prf;
    // linfit: the products of two default reals are FP32 products.
    sumx = sumx + (double)t;
    sumx2 = sumx2 + (double)(t * t);
    sumxy = sumxy + (double)(t * ang);
    sumy = sumy + (double)ang;
  }
  const double den = n * sumx2 - sumx * sumx;
  if (den == 0.0) return set_error(TDBP_ERR_ARG, "steering fit: degenerate times");
  const float c1 = (float)((n * sumxy - sumx * sumy) / den);
  const float c0 = (float)((sumy * sumx2 - sumx * sumxy) / den);
  *angc0 = c0;
  *angc1 = c1;
  return TDBP_OK;
}

int tdbp_aperture(const double* vel_mid_burst, double range0, double wvl, double prf,
                  double angc1) {
  const double pi = 4.0 * std::atan2(1.0, 1.0);
  const double re = 6378137.0, ht = 700000.0;
  double vmag = std::sqrt(vel_mid_burst[0] * vel_mid_burst[0] +
                          vel_mid_burst[1] * vel_mid_burst[1] +
                          vel_mid_burst[2] * vel_mid_burst[2]);
  vmag = vmag * (re / (re + ht));
  const double aperturetime =
      (range0 * wvl / 2 / vmag / vmag * prf) / (1.0 + range0 * angc1 * pi / 180.0 / vmag);
  return (int)(aperturetime * prf * 0.5);
}

int tdbp_bounds(double startime, double endtime, double startr, double endr,
                const double* timeorbit, const double* xx, const double* vv, int nsv,
                double* latlons) {
  if (!timeorbit || !xx || !vv || !latlons || nsv < 4 || nsv > kMaxStateVec)
    return set_error(TDBP_ERR_ARG, "bounds: %d state vectors", nsv);
  const double pi = 4.0 * std::atan2(1.0, 1.0);
  const double r2d = 180.0 / pi;
  Elp e;
  double xs[3], vs[3], xe[3], ve[3], xm[3], vm[3];
  orbit_at(timeorbit, xx, vv, nsv, startime, xs, vs);
  orbit_at(timeorbit, xx, vv, nsv, endtime, xe, ve);
  orbit_at(timeorbit, xx, vv, nsv, (startime + endtime) / 2.0, xm, vm);
  double llh[3];
  xyz_to_llh(e, xs, llh);
  const double lati = llh[0], loni = llh[1];
  double height = llh[2];
  xyz_to_llh(e, xe, llh);
  const double latf = llh[0], lonf = llh[1];
  height = (height + llh[2]) / 2.0;
  const double hdg = geo_hdg(e.a, e.e2, lati, loni, latf, lonf);
  xyz_to_llh(e, xm, llh);
  const double rcurv = rdir(e.a, e.e2, hdg, llh[0]);
  const double hgts[2] = {-100.0, 2000.0};
  double min_lat = 1.e10, min_lon = 1.e10, max_lat = -1.e10, max_lon = -1.e10;
  for (int line = 0; line < 2; ++line) {
    const double* xsat = line == 0 ? xs : xe;
    const double* vsat = line == 0 ? vs : ve;
    double vhat[3], that[3], chat[3], nhat[3];
    unitvec(vsat, vhat);
    tcnbasis(xsat, vsat, that, chat, nhat);
    for (int ind = 0; ind < 2; ++ind) {
      const double rng = ind == 0 ? startr : endr;
      for (int it = 0; it < 2; ++it) {
        const double zsch = hgts[it];
        const double aa = height + rcurv;
        const double bb = rcurv + zsch;
        const double costheta = 0.5 * ((aa / rng) + (rng / aa) - (bb / aa) * (bb / rng));
        const double sintheta = std::sqrt(1.0 - costheta * costheta);
        const double gamm = costheta * rng;
        const double nv = nhat[0] * vhat[0] + nhat[1] * vhat[1] + nhat[2] * vhat[2];
        const double vt = vhat[0] * that[0] + vhat[1] * that[1] + vhat[2] * that[2];
        const double alpha = (0.0 - gamm * nv) / vt;
        const double beta = std::sqrt(rng * rng * sintheta * sintheta - alpha * alpha);
        double xyz[3];
        for (int c = 0; c < 3; ++c)
          xyz[c] = xsat[c] + (gamm * nhat[c] + alpha * that[c] + beta * chat[c]);
        xyz_to_llh(e, xyz, llh);
        min_lat = std::fmin(min_lat, llh[0] * r2d);
        max_lat = std::fmax(max_lat, llh[0] * r2d);
        min_lon = std::fmin(min_lon, llh[1] * r2d);
        max_lon = std::fmax(max_lon, llh[1] * r2d);
      }
    }
  }
  latlons[0] = min_lat;
  latlons[1] = max_lat;
  latlons[2] = min_lon;
  latlons[3] = max_lon;
  return TDBP_OK;
}

int tdbp_burst_setup(const uint8_t* headers, size_t header_stride, int lines, int rangesamples,
                     const double* satvel_mid, const double* timeorbit, const double* xx,
                     const double* vv, int nsv, int demwidth, int demlength, double firstlat,
                     double firstlon, double deltalat, double deltalon, tdbp_geom* g,
                     double* latlons, int* in_scene) {
  if (!headers || !satvel_mid || !timeorbit || !xx || !vv || !g || !latlons || !in_scene ||
      lines < 42 || rangesamples < 2 || header_stride < 80)
    return set_error(TDBP_ERR_ARG, "burst setup: bad arguments");
  int st;
  tdbp_burst_params bp;
  if ((st = tdbp_burst_decode(headers, &bp))) return st;
  double angc0 = 0.0, angc1 = 0.0;
  if ((st = tdbp_steering_fit(headers, header_stride, lines, &angc0, &angc1))) return st;
  const double sol = kSol;
  const double slant_range_time = 2 * bp.range0 / sol;             // :145
  const int aperture = tdbp_aperture(satvel_mid, bp.range0, bp.wvl, bp.prf, angc1);
  const double tstart = bp.starttime;                                // :215-219
  const double dtaz = 1.0 / bp.prf;
  const double tend = tstart + (lines - 1) * dtaz;
  const double tmid = 0.5 * (tstart + tend);
  const double rngstart = slant_range_time * sol / 2.0;              // :222-224
  const double dmrg = sol / 2.0 / bp.samplefreq;
  const double rngend = rngstart + (rangesamples - 1) * dmrg;
  if ((st = tdbp_bounds(tstart - 2.0, tend + 2.0, rngstart, rngend, timeorbit, xx, vv, nsv,
                        latlons)))
    return st;
  const double m = std::fmax(std::fmax(std::fabs(latlons[0]), std::fabs(latlons[1])),
                             std::fmax(std::fabs(latlons[2]), std::fabs(latlons[3])));
  *in_scene = m <= 1000.0 ? 1 : 0;                                   // :228-232
  std::memset(g, 0, sizeof *g);
  g->rawdatalines = lines;
  g->samples = rangesamples;
  g->aperture = aperture;
  g->numstatevec = nsv;
  g->rngstart = rngstart;
  g->rngend = rngend;
  g->dmrg = dmrg;
  g->wvl = bp.wvl;
  g->prf = bp.prf;
  g->angc0 = angc0;
  g->angc1 = angc1;
  g->tmid = tmid;
  orbit_at(timeorbit, xx, vv, nsv, tmid, g->xyz_mid, g->vel_mid);   // :234-235
  g->timeorbit = timeorbit;
  g->xx = xx;
  g->vv = vv;
  g->demwidth = demwidth;
  int firstline = 0, lastline = 0, firstpix = 0, lastpix = 0;
  tdbp_window(latlons, firstlat, firstlon, deltalat, deltalon, demwidth, demlength, &firstline,
              &lastline, &firstpix, &lastpix);
  g->firstline = firstline;
  g->nlines = lastline - firstline + 1;
  g->firstpix = firstpix;
  g->lastpix = lastpix;
  g->firstlat = firstlat;
  g->firstlon = firstlon;
  g->deltalat = deltalat;
  g->deltalon = deltalon;
  g->fit_last_line = 0;
  return TDBP_OK;
}

int tdbp_pulse_orbit(const void* d_lines, int nlines, int stride, const double* timeorbit,
                     const double* xx, const double* vv, int nsv, int quantize,
                     double* d_satloc, double* d_satvel, double* d_times, void* stream) {
  if (!d_lines || nlines < 1 || stride < 10 || !timeorbit || !xx || !vv || nsv < 4 ||
      nsv > kMaxStateVec || !d_satloc || !d_satvel || !d_times)
    return set_error(TDBP_ERR_ARG, "pulse orbit: bad arguments");
  int st = current_device_ready();
  if (st) return st;
  Svs h;
  std::memset(&h, 0, sizeof h);
  for (int i = 0; i < nsv; ++i) {
    h.t[i] = timeorbit[i];
    for (int c = 0; c < 3; ++c) {
      h.x[i][c] = xx[3 * i + c];
      h.v[i][c] = vv[3 * i + c];
    }
  }
  hipStream_t s = (hipStream_t)stream;
  Svs* d_sv = nullptr;
  int* d_bad = nullptr;
  if ((st = hip_check(hipMallocAsync((void**)&d_sv, sizeof(Svs), s), "pulse orbit scratch")) ||
      (st = hip_check(hipMallocAsync((void**)&d_bad, sizeof(int), s), "pulse orbit scratch")))
    return st;
  int bad = 0;
  st = hip_check(hipMemcpyAsync(d_sv, &h, sizeof h, hipMemcpyHostToDevice, s), "sv upload");
  if (!st) st = hip_check(hipMemcpyAsync(d_bad, &bad, sizeof bad, hipMemcpyHostToDevice, s),
                          "counter upload");
  if (!st) {
    hipLaunchKernelGGL(k_pulse_orbit, dim3((nlines + 255) / 256), dim3(256), 0, s,
                       (const uint32_t*)d_lines, nlines, 2 * stride, d_sv, nsv,
                       quantize ? 1 : 0, d_satloc, d_satvel, d_times, d_bad);
    st = hip_check(hipGetLastError(), "pulse orbit launch");
  }
  if (!st) st = hip_check(hipMemcpyAsync(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost, s),
                          "counter download");
  (void)hipFreeAsync(d_sv, s);
  (void)hipFreeAsync(d_bad, s);
  const int st2 = hip_check(hipStreamSynchronize(s), "pulse orbit sync");
  if (st) return st;
  if (st2) return st2;
  if (bad)
    return set_error(TDBP_ERR_ARG, "pulse orbit: %d line times outside the state vectors", bad);
  return TDBP_OK;
}

}  // extern "C"
