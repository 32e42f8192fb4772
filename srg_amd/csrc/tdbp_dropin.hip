// tdbp_dropin.hip -- azimuth_compress_, the symbol the reference Fortran
// burst driver calls (sentinel/backprojectgpusub.f90:244-247), with the
// semantics of sentinel/azimuth_compress.cu:181-562:
//   * DEM lines and pixels of the burst from latlons (:295-301, :378-379),
//     the "Burst line limits" line on stdout (:302);
//   * the DEM slice read from fddem at firstline*W*2 (:400-403);
//   * aperture fit, azoff, pulse sum with 1/(azend-azstart) scaling;
//   * the .geo rows [firstline, firstline+nlines) read back from fdout,
//     pixels replaced where |Re(out)| > 1e-18, written back (:513-537).
// The DEM and .geo rows pass through pinned staging buffers; the .geo rows
// are read while the device integrates and merged on the device.
// Differences, all deliberate: pread/pwrite instead of a never-flushed FILE*
// (the bytes land at once and the fd offsets are untouched); a short DEM
// read is an error (the .geo is left unchanged) instead of integrating over
// uninitialised heights; .geo bytes past the end of the file read as zero
// instead of uninitialised memory; device buffers persist across calls.
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <mutex>
#include <vector>

#include "tdbp_api_internal.h"
#include "tdbp_internal.h"

using namespace tdbp;

namespace {

// Grow-only device buffers and pinned host staging reused by every call in
// the process.
enum { kBurst, kPacked, kScale, kSat, kDem, kCoef, kOut, kGeo, kNumDev };
enum { kHostDem, kHostGeo, kNumHost };
struct Pool {
  void* p[kNumDev] = {};
  size_t cap[kNumDev] = {};
  void* h[kNumHost] = {};
  size_t hcap[kNumHost] = {};
  hipStream_t s = nullptr;
  hipEvent_t ev[6] = {};
};

std::mutex g_mu;
Pool g_pool;
thread_local int g_status = TDBP_OK;
// Phase times of the last call, ms: total, DEM read, burst+satloc+DEM upload,
// device compute (pack, fit, sum), .geo read, .geo upload+merge+download,
// .geo write.
double g_timing[7] = {};

int need(int slot, size_t bytes) {
  if (g_pool.cap[slot] >= bytes && g_pool.p[slot]) return TDBP_OK;
  if (g_pool.p[slot]) {
    (void)hipStreamSynchronize(g_pool.s);
    (void)hipFree(g_pool.p[slot]);
    g_pool.p[slot] = nullptr;
    g_pool.cap[slot] = 0;
  }
  if (hipMalloc(&g_pool.p[slot], bytes) != hipSuccess) {
    g_pool.p[slot] = nullptr;
    return set_error(TDBP_ERR_NOMEM, "drop-in hipMalloc(%zu) failed", bytes);
  }
  g_pool.cap[slot] = bytes;
  return TDBP_OK;
}

int need_host(int slot, size_t bytes) {
  if (g_pool.hcap[slot] >= bytes && g_pool.h[slot]) return TDBP_OK;
  if (g_pool.h[slot]) {
    (void)hipStreamSynchronize(g_pool.s);
    (void)hipHostFree(g_pool.h[slot]);
    g_pool.h[slot] = nullptr;
    g_pool.hcap[slot] = 0;
  }
  if (hipHostMalloc(&g_pool.h[slot], bytes, hipHostMallocDefault) != hipSuccess) {
    g_pool.h[slot] = nullptr;
    return set_error(TDBP_ERR_NOMEM, "drop-in hipHostMalloc(%zu) failed", bytes);
  }
  g_pool.hcap[slot] = bytes;
  return TDBP_OK;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// TDBP_PRECISION: fp64 (default) | mixed_fp16 | mixed_bf16 | mixed_cf32.
void precision_from_env(int* prec, int* dtype) {
  *prec = TDBP_PLAIN_FP64;
  *dtype = TDBP_CF32;
  const char* e = std::getenv("TDBP_PRECISION");
  if (!e) return;
  if (!std::strcmp(e, "mixed_fp16")) {
    *prec = TDBP_MIXED;
    *dtype = TDBP_CF16;
  } else if (!std::strcmp(e, "mixed_bf16")) {
    *prec = TDBP_MIXED;
    *dtype = TDBP_CBF16;
  } else if (!std::strcmp(e, "mixed_cf32")) {
    *prec = TDBP_MIXED;
    *dtype = TDBP_CF32;
  }
}

ssize_t read_full(int fd, void* buf, size_t n, off_t off) {
  size_t got = 0;
  while (got < n) {
    ssize_t r = pread(fd, (char*)buf + got, n - got, off + (off_t)got);
    if (r < 0) return -1;
    if (r == 0) break;
    got += (size_t)r;
  }
  return (ssize_t)got;
}

int run(const float* burst, const double* satloc, int L, int S, int W,
        int demlength, int fdout, int fddem, double deltalat, double deltalon,
        double firstlat, double firstlon, const double* latlons,
        const double* timeorbit, const double* xx, const double* vv, int nsv,
        double rngstart, double rngend, double tmid, const double* xyz_mid,
        const double* vel_mid, double dmrg, double wvl, int aperture,
        double angc0, double angc1, double prf) {
  int st = ensure_device(0);  // the first visible device, as cudaSetDevice(0)
  if (st) return st;
  int firstline, lastline, firstpix, lastpix;
  tdbp_window(latlons, firstlat, firstlon, deltalat, deltalon, W, demlength,
              &firstline, &lastline, &firstpix, &lastpix);
  const int nlines = lastline - firstline + 1;
  std::printf("Burst line limits, size (first,last,nlines) %d %d %d\n",
              firstline, lastline, nlines);
  std::fflush(stdout);

  tdbp_geom g;
  std::memset(&g, 0, sizeof g);
  g.rawdatalines = L;
  g.samples = S;
  g.aperture = aperture;
  g.numstatevec = nsv;
  g.rngstart = rngstart;
  g.rngend = rngend;
  g.dmrg = dmrg;
  g.wvl = wvl;
  g.prf = prf;
  g.angc0 = angc0;
  g.angc1 = angc1;
  g.tmid = tmid;
  for (int i = 0; i < 3; ++i) {
    g.xyz_mid[i] = xyz_mid[i];
    g.vel_mid[i] = vel_mid[i];
  }
  g.timeorbit = timeorbit;
  g.xx = xx;
  g.vv = vv;
  g.demwidth = W;
  g.firstline = firstline;
  g.nlines = nlines;
  g.firstpix = firstpix;
  g.lastpix = lastpix;
  g.firstlat = firstlat;
  g.firstlon = firstlon;
  g.deltalat = deltalat;
  g.deltalon = deltalon;
  const char* fl = std::getenv("TDBP_FIT_LAST_LINE");
  g.fit_last_line = (fl && fl[0] == '1') ? 1 : 0;
  if ((st = validate_geom(&g))) return st;

  int prec, dtype;
  precision_from_env(&prec, &dtype);

  std::lock_guard<std::mutex> lk(g_mu);
  const double t0 = now_ms();
  if (!g_pool.s) {
    if ((st = hip_check(hipStreamCreateWithFlags(&g_pool.s, hipStreamNonBlocking),
                        "drop-in stream")))
      return st;
    for (auto& e : g_pool.ev)
      if ((st = hip_check(hipEventCreate(&e), "drop-in event"))) return st;
  }
  hipStream_t s = g_pool.s;
  hipEvent_t* ev = g_pool.ev;
  const size_t npx = (size_t)nlines * W;
  const size_t nsamp = (size_t)L * S;
  if ((st = need(kBurst, nsamp * 8)) || (st = need(kSat, (size_t)L * 24)) ||
      (st = need(kDem, npx * 2)) || (st = need(kCoef, (size_t)nlines * 24)) ||
      (st = need(kOut, npx * 8)) || (st = need(kGeo, npx * 8)) ||
      (st = need_host(kHostDem, npx * 2)) || (st = need_host(kHostGeo, npx * 8)))
    return st;
  if (dtype != TDBP_CF32 &&
      ((st = need(kPacked, nsamp * 4)) || (st = need(kScale, (size_t)L * 4))))
    return st;
  int16_t* hdem = (int16_t*)g_pool.h[kHostDem];
  float2* hgeo = (float2*)g_pool.h[kHostGeo];

  // DEM rows into pinned memory, then everything the device needs.
  const ssize_t got = read_full(fddem, hdem, npx * 2, (off_t)firstline * W * 2);
  if (got != (ssize_t)(npx * 2))
    return set_error(TDBP_ERR_IO, "DEM read of %zu bytes at line %d returned %zd",
                     npx * 2, firstline, got);
  const double t1 = now_ms();
  (void)hipEventRecord(ev[0], s);
  if ((st = hip_check(hipMemcpyAsync(g_pool.p[kDem], hdem, npx * 2,
                                     hipMemcpyHostToDevice, s),
                      "dem upload")) ||
      (st = hip_check(hipMemcpyAsync(g_pool.p[kSat], satloc, (size_t)L * 24,
                                     hipMemcpyHostToDevice, s),
                      "satloc upload")) ||
      (st = hip_check(hipMemcpyAsync(g_pool.p[kBurst], burst, nsamp * 8,
                                     hipMemcpyHostToDevice, s),
                      "burst upload")))
    return st;
  (void)hipEventRecord(ev[1], s);
  tdbp_samples smp;
  smp.d_data = g_pool.p[kBurst];
  smp.d_scale = nullptr;
  smp.dtype = TDBP_CF32;
  if (dtype != TDBP_CF32) {
    if ((st = tdbp_pack_samples(g_pool.p[kBurst], L, S, dtype, g_pool.p[kPacked],
                                (float*)g_pool.p[kScale], s)))
      return st;
    smp.d_data = g_pool.p[kPacked];
    smp.d_scale = (const float*)g_pool.p[kScale];
    smp.dtype = dtype;
  }
  if ((st = tdbp_fit_aperture(&g, (const int16_t*)g_pool.p[kDem],
                              (double*)g_pool.p[kCoef], s)))
    return st;
  if ((st = run_backproject(&g, &smp, (const double*)g_pool.p[kSat],
                            (const int16_t*)g_pool.p[kDem],
                            (const double*)g_pool.p[kCoef], g_pool.p[kOut], prec,
                            -1, s)))
    return st;
  (void)hipEventRecord(ev[2], s);

  // The .geo rows are read while the device integrates; they are merged on
  // the device and come back in the same pinned buffer.
  const off_t off = (off_t)firstline * W * 8;
  const double t2 = now_ms();
  const ssize_t gg = read_full(fdout, hgeo, npx * 8, off);
  if (gg < 0) {
    (void)hipStreamSynchronize(s);
    return set_error(TDBP_ERR_IO, ".geo read at line %d failed", firstline);
  }
  if ((size_t)gg < npx * 8)  // rows past the end of the file read as zero
    std::memset((char*)hgeo + gg, 0, npx * 8 - (size_t)gg);
  const double t3 = now_ms();
  (void)hipEventRecord(ev[3], s);
  if ((st = hip_check(hipMemcpyAsync(g_pool.p[kGeo], hgeo, npx * 8,
                                     hipMemcpyHostToDevice, s),
                      ".geo upload")) ||
      (st = hip_check(launch_merge((float2*)g_pool.p[kGeo], (const float2*)g_pool.p[kOut],
                                   npx, 0, s),
                      "merge launch")) ||
      (st = hip_check(hipMemcpyAsync(hgeo, g_pool.p[kGeo], npx * 8,
                                     hipMemcpyDeviceToHost, s),
                      ".geo download")))
    return st;
  (void)hipEventRecord(ev[4], s);
  if ((st = hip_check(hipStreamSynchronize(s), "drop-in sync"))) return st;
  const double t4 = now_ms();
  size_t put = 0;
  while (put < npx * 8) {
    const ssize_t w = pwrite(fdout, (const char*)hgeo + put, npx * 8 - put,
                             off + (off_t)put);
    if (w <= 0) return set_error(TDBP_ERR_IO, ".geo write failed");
    put += (size_t)w;
  }
  const double t5 = now_ms();
  float up = 0.f, comp = 0.f, rmw = 0.f;
  (void)hipEventElapsedTime(&up, ev[0], ev[1]);
  (void)hipEventElapsedTime(&comp, ev[1], ev[2]);
  (void)hipEventElapsedTime(&rmw, ev[3], ev[4]);
  g_timing[0] = t5 - t0;
  g_timing[1] = t1 - t0;
  g_timing[2] = up;
  g_timing[3] = comp;
  g_timing[4] = t3 - t2;
  g_timing[5] = rmw;
  g_timing[6] = t5 - t4;
  return TDBP_OK;
}

}  // namespace

extern "C" {

void azimuth_compress_(void* burstdata, double* satloc, int* rawdatalines,
                       int* samplesPerBurst, int* demwidth, int* demlength,
                       int* fdout, int* fddem, double* deltalat,
                       double* deltalon, double* firstlat, double* firstlon,
                       double* latlons, double* timeorbit, double* xx,
                       double* vv, int* numstatevec, double* rngstart,
                       double* rngend, double* tstart, double* tend,
                       double* tmid, double* xyz_mid, double* vel_mid,
                       double* t, double* dtaz, double* dmrg, double* wvl,
                       int* aperture, int* iaperture, double* angc0,
                       double* angc1, double* prf) {
  (void)tstart;
  (void)tend;
  (void)t;
  (void)dtaz;
  (void)iaperture;
  g_status = run((const float*)burstdata, satloc, *rawdatalines, *samplesPerBurst,
                 *demwidth, *demlength, *fdout, *fddem, *deltalat, *deltalon,
                 *firstlat, *firstlon, latlons, timeorbit, xx, vv, *numstatevec,
                 *rngstart, *rngend, *tmid, xyz_mid, vel_mid, *dmrg, *wvl,
                 *aperture, *angc0, *angc1, *prf);
  if (g_status != TDBP_OK)
    std::fprintf(stderr, "azimuth_compress: burst skipped: %s\n",
                 tdbp_last_error());
}

int tdbp_dropin_status(void) { return g_status; }

int tdbp_dropin_timing(double* ms, int n) {
  std::lock_guard<std::mutex> lk(g_mu);
  const int k = n < 7 ? n : 7;
  for (int i = 0; i < k; ++i) ms[i] = g_timing[i];
  return 7;
}

}  // extern "C"
