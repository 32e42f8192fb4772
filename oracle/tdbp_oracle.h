/*
 * tdbp_oracle.h -- CPU restatement of the reference TDBP path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (srg_amd/libsrg_tdbp.so) never links or calls
 * it, and has no CPU fallback.
 *
 * Every function restates one piece of the reference algorithm; the file:line
 * it follows is given at the definition in tdbp_oracle.c.  All arithmetic is
 * FP64 in the same evaluation order as the reference C/Fortran, so that with
 * gcc -O2 (no FMA contraction on x86-64) results are bit-identical to the
 * compiled reference CPU path (oracle/_ref).
 *
 * Layouts (row-major, C order):
 *   burst   : complex float [L][S]     (Fortran rangeprocdata(S, L) slice)
 *   satloc  : double [L][3]
 *   dem     : int16 [nlines][demwidth]
 *   xyz     : double [nlines][demwidth][3]
 *   xyzfit  : double [nlines][9]       (3 fit nodes x ECEF)
 *   coef    : double [nlines][3]       (parabola c0, c1, c2 per DEM line)
 *   azoff   : double [nlines][demwidth]
 *   out     : complex float [nlines][demwidth]
 *   orbit   : timeorbit[n], xx[n][3], vv[n][3]   (Fortran xx(3,n))
 */
#ifndef TDBP_ORACLE_H
#define TDBP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* pixelint flavours */
#define ORC_CPU_SEMANTICS 0  /* azimuth_compress_cpu.c: FP64 lerp, no 1/N   */
#define ORC_GPU_SEMANTICS 1  /* azimuth_compress.cu:   FP32 lerp, x 1/N     */

void orc_orbithermite(const double *x4 /*[4][3]*/, const double *v4 /*[4][3]*/,
                      const double *t4, double time, double *pos, double *vel);

int orc_intp_orbit(const double *timeorbit, const double *xx, const double *vv,
                   int numstatevec, double time, double *pos, double *vel);

void orc_orbitrangetime(const double *xyz, const double *timeorbit,
                        const double *xx, const double *vv, int numstatevec,
                        double tline0, const double *satx0,
                        const double *satv0, double *tline, double *range);

void orc_setxyz(const int16_t *dem, double *xyz, double *xyzfit, int demwidth,
                double firstlat, double deltalat, double firstlon,
                double deltalon, int firstline, int nlines, int firstpix,
                int lastpix);

/* Aperture-centre parabola for DEM lines [0, nfit) of the window.  The
 * reference fits nlines-1 lines and leaves the last one undefined; callers
 * here pass nfit = nlines to get a defined last line. */
void orc_fitcoef(const double *xyzfit, int nfit, const double *timeorbit,
                 const double *xx, const double *vv, int numstatevec,
                 double tmid, const double *xyz_mid, const double *vel_mid,
                 double wvl, double angc0, double angc1, double prf,
                 int rawdatalines, double *coef);

void orc_setazoff(const double *coef, double *azoff, int demwidth, int nlines,
                  int firstpix, int lastpix, int aperture, int rawdatalines);

/* The coherent sum.  `acc` (optional, complex double [nlines*demwidth]) gets
 * the FP64 accumulator before the cf32 rounding; `npairs` (optional) gets the
 * number of pixel-pulse pairs visited (aze-azs summed over azoff>0 pixels).
 * `nthreads` <= 0 uses the OpenMP default. */
void orc_pixelint(const float *burst, const double *satloc, const double *xyz,
                  const double *azoff, int demwidth, int aperture,
                  int rawdatalines, int samplesPerBurst, double rngstart,
                  double rngend, double dmrg, double wvl, float *out,
                  int nlines, int semantics, double *acc, int64_t *npairs,
                  int nthreads);

/* Window arithmetic of the entry point (azimuth_compress_cpu.c:246-284). */
void orc_window(const double *latlons, double firstlat, double firstlon,
                double deltalat, double deltalon, int demwidth, int demlength,
                int *firstline, int *lastline, int *firstpix, int *lastpix);

/* Whole burst in memory: dem slice -> xyz -> fit -> azoff -> pixelint.
 * `dem` is the [nlines][demwidth] slice starting at firstline.  nfit as in
 * orc_fitcoef.  Returns pairs visited. */
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
                        float *out, double *acc, int nthreads);

/* .geo merge of a burst window (npx cf32 pixels) into the grid rows it
 * covers: CPU twin abs((int)re) > 1e-18, GPU entry fabs(re) > 1e-18. */
void orc_merge_geo(float *grid, const float *out, long npx, int semantics);

/* Full entry point with file I/O (DEM read by fd, .geo read-modify-write by
 * fd).  semantics selects the CPU twin (int-abs merge, no 1/N) or the GPU
 * entry (fabs merge, 1/N).  nfit_last: 0 reproduces the reference loop bound
 * (last line's coef set to 0 instead of left undefined), 1 fits it. */
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
                          int nfit_last);

#ifdef __cplusplus
}
#endif
#endif
