// tdbp_api_internal.h -- host helpers shared by the API, batch and drop-in
// translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "tdbp_internal.h"

namespace tdbp {

int set_error(int code, const char* fmt, ...);
int hip_check(hipError_t e, const char* what);
int ensure_device(int device);
int current_device_ready();
int cu_count();
int validate_geom(const tdbp_geom* g);
void fill_burst(const tdbp_geom* g, BurstDev* b);
void set_tiles(BurstDev* b, int variant);
std::vector<double> pack_orbit(const tdbp_geom* g);
void point_orbit(BurstDev* b, const double* d_orbit, int n);
int check_samples(const tdbp_samples* s, int precision);
int default_variant(int precision);
int fit_on_stream(BurstDev& b, const tdbp_geom* g, double* d_coef,
                  hipStream_t s);
int run_backproject(const tdbp_geom* g, const tdbp_samples* smp,
                    const double* d_satloc, const int16_t* d_dem,
                    const double* d_coef, void* d_out, int precision,
                    int variant, hipStream_t s);

}  // namespace tdbp
