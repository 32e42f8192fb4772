// tdbp_fft.hip -- cached rocFFT plans for the batched 1-D complex FP32
// transforms of range compression and interferogram formation.  One plan per
// (device, stream, length, batch, direction), created on first use with its
// own execution info and work buffer -- a plan's work buffer must not be
// shared by transforms running concurrently on two streams; every use holds
// the module mutex while it enqueues (rocfft_execute is asynchronous).
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <mutex>
#include <vector>

#include "tdbp_api_internal.h"
#include "tdbp_fft.h"

namespace tdbp {

struct FftPlan {
  int dev;
  hipStream_t stream;
  size_t len;
  int batch;
  bool inverse;
  rocfft_plan plan = nullptr;
  rocfft_execution_info info = nullptr;
  void* work = nullptr;
};

namespace {

std::mutex g_fft_mu;
std::vector<FftPlan*> g_plans;
bool g_setup = false;

int fft_check(rocfft_status s, const char* what) {
  if (s == rocfft_status_success) return TDBP_OK;
  return set_error(TDBP_ERR_HIP, "%s: rocFFT status %d", what, (int)s);
}

int make_plan(FftPlan* p) {
  int st;
  if (!g_setup) {
    if ((st = fft_check(rocfft_setup(), "rocfft_setup"))) return st;
    g_setup = true;
  }
  const size_t len = p->len;
  if ((st = fft_check(rocfft_plan_create(&p->plan, rocfft_placement_inplace,
                                         p->inverse ? rocfft_transform_type_complex_inverse
                                                    : rocfft_transform_type_complex_forward,
                                         rocfft_precision_single, 1, &len, (size_t)p->batch,
                                         nullptr),
                      "plan create")) ||
      (st = fft_check(rocfft_execution_info_create(&p->info), "execution info")))
    return st;
  size_t w = 0;
  if ((st = fft_check(rocfft_plan_get_work_buffer_size(p->plan, &w), "work size"))) return st;
  if (w) {
    if (hipMalloc(&p->work, w) != hipSuccess)
      return set_error(TDBP_ERR_NOMEM, "rocFFT work buffer (%zu B)", w);
    if ((st = fft_check(rocfft_execution_info_set_work_buffer(p->info, p->work, w),
                        "work buffer")))
      return st;
  }
  return TDBP_OK;
}

void drop_plan(FftPlan* p) {
  if (p->info) rocfft_execution_info_destroy(p->info);
  if (p->plan) rocfft_plan_destroy(p->plan);
  if (p->work) (void)hipFree(p->work);
  delete p;
}

}  // namespace

int fft_run(size_t len, int batch, bool inverse, void* data, hipStream_t s) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_fft_mu);
  FftPlan* p = nullptr;
  for (FftPlan* q : g_plans)
    if (q->dev == dev && q->stream == s && q->len == len && q->batch == batch &&
        q->inverse == inverse)
      p = q;
  if (!p) {
    p = new FftPlan{dev, s, len, batch, inverse};
    const int st = make_plan(p);
    if (st) {
      drop_plan(p);
      return st;
    }
    g_plans.push_back(p);
  }
  int st = fft_check(rocfft_execution_info_set_stream(p->info, s), "set stream");
  if (st) return st;
  void* in[1] = {data};
  return fft_check(rocfft_execute(p->plan, in, nullptr, p->info), "execute");
}

}  // namespace tdbp
