// tdbp_fft.h -- batched in-place 1-D complex FP32 FFTs on the current device
// (rocFFT plans cached per device, length, batch and direction).  Forward is
// e^{-2 pi i jk/n}, inverse e^{+2 pi i jk/n} without normalisation, as
// FFTW's FFTW_FORWARD / FFTW_BACKWARD.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace tdbp {

// data: batch x len complex float, contiguous rows.  Asynchronous on s.
int fft_run(size_t len, int batch, bool inverse, void* data, hipStream_t s);

}  // namespace tdbp
