// howmanygpus -- prints the number of visible GPUs on one line, the contract
// of sentinel/howmanygpus.cu that sentinel/sentinel_gpu.py:98-100 reads to
// split its scene list.  A failed query (no device, no driver) prints 0.
#include <hip/hip_runtime.h>

#include <cstdio>

int main() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  std::printf("%d\n", n);
  return 0;
}
