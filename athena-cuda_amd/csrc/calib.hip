// calib.hip -- a plain streaming copy of known size. Its HBM counters
// (FETCH_SIZE / WRITE_SIZE under rocprofv3 --pmc) calibrate the per-kernel
// traffic figures of the step kernels. Measurement hook only.
#include <hip/hip_runtime.h>

#include "athmi.h"

namespace {
__global__ void __launch_bounds__(256) k_copy_calibration(const double2* __restrict__ src,
                                                          double2* __restrict__ dst, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    dst[k] = src[k];
}
}  // namespace

extern "C" int athmi_copy_calibration(size_t nbytes, int reps) {
  if (nbytes < 16 || reps < 1) return ATHMI_EINVAL;
  const size_t n = nbytes / 16;
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess) return ATHMI_ENOMEM;
  if (hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipFree(a);
    return ATHMI_ENOMEM;
  }
  int rc = ATHMI_OK;
  if (hipMemset(a, 0, n * 16) != hipSuccess) rc = ATHMI_EHIP;
  for (int r = 0; r < reps && rc == ATHMI_OK; r++) {
    hipLaunchKernelGGL(k_copy_calibration, dim3(4096), dim3(256), 0, 0, a, b, n);
    if (hipGetLastError() != hipSuccess) rc = ATHMI_EHIP;
  }
  if (hipDeviceSynchronize() != hipSuccess) rc = ATHMI_EHIP;
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
