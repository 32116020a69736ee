// grn_probe.hip -- a plain device-to-device copy of known size.  Profilers
// read the HBM byte counters (FETCH_SIZE, WRITE_SIZE) of this kernel to
// calibrate the same counters on the simulation kernels.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "grn.h"

__global__ __launch_bounds__(256) void grn_copy_probe_kernel(const float4 *__restrict__ in,
                                                             float4 *__restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = in[i];
}

extern "C" int grn_copy_probe(long bytes, double *ms) {
  if (bytes <= 0 || bytes % 16) return GRN_E_INVALID;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  if (hipMalloc(&a, bytes) != hipSuccess) return GRN_E_NOMEM;
  if (hipMalloc(&b, bytes) != hipSuccess) {
    (void)hipFree(a);
    return GRN_E_NOMEM;
  }
  int rc = GRN_OK;
  if (hipMemset(a, 0, bytes) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = GRN_E_DEVICE;
  } else {
    const size_t n = (size_t)bytes / 16;
    (void)hipEventRecord(e0, 0);
    grn_copy_probe_kernel<<<dim3(2048), dim3(256), 0, 0>>>((const float4 *)a, (float4 *)b, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) rc = GRN_E_DEVICE;
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}
