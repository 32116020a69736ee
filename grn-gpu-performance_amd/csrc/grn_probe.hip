// grn_probe.hip -- small utility kernels:
//  * the result-clearing kernel that runs before every batch's simulation
//    launches (a kernel rather than a memset so that it can sit in a captured
//    graph; the batch size comes from the device-side batch table);
//  * a plain device-to-device copy of known size, whose HBM byte counters
//    (FETCH_SIZE, WRITE_SIZE) calibrate the same counters on the simulation
//    kernels.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "grn.h"
#include "grn_kernels.h"
#include "grn_program.h"

__global__ __launch_bounds__(256) void grn_clear_results_kernel(GrnBatch *__restrict__ batch,
                                                                GrnModelOut *__restrict__ out) {
  const int n = batch->n_total;
  // rewind the work queues: the same batch table can be launched again
  if (blockIdx.x == 0 && (int)threadIdx.x < batch->n_queues) batch->q[threadIdx.x].head = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    GrnModelOut z;
    z.err = 0.0;
    z.change = 0.0;
    z.steps = 0;
    z.exps = 0;
    z.done = 0;
    z.pad = 0;
    out[i] = z;
  }
}

hipError_t grn_launch_clear_results(GrnBatch *batch, GrnModelOut *out, hipStream_t stream) {
  hipLaunchKernelGGL(grn_clear_results_kernel, dim3(64), dim3(256), 0, stream, batch, out);
  return hipGetLastError();
}

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
