// grn_kernels.h -- the kernel instantiations the launcher can pick from.
#ifndef GRN_KERNELS_H
#define GRN_KERNELS_H

#include <hip/hip_runtime.h>

#include "grn_program.h"

struct GrnKernelEntry {
  int wc;     // compile-time grid width, 0 = any width, -1 = halo-tiled kernel, -2 = FP64 kernel
  int cpt;    // cells per thread (vertical strip length)
  int ntmax;  // workgroup size bound
  int glo;    // smallest gene count handled
  int ghi;    // largest gene count handled
  bool guard; // applies the diffusion guard (needed for fields that can go negative)
  const void *fn;
  int pitch;  // halo-tiled: LDS window rows have the compile-time pitch pitch + 2, so the
              // window is at most pitch cells wide; 0 = the window width + 2 (runtime)
};

// Instantiations of the on-chip field kernel, one table per translation unit.
extern "C" const GrnKernelEntry *grn_field_table_w32(int *n);
extern "C" const GrnKernelEntry *grn_field_table_w32wide(int *n);
extern "C" const GrnKernelEntry *grn_field_table_w32wide9(int *n);
extern "C" const GrnKernelEntry *grn_field_table_w32wide13(int *n);
extern "C" const GrnKernelEntry *grn_field_table_w8(int *n);
extern "C" const GrnKernelEntry *grn_field_table_w48(int *n);
extern "C" const GrnKernelEntry *grn_field_table_w64(int *n);
extern "C" const GrnKernelEntry *grn_field_table_wany(int *n);
extern "C" const GrnKernelEntry *grn_field_table_wanyguard(int *n);
extern "C" const GrnKernelEntry *grn_field_table_tiled(int *n);
extern "C" const GrnKernelEntry *grn_field_table_tiled512(int *n);
extern "C" const GrnKernelEntry *grn_field_table_tiled512s(int *n);
extern "C" const GrnKernelEntry *grn_field_table_tiled512g(int *n);
extern "C" const GrnKernelEntry *grn_field_table_tiled1024(int *n);
// The FP64 kernel (grn_exact.hip): wc = -2.
extern "C" const GrnKernelEntry *grn_field_table_exact(int *n);

// Zeroes out[0 .. batch->n_total) on stream (grn_probe.hip).
hipError_t grn_launch_clear_results(GrnBatch *batch, GrnModelOut *out, hipStream_t stream);

#endif
