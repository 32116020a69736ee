// Halo-tiled kernels with 512-thread workgroups: twice the registers per
// thread of the 1024-thread ones, so the (hi, lo) state of more cells per
// thread stays in registers and link passes serve larger cell blocks.
#include "grn_tiled.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_TILED_ENTRY(4, 512, 2, 8, false),
    GRN_TILED_ENTRY(4, 512, 9, 10, false),  // 11+ genes would spill at 4 cells per thread
    GRN_TILED_ENTRY(2, 512, 9, 16, false),
};

extern "C" const GrnKernelEntry *grn_field_table_tiled512(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
