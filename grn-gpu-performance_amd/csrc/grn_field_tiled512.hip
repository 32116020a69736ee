// Halo-tiled kernels with 512-thread workgroups: twice the registers per
// thread of the 1024-thread ones, so the (hi, lo) state of more cells per
// thread stays in registers and link passes serve larger cell blocks.  Their
// LDS window rows have a fixed pitch of 54 floats (windows up to 52 cells
// wide: 33 at 100x100, 51 at 128x128), so neighbour reads use immediate
// offsets.
#include "grn_tiled.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_TILED_ENTRY(4, 512, 2, 8, false, 52),
    GRN_TILED_ENTRY(4, 512, 9, 10, false, 52),  // 11+ genes would spill at 4 cells per thread
    GRN_TILED_ENTRY(2, 512, 9, 16, false, 52),
};

extern "C" const GrnKernelEntry *grn_field_table_tiled512(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
