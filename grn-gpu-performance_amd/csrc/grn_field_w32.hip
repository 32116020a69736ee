// On-chip field kernels for 32-wide grids.
#include "grn_field.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_FIELD_ENTRY(32, 2, 512, 2, 8, false),
    // 9-12 genes: half the cells per thread of the 256-thread kernel, so a
    // model's Euler step (its latency, when few models share the device)
    // takes half as long
    GRN_FIELD_ENTRY(32, 2, 512, 9, 12, false),
};

extern "C" const GrnKernelEntry *grn_field_table_w32(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
