// On-chip field kernels for 32-wide grids.
#include "grn_field.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_FIELD_ENTRY(32, 2, 512, 2, 8, false),
};

extern "C" const GrnKernelEntry *grn_field_table_w32(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
