// On-chip field kernels for 64-wide grids.
#include "grn_field.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_FIELD_ENTRY(64, 4, 1024, 2, 8, false),
    GRN_FIELD_ENTRY(64, 8, 512, 2, 9, false),
};

extern "C" const GrnKernelEntry *grn_field_table_w64(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
