// On-chip field kernels for 48-wide grids.
#include "grn_field.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_FIELD_ENTRY(48, 3, 768, 2, 8, false),
    GRN_FIELD_ENTRY(48, 4, 576, 2, 12, false),
};

extern "C" const GrnKernelEntry *grn_field_table_w48(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
