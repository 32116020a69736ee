// On-chip field kernels for any grid width (runtime LDS pitch).
#include "grn_field.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_FIELD_ENTRY(0, 4, 1024, 2, 8, false),
    GRN_FIELD_ENTRY(0, 1, 1024, 2, 16, false),
};

extern "C" const GrnKernelEntry *grn_field_table_wany(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
