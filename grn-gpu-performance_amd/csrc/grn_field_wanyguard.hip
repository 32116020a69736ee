// On-chip field kernels for any grid width with the diffusion guard: they
// serve fields that can turn negative.
#include "grn_field.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_FIELD_ENTRY(0, 4, 1024, 2, 8, true),
    GRN_FIELD_ENTRY(0, 1, 1024, 2, 16, true),
};

extern "C" const GrnKernelEntry *grn_field_table_wanyguard(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
