// On-chip field kernels for 32-wide grids with 13-16 genes, 256-thread
// workgroups of 4 cells per thread.
#include "grn_field.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_FIELD_ENTRY(32, 4, 256, 13, 16, false),
};

extern "C" const GrnKernelEntry *grn_field_table_w32wide13(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
