// On-chip field kernels for 32-wide grids with up to 16 genes, 256-thread
// workgroups of 4 cells per thread (one kernel per gene range, so that the
// kernel files build in parallel).
#include "grn_field.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_FIELD_ENTRY(32, 4, 256, 2, 8, false),
};

extern "C" const GrnKernelEntry *grn_field_table_w32wide(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
