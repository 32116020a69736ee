// Halo-tiled kernels for grids whose field does not fit one workgroup's LDS.
#include "grn_tiled.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_TILED_ENTRY(4, 1024, 2, 8, false, 0),
    GRN_TILED_ENTRY(2, 1024, 9, 16, false, 0),
    GRN_TILED_ENTRY(2, 1024, 2, 16, true, 0),
};

extern "C" const GrnKernelEntry *grn_field_table_tiled(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
