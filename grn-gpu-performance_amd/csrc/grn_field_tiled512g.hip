// Halo-tiled kernels for 11-12-gene models with more cells per thread than
// the 9-16-gene kernel (3 and 4 against 2): larger windows, fewer recomputed
// halo cells per interior cell.
#include "grn_tiled.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_TILED_ENTRY(3, 512, 11, 12, false, 52),
    GRN_TILED_ENTRY(4, 512, 11, 12, false, 52),
};

extern "C" const GrnKernelEntry *grn_field_table_tiled512g(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
