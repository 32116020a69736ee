// Halo-tiled kernels with larger windows for models with up to 8 genes: more
// cells per thread than the 2-8-gene kernel's 4 (8 for up to 4 genes, 6 for
// up to 5, 5 for 6-8), so fewer halo cells are recomputed per interior cell.
// Their LDS window rows have fixed pitches of 60 and 74 floats (windows up to
// 58 and 72 cells wide).  What little these spill is spilled around the tile
// loads and stores, not inside the Euler steps.
#include "grn_tiled.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_TILED_ENTRY(6, 512, 2, 5, false, 58),
    GRN_TILED_ENTRY(8, 512, 2, 4, false, 72),
    GRN_TILED_ENTRY(5, 512, 6, 8, false, 58),
};

extern "C" const GrnKernelEntry *grn_field_table_tiled512s(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
