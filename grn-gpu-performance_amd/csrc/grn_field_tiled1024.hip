// Halo-tiled kernels with 1024-thread workgroups and few cells per thread:
// four wavefronts per SIMD (twice the 512-thread kernels' occupancy), so
// LDS reads, transcendentals and barriers of one wave hide behind the
// others, with windows of up to 2,048 cells (1,024 for 9-12 genes).  LDS
// window rows have the fixed pitch 54 (windows up to 52 cells wide).
#include "grn_tiled.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_TILED_ENTRY(2, 1024, 2, 5, false, 52),
    GRN_TILED_ENTRY(2, 1024, 6, 8, false, 52),
    GRN_TILED_ENTRY(2, 1024, 9, 10, false, 52),
    GRN_TILED_ENTRY(1, 1024, 9, 12, false, 52),
};

extern "C" const GrnKernelEntry *grn_field_table_tiled1024(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
