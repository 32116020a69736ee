// On-chip field kernels for 8- and 16-wide grids (an 8x8 grid is one
// wavefront, one cell per lane; 16x16 four): the coarse grids of the
// multi-fidelity screen and of the coarse-to-fine search.  Up to 10 genes,
// so the gene bodies fit the 1024-thread register budget (125 VGPRs) without
// the spills of the any-width 2-16-gene kernel (400 bytes per lane; 48 with
// 12 genes); larger models take that kernel.
#include "grn_field.hpp"

static const GrnKernelEntry kTable[] = {
    GRN_FIELD_ENTRY(8, 1, 1024, 2, 10, false),
    GRN_FIELD_ENTRY(16, 1, 1024, 2, 10, false),
};

extern "C" const GrnKernelEntry *grn_field_table_w8(int *n) {
  *n = (int)(sizeof(kTable) / sizeof(kTable[0]));
  return kTable;
}
