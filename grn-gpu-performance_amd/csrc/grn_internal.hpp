// grn_internal.hpp -- calls between the library's own translation units that
// are not part of the C ABI.
#pragma once

#include <string>

#include "grn.h"

namespace grn {

// A host stage run after a batch's launches, in stream order: it receives the
// converted results of all n models and returns 0 or a grn_status code with
// *err set.  In graph mode it is captured in the batch's graph as a host
// node; it must not call the HIP runtime.
typedef int (*BatchHook)(void *user, const grn_result *results, int n, std::string *err);

// grn_evaluate_batch with a host stage: the results are copied back and
// `hook` runs within the batch's stream work (a host node of the captured
// graph in GRN_LAUNCH_GRAPH mode), before this call returns.
int evaluate_batch_hooked(void *ctx, int n, const int *dataset_ids, const grn_model *models,
                          const double *max_errors, const double *dt_override, int path, grn_result *out,
                          BatchHook hook, void *user);

}  // namespace grn
