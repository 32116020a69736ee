// grn_model.hpp -- host-side GRN model: genes (products), links, the in-use
// closure, the reference's text format, and the compiler that turns a model
// into the kernels' GrnProgram.
#pragma once

#include <string>
#include <vector>

#include "grn.h"
#include "grn_program.h"

namespace grn {

// ModelProd (Src/Model/modelprod.h:10-48).
struct Gene {
  int label = -1;
  double init = 0, lim = 0, deg = 0, dif = 0;
};

// ModelLink (Src/Model/modellink.h:10-46).
struct Link {
  int regulator = 0, regulated = 0;
  double dis_const = 0, hill_coef = 0;
  bool is_and = false;
};

// Model (Src/Model/model.h:14-91): products and links in insertion order.
struct Model {
  std::vector<Gene> genes;
  std::vector<Link> links;

  // Model::calcProductLabelsInUse (model.cpp:243-270), ascending.
  std::vector<int> labels_in_use(int n_morphogens) const;
  // Model::linksToLabel (model.cpp:391-399): indices, model order.
  std::vector<int> links_to_label(int label) const;
  // Model::prodWithLabel (model.cpp:227-240): first product, or nullptr.
  const Gene *gene_with_label(int label) const;
  // Model::calcComplexityInUse (model.cpp:590-611).
  int complexity_in_use(int n_morphogens) const;

  // Text format of Model::operator<< / operator>> (model.cpp:627-706,
  // modelprod.cpp:74-106, modellink.cpp:56-90):
  //   (label init lim deg dif|...*regulator regulated K n isAnd|...)
  // Numbers are written with `precision` significant digits (the reference
  // stream's default is 6).
  std::string to_string(int precision = 6) const;
  // Returns false and sets *err on malformed text.
  static bool parse(const std::string &text, Model *out, std::string *err);
};

// The C ABI's grn_model -> Model; false when the struct is malformed
// (negative counts, null arrays with nonzero counts).
bool model_from_abi(const grn_model *gm, Model *out);

// Work counts of a compiled model, for ordering and the roofline.
struct ProgramInfo {
  int G = 0;       // genes in use
  int L = 0;       // Hill links evaluated (links into genes with an activator)
  int G_h = 0;     // genes whose h comes from Hill links
  int D = 0;       // genes with dif > 0
  double flops_per_cell_step = 0;  // F = 5L + (L - G_h) + 8G + 7D
};

// CModelSimulator::loadModel (Src/GPUSimulator/cmodelsimulator.cpp:114-227)
// and ModelSimulator::loadModel (Src/Simulator/modelsimulator.cpp:71-203),
// compiled into exponent-form Hill links (grn_program.h).  Returns 0 or a
// grn_status code with *err set.
// out64 (optional) receives the same program's constants in double
// precision for the FP64 kernel, with dt64 as its time step.
int compile_model(const Model &m, int n_morphogens, int dataset, float dt, GrnProgram *out,
                  ProgramInfo *info, std::string *err, GrnProgram64 *out64 = nullptr, double dt64 = 0.0);

// Structural check of a compiled program before it is uploaded (the host
// side of the reference's debug-build GPU_ASSERT bounds checks,
// Src/GPUSimulator/cerrorhandle.h:24-28): gene count within the kernels'
// capacity, every gene's link range inside the stored links, every
// regulator index a gene in use, masks consistent with the genes, the
// dataset index below n_datasets and a finite time step.  Returns 0 or
// GRN_E_MODEL with *err naming the first violation.
int validate_program(const GrnProgram &p, int n_datasets, std::string *err);

}  // namespace grn
