/*
 * grn_oracle.h -- FP64 CPU restatement of the reference GRN simulator and
 * fitness calculator.  TEST INFRASTRUCTURE ONLY: the parity tests, the smoke
 * check and the CPU-baseline leg of the benchmark load this library as the
 * checker.  The product library never links or calls it.
 *
 * Reference behaviour restated (paths relative to the reference tree):
 *   in-use closure ........ Src/Model/model.cpp:243-283
 *   op compile ............ Src/Simulator/modelsimulator.cpp:71-203
 *   Hill / AND / OR / zero  Src/Simulator/simophillact.cpp:9-20,
 *                           simophillrep.cpp:9-20, simopand.cpp:20-22,
 *                           simopor.cpp:20-22, simopzero.cpp:15-17
 *   Euler step ............ Src/Simulator/modelsimulator.cpp:205-280
 *   stop rule ............. Src/Simulator/simulator.cpp:54-63
 *   initial state ......... Src/Simulator/simulator.cpp:35-52, simstate.cpp:64-90
 *   log fitness ........... Src/Simulator/simstate.cpp:135-158
 *   evaluate contract ..... Src/Search/evaluatormorphogens.cpp:48-76
 *   ceilS, min, max ....... Src/Common/mathalgo.h:55-62,208-216
 *   experiment images ..... Src/Simulator/simstate.cpp:71-90
 *
 * Field layout used by every function here: plane k, row y, column x at
 * k*H*W + y*W + x (x fastest), which is the element order of the reference's
 * column-major Eigen::MatrixXd(W,H) planes (simstate.h:63-66).
 */
#ifndef GRN_ORACLE_H
#define GRN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int label;
  double init, lim, deg, dif;
} gro_gene;

typedef struct {
  int regulator, regulated;
  double dis_const, hill_coef;
  int is_and;
} gro_link;

enum { GRO_FITNESS_REF = 0, GRO_FITNESS_MSE = 1 };

typedef struct {
  int dim_x, dim_y;
  int n_in, n_tgt;
  int n_steps;
  double dt;
  double min_conc;
  double min_conc_change;
  double dist_error_threshold;
  int error_precision; /* significant digits of ceilS; <= 0 disables rounding */
  int fitness_kind;    /* GRO_FITNESS_REF or GRO_FITNESS_MSE */
} gro_params;

typedef struct {
  double fitness;     /* after the 100 sentinel and ceilS */
  double raw_error;   /* accumulated error before sentinel / rounding */
  double change;      /* accumulated last-step max change over experiments */
  int n_products;     /* genes in use */
  int steps;          /* Euler steps executed, summed over experiments */
  int experiments;    /* experiments simulated before the maxError cut */
} gro_result;

/* MathAlgo::ceilS (mathalgo.h:55-62). */
double gro_ceil_s(double n, int d);

/* Model::calcProductLabelsInUse (model.cpp:243-270), sorted ascending as in
 * ModelSimulator::loadModel (modelsimulator.cpp:73-75).  Returns the count or
 * a negative value when out_cap is too small. */
int gro_labels_in_use(int n_genes, const gro_gene *genes, int n_links,
                      const gro_link *links, int n_morphogens, int *out_labels,
                      int out_cap);

/* EvaluatorMorphogens::evaluate for one model.  init: [n_exp][n_in][H][W],
 * target: [n_exp][n_tgt][H][W].  final_fields (optional): receives the last
 * simulated experiment's field, [n_products][H][W] with n_products <= 64.
 * Returns 0 on success, a negative code for a malformed model. */
int gro_evaluate(const gro_params *p, int n_exp, const double *init,
                 const double *target, int n_genes, const gro_gene *genes,
                 int n_links, const gro_link *links, double max_error,
                 gro_result *out, double *final_fields);

/* Many models on n_threads host threads (model-parallel, one model per
 * worker at a time, as the reference CPU workers do,
 * errorcalculatormultithread.cpp:30-36,223-229).  Models are given as
 * concatenated gene / link arrays with offsets of length n_models+1. */
int gro_evaluate_batch(const gro_params *p, int n_exp, const double *init,
                       const double *target, int n_models,
                       const int *gene_off, const gro_gene *genes,
                       const int *link_off, const gro_link *links,
                       const double *max_error, int n_threads,
                       gro_result *out);

/* As gro_evaluate_batch with a per-model time step dt[m] (NULL: p->dt). */
int gro_evaluate_batch_dt(const gro_params *p, int n_exp, const double *init,
                          const double *target, int n_models,
                          const int *gene_off, const gro_gene *genes,
                          const int *link_off, const gro_link *links,
                          const double *max_error, const double *dt,
                          int n_threads, gro_result *out);

/* SimState::loadMorphologyImage (simstate.cpp:71-90): decodes n_planes
 * (<= 3) channels of a width x height image of packed 0xAARRGGBB pixels,
 * scan lines top to bottom, into planes out[n_planes][height][width] (x
 * fastest): plane k is byte 2 - k / 255, image line j is plane row
 * height - 1 - j.  Returns 0, or -1 for bad sizes. */
int gro_decode_rgb(int width, int height, const unsigned int *rgb, int n_planes,
                   double *out);

#ifdef __cplusplus
}
#endif
#endif
