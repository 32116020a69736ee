/*
 * grn.h -- C ABI of the MI355X-native GRN simulation-and-fitness library
 * (libgrn_hip.so).  Plain C types only: no HIP, torch or C++ types cross it.
 *
 * Every entry point returns GRN_OK (0) or a negative grn_status; the message
 * of the last failure on the calling thread is available from
 * grn_last_error().  No C++ exception crosses the ABI.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   grn_create + grn_add_dataset
 *       CEvaluatorMorphogens::CEvaluatorMorphogens(int threadId, const Search&)
 *       Src/GPUSimulator/cevaluatormorphogens.h:23, .cpp:35-56 (device select,
 *       stream) and createGPUExperimentDataset/copySimState (.cpp:92-191:
 *       SimParams scalars + initial and target planes to the device)
 *   grn_evaluate
 *       double CEvaluatorMorphogens::evaluate(const Model&, double maxError)
 *       cevaluatormorphogens.h:26, .cpp:233-245; with
 *       CModelSimulator::loadModel (cmodelsimulator.cpp:114-227) and
 *       CModelSimulator::simulateExperiment + launchKernel
 *       (cmodelsimulator.cpp:229-242, modelsimulatordevice.cu:253-264):
 *       one model per launch, synchronous, the reference's calling pattern
 *   grn_evaluate_batch
 *       no single reference call: it replaces the per-model loop that
 *       ErrorCalculatorMultiThread's GPU workers run over one deme's queue
 *       (Src/UI/Evolution/errorcalculatormultithread.cpp:175-199,248-254);
 *       results are per-model identical to grn_evaluate
 *   grn_simulate_fields
 *       the final SimState a Simulator leaves behind (Src/Simulator/
 *       simulator.cpp:54-63) -- used for per-cell parity checks
 *   grn_destroy
 *       CEvaluatorMorphogens::~CEvaluatorMorphogens / deleteGPUExperimentDatasets
 *       (cevaluatormorphogens.cpp:58-61,193-231)
 */
#ifndef GRN_H
#define GRN_H

#ifdef __cplusplus
extern "C" {
#endif

#define GRN_ABI_VERSION 2

/* Capacity of the device program (compile-time limits of the kernels). */
#define GRN_MAX_GENES 16  /* genes in use per model */
#define GRN_MAX_LINKS 96  /* links into in-use genes per model */

typedef enum {
  GRN_OK = 0,
  GRN_E_INVALID = -1,  /* bad argument */
  GRN_E_MODEL = -2,    /* malformed model (missing product for a label) */
  GRN_E_CAPACITY = -3, /* model or grid beyond the kernels' limits */
  GRN_E_DEVICE = -4,   /* HIP runtime error */
  GRN_E_NOMEM = -5,    /* host or device allocation failed */
  GRN_E_RESULT = -6    /* the kernel did not write a result */
} grn_status;

typedef enum {
  GRN_FITNESS_REF = 0, /* thresholded log error + non-convergence penalty
                          (simstate.cpp:135-158, evaluatormorphogens.cpp:48-76) */
  GRN_FITNESS_MSE = 1  /* mean squared error over cells and target planes */
} grn_fitness_kind;

/* Path selection for grn_evaluate_batch. */
typedef enum {
  GRN_PATH_AUTO = 0,  /* tuned per size class */
  GRN_PATH_LDS = 1,   /* field resident on chip (registers + LDS halo) */
  GRN_PATH_TILED = 2  /* field in device memory, halo-tiled */
} grn_path;

/* The SimParams fields the simulation and fitness read
 * (Src/Simulator/simparams.h:25-43). */
typedef struct {
  int dim_x, dim_y;            /* size.width(), size.height() */
  int n_input_morphogens;      /* nInputMorphogens */
  int n_target_morphogens;     /* nTargetMorphogens */
  int num_sim_steps;           /* NumSimSteps */
  double dt;                   /* dt */
  double min_conc;             /* minConc */
  double min_conc_change;      /* minConcChange (< 0: run every step) */
  double dist_error_threshold; /* distErrorThreshold */
  int error_precision;         /* errorPrecision (ceilS digits; <= 0: none) */
  int fitness_kind;            /* grn_fitness_kind */
} grn_sim_params;

/* ModelProd (Src/Model/modelprod.h:10-48). init is carried but unused by the
 * simulation, as in the reference (modelprod.cpp:61-70). */
typedef struct {
  int label;
  double init, lim, deg, dif;
} grn_gene;

/* ModelLink (Src/Model/modellink.h:10-46): hill_coef >= 0 activates,
 * < 0 represses with Hill exponent |hill_coef|. */
typedef struct {
  int regulator, regulated; /* product labels */
  double dis_const, hill_coef;
  int is_and; /* AND (1) or OR (0) regulation */
} grn_link;

/* Model (Src/Model/model.h:14-91): caller-owned, read-only. */
typedef struct {
  int n_genes;
  const grn_gene *genes;
  int n_links;
  const grn_link *links;
} grn_model;

typedef struct {
  double fitness;   /* 100 if raw_error > max_error, else ceilS(raw_error) */
  double raw_error; /* error accumulated over experiments */
  double change;    /* accumulated last-step max |dc/dt| over experiments */
  int n_products;   /* genes in use */
  int steps;        /* Euler steps executed, summed over experiments */
  int experiments;  /* experiments simulated before the max_error cut */
  int reserved;
} grn_result;

typedef struct {
  double kernel_ms;       /* device time of the last evaluate call (events) */
  double host_ms;         /* wall time of the last evaluate call */
  double cell_steps;      /* sum over models of W*H*steps executed */
  double alg_flops;       /* sum over models of cell-steps * F(model) */
  int n_launches;         /* simulation launches of the last call (one per kernel class) */
  int n_models;
  int launch_mode;        /* grn_launch_mode the last call ran with */
  int graph_replayed;     /* 1: the last call replayed a graph captured earlier */
} grn_stats;

typedef enum {
  GRN_LAUNCH_EAGER = 0, /* launches issued one by one on the context's stream */
  GRN_LAUNCH_GRAPH = 1  /* launches captured once per kernel-class layout, then replayed */
} grn_launch_mode;

/* Library ABI version (GRN_ABI_VERSION). */
int grn_abi_version(void);

/* Message of the last failure on this thread ("" if none). */
const char *grn_last_error(void);

/* Number of visible HIP devices. */
int grn_device_count(int *n);

/* Selects the device, creates the context's stream and device buffers. */
int grn_create(int device, void **ctx);

/* Frees everything the context owns. Safe on NULL. */
void grn_destroy(void *ctx);

/* Uploads one experiment set: params plus initial planes
 * init[n_experiments][n_input_morphogens][dim_y][dim_x] and target planes
 * target[n_experiments][n_target_morphogens][dim_y][dim_x] (x fastest).
 * Returns the dataset id used by the evaluate calls. */
int grn_add_dataset(void *ctx, const grn_sim_params *params, int n_experiments,
                    const double *init, const double *target, int *dataset_id);

/* Plain path: compile, launch one workgroup, synchronize -- one model per
 * launch as the reference does. */
int grn_evaluate(void *ctx, int dataset_id, const grn_model *model,
                 double max_error, grn_result *out);

/* Batched path: every model of the batch in persistent launches (one per
 * size class), synchronizing once.  dataset_ids[i] selects model i's
 * experiment set; dt_override (optional) replaces the dataset's dt for model
 * i when > 0; max_errors (optional, default +inf). */
int grn_evaluate_batch(void *ctx, int n_models, const int *dataset_ids,
                       const grn_model *models, const double *max_errors,
                       const double *dt_override, int path, grn_result *out);

/* Simulates experiment `experiment` of a dataset for one model and writes
 * the final field, fields[n_products][dim_y][dim_x] as double, capacity in
 * elements.  Runs the same kernel code as grn_evaluate. */
int grn_simulate_fields(void *ctx, int dataset_id, const grn_model *model,
                        double dt_override, int experiment, int path,
                        double *fields, long capacity, int *n_products);

/* No reference counterpart: how grn_evaluate_batch (and the search calls
 * that use it) issue a batch's launches.  GRN_LAUNCH_GRAPH captures the
 * result clearing and the persistent launches in a HIP graph the first time
 * a layout of kernel classes is seen and replays it for later batches of the
 * same layout; every launch reads its work queue from a device table that is
 * uploaded per batch, so results are identical to GRN_LAUNCH_EAGER. */
int grn_set_launch_mode(void *ctx, int mode);

/* Timing and work counters of the last evaluate call on this context. */
int grn_get_stats(void *ctx, grn_stats *out);

/* No reference counterpart: copies `bytes` (a multiple of 16) device to
 * device with a plain copy kernel on the current device and returns its
 * time.  Profilers use it to calibrate HBM byte counters. */
int grn_copy_probe(long bytes, double *ms);

/* MathAlgo::ceilS (Src/Common/mathalgo.h:55-62). */
double grn_ceil_s(double n, int digits);

/* Model text format, Model::operator>> (Src/Model/model.cpp:659-706,
 * modelprod.cpp:81-106, modellink.cpp:175-202):
 *   (label init lim deg dif|...*regulator regulated K n isAnd|...)
 * Writes up to gene_cap genes and link_cap links; *n_genes / *n_links get
 * the model's counts even when the capacity is too small (GRN_E_CAPACITY). */
int grn_model_parse(const char *text, grn_gene *genes, int gene_cap, int *n_genes, grn_link *links,
                    int link_cap, int *n_links);

/* Model::operator<< (Src/Model/model.cpp:627-657): numbers with `precision`
 * significant digits (<= 0: 6, the reference stream's default).  *len gets
 * the text length without the terminating NUL. */
int grn_model_format(const grn_model *model, int precision, char *buf, long cap, long *len);

/* Compiles a model for a dataset with n_morphogens structural genes and
 * reports genes in use, Hill links evaluated and the algorithmic FLOPs per
 * cell-step, F = 5L + (L - G_h) + 8G + 7D. */
int grn_model_info(const grn_model *model, int n_morphogens, int *genes_in_use, int *hill_links,
                   double *flops_per_cell_step);

/* ---- evolutionary search: the caller of the path ------------------------
 * Island populations evolved by deterministic crowding with the reference's
 * operators (Src/Search/searchalgodetcrowd.cpp:54-147, Src/Model/model.cpp
 * 21-185, 285-333, 472-574) and migration between islands
 * (Src/Search/search.cpp:263-311).  The search is driven in phases so any
 * evaluator can serve it: grn_search_pending() lists the models that need an
 * error, grn_search_advance() takes their errors and produces the next
 * generation's children; grn_search_run_generation() does both with
 * grn_evaluate_batch on a context. */

typedef enum {
  GRN_MIGRATION_PAIRWISE = 0, /* random island pairs, union shuffled and split (search.cpp:263-311) */
  GRN_MIGRATION_RING = 1,     /* best fraction of island i replaces the worst of island i+1 */
  GRN_MIGRATION_NONE = 2
} grn_migration;

typedef struct {
  int n_islands;            /* SearchParams.nDemes */
  int island_size;          /* SearchParams.demesSize, rounded up to even */
  int n_morphogens;         /* nInputMorphogens + nTargetMorphogens (= nMinProducts) */
  int n_max_products;       /* SimParams.nMaxProducts, -1 = no limit */
  int n_max_links;          /* SimParams.nMaxLinks, -1 = no limit */
  int init_products_min;    /* Model::createRandom sizes of the initial population */
  int init_products_max;
  int migration;            /* grn_migration */
  int migration_period;     /* generations between migrations (SearchParams.migrationPeriod) */
  double migration_fraction;/* ring: fraction of an island that emigrates */
  double crossover_rate;    /* 0.75 in the reference (searchalgodetcrowd.cpp:74) */
  /* distributions of new genes and links and of parameter mutations; the
   * reference draws lim, deg, dif, K from U(0,1), |n| from U(1,5), and makes
   * a new gene diffusing with probability 0.1 (modelprod.cpp:14-24) */
  double lim_lo, lim_hi, deg_lo, deg_hi;
  double dif_prob, dif_lo, dif_hi;
  double k_lo, k_hi, n_lo, n_hi;
  int dt_rule;              /* 0: the dataset's dt; 1: dt_scale / max dif, at most dt_cap */
  double dt_scale, dt_cap;
  unsigned long long seed;
} grn_search_params;

typedef struct {
  int generation;           /* generations completed */
  int n_evaluated;          /* models evaluated by the last advance */
  double min_fit, mean_fit, max_fit; /* Generation::calcPopulSta over all islands */
  int best_complexity;      /* complexity in use of the best individual */
  int best_island;
  int pareto_size;          /* size of the (error, complexity) Pareto front */
  int reserved;
} grn_generation_stats;

/* The reference's defaults for everything grn_search_params holds. */
void grn_search_default_params(grn_search_params *p);

/* island_datasets[i]: the dataset id every model of island i is evaluated on. */
int grn_search_create(const grn_search_params *p, const int *island_datasets, void **search);
void grn_search_destroy(void *search);

/* Models that need an error now, with the dataset, maxError (the parent's
 * error, 1000 for the initial population, individual.cpp:19-25) and time
 * step (0 = the dataset's) of each.  Pointers stay valid until the next
 * grn_search_advance. */
int grn_search_pending(void *search, int *n, const grn_model **models, const int **datasets,
                       const double **max_errors, const double **dt);

/* Takes the pending models' errors (fitness as grn_result.fitness returns
 * it), selects (a child replaces its parent when child <= parent,
 * searchalgodetcrowd.cpp:119,128), migrates when due and reproduces. */
int grn_search_advance(void *search, const double *errors, grn_generation_stats *stats);

/* grn_search_pending + grn_evaluate_batch on ctx + grn_search_advance. */
int grn_search_run_generation(void *search, void *ctx, int path, grn_generation_stats *stats);

/* The lowest-error individual over all islands (ties: lower complexity). */
int grn_search_best(void *search, grn_gene *genes, int gene_cap, int *n_genes, grn_link *links,
                    int link_cap, int *n_links, double *error);

#ifdef __cplusplus
}
#endif
#endif /* GRN_H */
