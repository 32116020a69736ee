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

#define GRN_ABI_VERSION 3

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
  GRN_E_RESULT = -6,   /* the kernel did not write a result */
  GRN_E_WRONG = -11    /* a known-answer check failed (testPerformance's -11) */
} grn_status;

typedef enum {
  GRN_FITNESS_REF = 0, /* thresholded log error + non-convergence penalty
                          (simstate.cpp:135-158, evaluatormorphogens.cpp:48-76) */
  GRN_FITNESS_MSE = 1  /* mean squared error over cells and target planes */
} grn_fitness_kind;

/* Path selection for grn_evaluate_batch. */
typedef enum {
  GRN_PATH_AUTO = 0,  /* FP64 kernel for datasets with REF fitness and an early-stop
                         threshold (min_conc_change >= 0), else the FP32 kernels
                         tuned per size class */
  GRN_PATH_LDS = 1,   /* FP32, field resident on chip (registers + LDS halo) */
  GRN_PATH_TILED = 2, /* FP32, field in device memory, halo-tiled */
  GRN_PATH_EXACT = 3  /* FP64 state and arithmetic throughout, as the reference
                         (modelsimulator.cpp, cmodeldata.h: double everywhere) */
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

/* SimState::loadMorphologyImage (Src/Simulator/simstate.cpp:71-90): decodes
 * n_planes (<= 3) channels of a width x height image -- packed 0xAARRGGBB
 * pixels (QRgb), scan lines top to bottom -- into planes
 * out[n_planes][height][width] (x fastest): plane k is byte 2 - k of the pixel
 * (red, green, blue) / 255 and image scan line j is plane row
 * height - 1 - j.  Host only (no device needed). */
int grn_decode_rgb(int width, int height, const unsigned int *rgb, int n_planes, double *out);

/* grn_add_dataset from the reference's experiment images
 * (CEvaluatorMorphogens::createGPUExperimentDataset,
 * cevaluatormorphogens.cpp:113-131, which loads the input and output
 * morphologies with loadMorphologyImage): init_images[n_experiments][dim_y]
 * [dim_x] give the n_input_morphogens input planes of each experiment and
 * target_images the n_target_morphogens target planes (each at most 3).
 * Images are dim_x x dim_y pixels, decoded as grn_decode_rgb. */
int grn_add_dataset_rgb(void *ctx, const grn_sim_params *params, int n_experiments,
                        const unsigned int *init_images, const unsigned int *target_images, int *dataset_id);

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

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  grn_evaluate_batch with GRN_PATH_LDS on at most
 * GRN_NOALLOC_MAX_MODELS models, with every batch buffer (programs, work
 * order, maxError bounds, results, work-queue table) placed in one fixed
 * array per device of GRN_NOALLOC_SCRATCH_BYTES bytes, which each device
 * holds until the process ends.  It calls no hipMalloc, hipFree or pinned
 * allocation; its host staging is sized when the context is created.
 * Results equal grn_evaluate_batch(..., GRN_PATH_LDS, ...) on the same
 * arguments.  More models, or a model that needs the halo-tiled or the FP64
 * kernel (device memory per workgroup), fail with GRN_E_CAPACITY.  Every
 * call on a device shares that device's array, which keeps the last call's
 * contents: calls that overlap on one GPU (any contexts, any threads) race.
 * Launches are always eager (no graph is captured). */
#define GRN_NOALLOC_MAX_MODELS 256
#define GRN_NOALLOC_SCRATCH_BYTES 545024
int grn_evaluate_batch_noalloc(void *ctx, int n_models, const int *dataset_ids, const grn_model *models,
                               const double *max_errors, const double *dt_override, grn_result *out);

/* A batch compiled, planned and uploaded once: grn_run_staged evaluates it
 * again and again with the programs resident in device memory (only the
 * results cross the bus), giving the same results as grn_evaluate_batch on
 * the same arguments.  For repeated evaluation of one population (timing,
 * the known-answer self-test); no counterpart in the reference, which
 * uploads every batch.  The handle belongs to ctx; grn_staged_destroy frees
 * it, and grn_destroy of the context frees those still alive.  A handle
 * that was freed either way is rejected by grn_run_staged (GRN_E_INVALID)
 * and ignored by grn_staged_destroy. */
int grn_stage_batch(void *ctx, int n_models, const int *dataset_ids, const grn_model *models,
                    const double *max_errors, const double *dt_override, int path, void **staged);
int grn_run_staged(void *ctx, void *staged, grn_result *out);
void grn_staged_destroy(void *staged);

/* Simulates experiment `experiment` of a dataset for one model and writes
 * the final field, fields[n_products][dim_y][dim_x] as double, capacity in
 * elements.  Runs the same kernel code as grn_evaluate. */
int grn_simulate_fields(void *ctx, int dataset_id, const grn_model *model,
                        double dt_override, int experiment, int path,
                        double *fields, long capacity, int *n_products);

/* ErrorCalculatorMultiThread::testPerformance
 * (Src/UI/Evolution/errorcalculatormultithread.cpp:59-90): evaluates n
 * copies of the model in `model_text` as one batch (maxError 1000, a
 * parentless individual), each fitness compared for exact equality with
 * `expected`.  *max_dif gets the largest |fitness - expected|, *ms the
 * kernel time; returns GRN_E_WRONG when any copy differs. */
int grn_test_performance(void *ctx, int dataset_id, const char *model_text, int n, double expected,
                         double *max_dif, double *ms);

/* No reference counterpart: how grn_evaluate_batch (and the search calls
 * that use it) issue a batch's launches.  GRN_LAUNCH_GRAPH captures the
 * result clearing and the persistent launches in a HIP graph the first time
 * a layout of kernel classes is seen and replays it for later batches of the
 * same layout; every launch reads its work queue from a device table that is
 * uploaded per batch, so results are identical to GRN_LAUNCH_EAGER. */
int grn_set_launch_mode(void *ctx, int mode);

/* ---- autotuner ----------------------------------------------------------
 * No reference counterpart (the reference launches one fixed shape).  A size
 * class is (grid, genes in use, target genes, guard); its candidates are the
 * kernel configurations that can run it: on-chip instantiations (cells per
 * thread, workgroup size) and halo-tiled ones (tile window, halo width). */
typedef enum {
  GRN_TUNE_ONCHIP_EXACT = 0, /* on-chip kernel compiled for this grid width */
  GRN_TUNE_ONCHIP_ANY = 1,   /* on-chip kernel for any width */
  GRN_TUNE_TILED = 2         /* halo-tiled kernel */
} grn_tune_kind;

#define GRN_TUNE_MAX_CANDIDATES 24

typedef struct {
  int kind;                 /* grn_tune_kind */
  int cells_per_thread;
  int block;                /* threads per workgroup */
  int genes_lo, genes_hi;   /* gene range of the instantiation */
  int tile_w, tile_h;       /* halo-tiled: extended tile window (cells) */
  int tile_sx, tile_sy;     /* halo-tiled: interior written back per tile */
  int halo;                 /* halo-tiled: halo width = Euler steps per tile load */
  double ms;                /* kernel time on the class sample (< 0: failed) */
  double cell_steps_per_s;
} grn_tune_candidate;

typedef struct {
  int dim_x, dim_y, genes, n_target_morphogens, guard;
  int n_models;             /* models of the class in the batch */
  int n_timed;              /* models in the timed sample */
  int n_candidates;
  int chosen;               /* index into cand */
  int reserved;
  grn_tune_candidate cand[GRN_TUNE_MAX_CANDIDATES];
} grn_tune_class;

/* Times every candidate of every size class present in the batch on at most
 * sample_models of the class's models, with at most step_cap Euler steps
 * (0: the datasets' own), and keeps the fastest for later GRN_PATH_AUTO
 * evaluations on this context.  Writes up to cap class reports; *n_classes
 * gets the number of classes. */
int grn_autotune(void *ctx, int n_models, const int *dataset_ids, const grn_model *models, int sample_models,
                 int step_cap, grn_tune_class *report, int cap, int *n_classes);

/* Applies saved reports (e.g. from an earlier process): for each class the
 * candidate cand[chosen] is looked up among the configurations this library
 * can run for the class and kept as its choice; GRN_E_INVALID if one is not
 * available.  Nothing is timed. */
int grn_apply_tuning(void *ctx, const grn_tune_class *classes, int n);

/* Forgets the autotuner's choices (the default heuristic applies again). */
int grn_clear_tuning(void *ctx);

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
  int n_generations;        /* SearchParams.nGenerations: the last generation (<= 0: no limit) */
  int reserved;
} grn_search_params;

typedef struct {
  int generation;           /* generations completed */
  int n_evaluated;          /* models evaluated by the last advance */
  double min_fit, mean_fit, max_fit; /* Generation::calcPopulSta over all islands */
  int best_complexity;      /* complexity in use of the best individual */
  int best_island;
  int pareto_size;          /* size of the (error, complexity) Pareto front */
  int done;                 /* 1: the search has reached its last generation */
  int last_generation;      /* n_generations, or 250 generations after the least complex
                               perfect model found (search.cpp:157-167); -1: no limit */
  int best_perfect_complexity; /* complexity of the least complex error-0 model, -1: none yet */
} grn_generation_stats;

/* One member of the Pareto front on (error, complexity in use) that
 * Search::recalculateParetoFront keeps (search.cpp:348-377): sorted by
 * error, then complexity; each member less complex than the one before it,
 * or an exact duplicate (same error and complexity) of it. */
typedef struct {
  double error;
  int complexity;           /* Model::calcComplexityInUse */
  int generation;           /* generation whose evaluation gave its error (0: initial population) */
  int island;               /* island it was bred on */
  int n_genes, n_links;     /* size of its model (grn_search_pareto_model) */
  int reserved;
} grn_pareto_entry;

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

/* The current Pareto front: writes up to cap entries, *n gets its size. */
int grn_search_pareto(void *search, grn_pareto_entry *entries, int cap, int *n);

/* The model of front member i (as grn_search_best writes models). */
int grn_search_pareto_model(void *search, int i, grn_gene *genes, int gene_cap, int *n_genes, grn_link *links,
                            int link_cap, int *n_links);

/* The lowest-error individual over all islands (ties: lower complexity). */
int grn_search_best(void *search, grn_gene *genes, int gene_cap, int *n_genes, grn_link *links,
                    int link_cap, int *n_links, double *error);

/* ---- learned pre-screen of offspring -------------------------------------
 * No reference counterpart (the reference simulates every child).  An MLP
 * (468 features, four GELU layers of 512, one output) with BF16 weights and
 * FP32 accumulation, trained online with Adam on (model, ln error) samples
 * that the simulator produced, ranks candidate offspring so that only the
 * most promising are simulated.  The simulator has the last word: every
 * error the search keeps is a simulated one. */
typedef struct {
  int n_morphogens;         /* structural genes of the encoded models */
  int batch;                /* training rows per Adam step, a multiple of 128 (4096) */
  int capacity;             /* samples kept; the oldest are overwritten (1 << 20) */
  int reserved;
  double lr, beta1, beta2, eps; /* Adam: 1e-3, 0.9, 0.999, 1e-8 */
  unsigned long long seed;  /* weight initialisation and minibatch draws */
} grn_surrogate_params;

typedef struct {
  long samples;             /* samples in the store */
  long train_steps;         /* Adam steps taken */
  double last_loss;         /* mean squared error of the last step (standardised ln error) */
  double mu, sigma;         /* standardisation: target = (ln error - mu) / sigma */
  double train_ms;          /* device time of the last grn_surrogate_train call */
  double predict_ms;        /* device time of the last prediction */
} grn_surrogate_stats;

void grn_surrogate_default_params(grn_surrogate_params *p);
int grn_surrogate_create(int device, const grn_surrogate_params *p, void **surrogate);
void grn_surrogate_destroy(void *surrogate);
/* Length of the flat FP32 parameter vector: W_l^T (512 x 512, out x in) for
 * l = 1..4, b_1..b_4 (512 each), output weights (512), output bias. */
long grn_surrogate_n_params(void);
/* The 468 features of a model (genes in use in label order, at most 12):
 * gene s: lim, deg, 10 dif at [3s..3s+2]; link r -> t at 36 + 3 (12 r + t):
 * type (+1 activator OR, +0.5 activator AND, -1 repressor, 0 none), K, |n|/4. */
int grn_surrogate_encode(const grn_model *model, int n_morphogens, float *features);
/* Adds simulated samples; errors must be finite and > 0 (others are skipped). */
int grn_surrogate_add(void *surrogate, int n, const grn_model *models, const double *errors);
/* Adam steps on minibatches drawn uniformly from the store. */
int grn_surrogate_train(void *surrogate, int steps, double *last_loss);
/* Pair mode (weight > 0; set while the store is empty): the store keeps each
 * simulated child beside its slot parent, and training adds weight x the
 * squared error of the predicted change of ln error from parent to child,
 * the quantity a screen compares among one slot's candidates.  Samples from
 * grn_surrogate_add are stored as pairs of a model with itself.  No
 * counterpart in the reference. */
int grn_surrogate_set_pair_weight(void *surrogate, double weight);
/* Predicted ln error of each model. */
int grn_surrogate_predict(void *surrogate, int n, const grn_model *models, double *log_errors);
/* The same on encoded rows (n x 468 floats). */
int grn_surrogate_predict_features(void *surrogate, int n, const float *features, double *log_errors);
/* One forward and backward pass on n encoded rows with standardised targets,
 * no update: the loss (mean squared error) and its gradient. */
int grn_surrogate_gradient(void *surrogate, int n, const float *features, const double *targets, float *grad,
                           double *loss);
int grn_surrogate_get_params(void *surrogate, float *params);
int grn_surrogate_set_params(void *surrogate, const float *params);
int grn_surrogate_get_stats(void *surrogate, grn_surrogate_stats *out);
/* The surrogate's GEMM on its own: C[M][N] = A[M][K] . Bt[N][K]^T with BF16
 * operands given as raw 16-bit patterns, FP32 result; M, N multiples of 128,
 * K of 32.  For tests. */
int grn_mlp_gemm_bf16(int device, int M, int N, int K, const unsigned short *A, const unsigned short *Bt, float *C);

typedef struct {
  int n_candidates;         /* offspring scored by the surrogate */
  int n_simulated;          /* offspring simulated */
  double spearman;          /* rank correlation, predicted vs simulated ln error (simulated offspring) */
  double mae;               /* mean |predicted - simulated| ln error (simulated offspring) */
  double loss;              /* loss of the last completed training call */
  double predict_ms;        /* wall time of encoding and scoring the candidates */
  double train_ms;          /* device time of the last completed training call (training
                               overlaps the host's breeding and finishes before the next prediction) */
  double simulate_ms;       /* wall time of simulating the kept children */
  double breed_ms;          /* wall time of selection, migration and breeding */
  long samples;             /* samples in the surrogate's store */
  double record_ms;         /* wall time of storing the simulated children as samples and
                               enqueueing the training steps */
} grn_screen_stats;

/* Candidates per offspring slot (1 = the reference's search).  With k > 1
 * every parent pair breeds k candidate pairs; grn_search_screen keeps, per
 * slot, the candidate with the lowest score. */
int grn_search_set_oversample(void *search, int k);
/* Candidates awaiting grn_search_screen (valid until the next call). */
int grn_search_candidates(void *search, int *n, const grn_model **models);
/* Scores (lower is better) of the candidates, in order; fills pending. */
int grn_search_screen(void *search, const double *scores);
/* Fraction (0, 1] of each island's slots that grn_search_screen sends to the
 * simulator: those whose kept candidate's score most undercuts the parent's
 * recorded ln raw error (grn_search_record_raw).  The other slots keep their
 * parent, as a losing child would.  Default 1. */
int grn_search_set_screen_fraction(void *search, double fraction);
/* What the surrogate scores in grn_search_run_generation_screened, and how a
 * slot's child is picked:
 * GRN_SCREEN_ABSOLUTE, a candidate's ln raw error from its own features; the
 *   lowest score is kept;
 * GRN_SCREEN_DELTA, the change of ln raw error from its slot parent, from the
 *   difference of their features and the parent's error (the surrogate then
 *   learns from children with a recorded parent only); the lowest is kept;
 * GRN_SCREEN_HALF, scored as ABSOLUTE, but the child is drawn uniformly
 *   (a seeded hash of generation, island and slot) from the better half of
 *   the slot's non-clone candidates: the surrogate only rejects.
 * Default absolute. */
enum { GRN_SCREEN_ABSOLUTE = 0, GRN_SCREEN_DELTA = 1, GRN_SCREEN_HALF = 2 };
int grn_search_set_screen_policy(void *search, int policy);
/* Records the raw errors of the pending models (call before
 * grn_search_advance): their ln is what grn_search_set_screen_fraction
 * compares a child's score with. */
int grn_search_record_raw(void *search, const double *raw_errors);
/* One screened generation: predict the candidates, keep the best per slot
 * (passing over clones of the parent), simulate them on ctx, add them to the
 * surrogate's samples, enqueue train_steps Adam steps -- they run on the
 * device while the host selects, migrates and breeds the next candidates --
 * and return.  The initial population is simulated unscreened. */
int grn_search_run_generation_screened(void *search, void *ctx, void *surrogate, int train_steps, int path,
                                       grn_generation_stats *stats, grn_screen_stats *screen);
/* One generation screened by a coarse simulation instead of the learned
 * surrogate (a multi-fidelity screen; the reference has no counterpart, its
 * search simulates every child): every candidate is simulated on
 * coarse_datasets[island] -- the island's experiments on a coarser grid --
 * with each diffusion constant multiplied by dif_scale (for a grid coarser
 * by a factor f in both directions, 1/f^2 keeps the physical diffusion) and
 * its full-size time step; its ln raw error is the score grn_search_screen
 * ranks (under the screen policy); the kept children are then simulated on
 * their own datasets and the search advances.  The initial population is
 * simulated unscreened.  coarse_ms (optional) receives the wall time of the
 * coarse pass and the screen. */
int grn_search_run_generation_coarse(void *search, void *ctx, const int *coarse_datasets, double dif_scale,
                                     int path, grn_generation_stats *stats, double *coarse_ms);
/* One plain generation simulated at another fidelity (no counterpart in the
 * reference, whose search simulates every model at one size): the pending
 * models run on island_datasets[island] with each diffusion constant
 * multiplied by dif_scale, their own time steps and maxErrors, and the
 * search advances on those errors.  A search run this way for its first
 * generations on a coarse grid, then grn_search_rescore'd, continues at full
 * size from the coarse population (a coarse-to-fine schedule). */
int grn_search_run_generation_scaled(void *search, void *ctx, const int *island_datasets, double dif_scale,
                                     int path, grn_generation_stats *stats);
/* Makes every island member pending again, as an initial population is:
 * the children or candidates bred for the next generation are dropped, the
 * Pareto front and the perfect-model end are cleared, and the next
 * generation records the members' errors without selection, then breeds.
 * Used after a change of fidelity (grn_search_run_generation_scaled). */
int grn_search_rescore(void *search);

#ifdef __cplusplus
}
#endif
#endif /* GRN_H */
