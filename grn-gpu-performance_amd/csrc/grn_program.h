// grn_program.h -- device-side layout of a compiled GRN model and of an
// uploaded experiment set.  Shared by the host compiler (grn_compile.cpp),
// the host launcher (grn_host.cpp) and the kernels (grn_kernels.hip).
//
// A model is compiled on the host into a GrnProgram: the in-use genes in
// ascending label order (Src/Simulator/modelsimulator.cpp:71-80), their
// lim/deg/dif constants, and for every gene whose links include an activator
// (modelsimulator.cpp:120-138) its OR links followed by its AND links, each
// in model order (modelsimulator.cpp:156-192).
//
// Hill terms are carried in exponent form.  For a link with regulator
// concentration x, dissociation constant K and exponent n = |hillCoef|:
//   activation  x^n/(x^n+K^n) = 1/(1+2^t),  t = -n*log2(x) + n*log2(K)
//   repression  K^n/(x^n+K^n) = 1/(1+2^t),  t = +n*log2(x) - n*log2(K)
// so a link is (r, a, b) with t = a*log2(x_r) + b.  Because 1/(1+2^t) falls
// monotonically in t, the reference's fold h = min(max_OR H, min_AND H)
// (simopor.cpp, simopand.cpp) becomes one exponential per gene:
//   h = 1/(1 + 2^T),  T = max(min_OR t, max_AND t)
// with an empty OR or AND side dropped.
#ifndef GRN_PROGRAM_H
#define GRN_PROGRAM_H

#define GRN_PMAX_GENES 16
#define GRN_PMAX_LINKS 96

struct GrnLink {
  int r;    // regulator gene index (0..G-1)
  float a;  // t = a*log2(x) + b
  float b;
  int pad;
};

// Per-gene constants.  tor0 / tand0 seed the OR-min and AND-max folds so
// that T = max(tor, tand) needs no branches:
//   OR links present:  tor0 = +inf (min identity), else -inf (side dropped)
//   Hill links present: tand0 = -inf (max identity)
//   no activator among the gene's links (modelsimulator.cpp:120-138: h = 0):
//     no links stored, tor0 = -inf, tand0 = +inf, so T = +inf and h = 0.
struct GrnGene {
  float lim, deg, dif;
  float tor0, tand0;
  int link_begin, n_or, n_and;
};

struct GrnProgram {
  int G;              // genes in use (every gene has an LDS plane)
  int n_links;        // links stored below
  int dataset;        // experiment set
  float dt;           // Euler time step
  int reg_mask;       // bit g: gene g regulates a stored link (its log2 is needed)
  int hill_mask;      // bit g: gene g has an activator (else h = 0: no Hill terms)
  int dif_mask;       // bit g: gene g diffuses (its LDS plane is read)
  int fast;           // 1: the state provably stays >= 0 and no Hill exponent is 0,
                      // so the step needs neither the minConc clamp nor the log2 floor
  GrnGene gene[GRN_PMAX_GENES];
  GrnLink link[GRN_PMAX_LINKS];
};

// Double-precision constants of a compiled model for the FP64 kernel
// (grn_exact.hip), in the same gene and link order as its GrnProgram (which
// still provides G, the link ranges, the regulator indices and the masks):
// the same exponent-form Hill links, t = a*log2(x) + b, without rounding a
// and b to float.
struct GrnLink64 {
  double a, b;
};

struct GrnGene64 {
  double lim, deg, dif;
};

struct GrnProgram64 {
  double dt;
  double pad;
  GrnGene64 gene[GRN_PMAX_GENES];
  GrnLink64 link[GRN_PMAX_LINKS];
};

struct GrnDataset {
  int W, H;
  int n_in, n_tgt;
  int n_exp, steps;
  int kind;       // 0 = REF, 1 = MSE
  int nonneg;     // inputs >= 0 and min_conc >= 0: the diffusion guard is moot
  float min_conc;
  float min_conc_change;
  double min_conc_d;         // the same two thresholds as given (double)
  double min_conc_change_d;
  double thr;     // distErrorThreshold
  const double *init;    // [n_exp][n_in][H][W]
  const double *target;  // [n_exp][n_tgt][H][W]
};

struct GrnModelOut {
  double err;     // accumulated error (before sentinel / rounding)
  double change;  // accumulated last-step max change
  int steps;      // Euler steps executed, summed over experiments
  int exps;       // experiments simulated
  int done;       // 1 once written
  int pad;
};

// One launch's work queue, in device memory so that a captured launch can be
// replayed on a new batch: only this table and the program arrays change.
struct GrnQueue {
  int head;   // next queue position to take; zero at launch
  int n;      // models in this queue
  int begin;  // first slot of this queue in the work-order array
  int pad;
};

#define GRN_MAX_QUEUES 64

// Per-batch table uploaded before the launches.
struct GrnBatch {
  int n_total;  // models in the batch (results to clear)
  int n_queues;
  int pad[2];
  GrnQueue q[GRN_MAX_QUEUES];
};

// Halo-tiled kernels keep a model's state in device memory as one record of
// grn_cell_rec(G) floats per cell: hi[0..G), lo[0..G), zero padding to a
// multiple of four floats (16-byte loads and stores).  (constexpr: callable
// from host and device code.)
constexpr int grn_cell_rec(int G) { return (2 * G + 3) & ~3; }

struct GrnLaunch {
  const GrnProgram *progs;  // [n_total]
  const GrnProgram64 *progs64;  // [n_total] (FP64 kernel only; entries of other models unset)
  GrnQueue *queue;          // this launch's queue; the work order comes as a kernel argument
  const GrnDataset *datasets;
  const double *max_err;    // [n]
  GrnModelOut *out;         // [n]
  double *fields;           // optional: [n][G][H][W] final field
  long field_stride;        // elements per model in fields
  int field_exp;            // experiment whose final field is written
  int step_cap;             // > 0: at most this many Euler steps (launch timing); 0: the dataset's
  // halo-tiled kernel only: device-memory state and tile geometry
  float *scratch;           // per workgroup: 4 planes sets (hi, lo) x (A, B)
  long slot_floats;         // floats per workgroup slot
  int tile_sx, tile_sy;     // interior of a tile (cells written back)
  int tile_halo;            // halo width = Euler steps per tile load
  int tile_w, tile_h;       // extended tile (interior + halos, padded)
  int pad2;
};

#endif
