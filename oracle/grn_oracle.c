/*
 * grn_oracle.c -- FP64 CPU restatement of the reference GRN simulator and
 * fitness calculator (see grn_oracle.h for the file:line map).
 * TEST INFRASTRUCTURE ONLY: this is the checker, never the product.
 *
 * The code keeps the reference's loop order, double arithmetic, pow() Hill
 * terms and the op sequence that ModelSimulator::loadModel builds, so its
 * results are those of the reference CPU path up to libm differences.
 */
#include "grn_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- MathAlgo helpers (mathalgo.h:55-62, 208-216) ---------------------- */

double gro_ceil_s(double n, int d) {
  if (n == 0.0) return 0.0;
  double factor = pow(10.0, d - ceil(log10(fabs(n))));
  return ceil(n * factor) / factor;
}

static inline double ma_max(double a, double b) { return (a < b) ? b : a; }
static inline double ma_min(double a, double b) { return (!(b < a)) ? a : b; }

/* ---- in-use closure (model.cpp:243-283) -------------------------------- */

static int contains(const int *v, int n, int x) {
  for (int i = 0; i < n; ++i)
    if (v[i] == x) return 1;
  return 0;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

int gro_labels_in_use(int n_genes, const gro_gene *genes, int n_links,
                      const gro_link *links, int n_morphogens, int *out,
                      int cap) {
  (void)n_genes;
  (void)genes;
  int n = 0;
  /* All structural labels 0..nMorphogens-1 are in use (model.cpp:249-252). */
  for (int i = 0; i < n_morphogens; ++i) {
    if (n >= cap) return -1;
    out[n++] = i;
  }
  /* Breadth-first over regulators (model.cpp:254-267): the visit queue is
   * the tail of out[] not yet expanded. */
  int head = 0;
  while (head < n) {
    int lab = out[head++];
    for (int l = 0; l < n_links; ++l) {
      if (links[l].regulated != lab) continue;
      int r = links[l].regulator;
      if (!contains(out, n, r)) {
        if (n >= cap) return -1;
        out[n++] = r;
      }
    }
  }
  qsort(out, (size_t)n, sizeof(int), cmp_int);
  return n;
}

/* ---- op program (modelsimulator.cpp:71-203) ---------------------------- */

enum { OP_ZERO, OP_HILL_ACT, OP_HILL_REP, OP_OR, OP_AND };
#define TO_TEMP (-1)

typedef struct {
  int type;
  int from;      /* product index of the regulator (Hill ops) */
  int to;        /* product index, or TO_TEMP for ratiosTempHill_ */
  double hill;   /* |hillCoef| */
  double dis_n;  /* pow(disConst, |hillCoef|), simophillact.cpp:11 */
} op_t;

typedef struct {
  int n_products;
  int labels[64];
  double lim[64], deg[64];
  int n_dif;
  int dif_ind[64];
  double dif_const[64];
  int n_ops;
  op_t *ops;
} program_t;

static int label_index(const program_t *pg, int label) {
  for (int i = 0; i < pg->n_products; ++i)
    if (pg->labels[i] == label) return i;
  return -1;
}

/* createHillOpForLink (modelsimulator.cpp:194-203). */
static op_t hill_op(const program_t *pg, const gro_link *l, int to) {
  op_t o;
  o.from = label_index(pg, l->regulator);
  o.to = to;
  if (l->hill_coef >= 0) {
    o.type = OP_HILL_ACT;
    o.hill = l->hill_coef;
  } else {
    o.type = OP_HILL_REP;
    o.hill = -1.0 * l->hill_coef;
  }
  o.dis_n = pow(l->dis_const, o.hill);
  return o;
}

static int compile_model(int n_genes, const gro_gene *genes, int n_links,
                         const gro_link *links, int n_morph, program_t *pg) {
  pg->ops = NULL;
  int n = gro_labels_in_use(n_genes, genes, n_links, links, n_morph,
                            pg->labels, 64);
  if (n < 0) return -2;
  pg->n_products = n;
  pg->n_dif = 0;
  pg->n_ops = 0;
  pg->ops = (op_t *)malloc(sizeof(op_t) * (size_t)(2 * n_links + n + 1));
  if (!pg->ops) return -3;
  int *or_l = (int *)malloc(sizeof(int) * (size_t)(n_links + 1));
  int *and_l = (int *)malloc(sizeof(int) * (size_t)(n_links + 1));
  if (!or_l || !and_l) {
    free(or_l);
    free(and_l);
    return -3;
  }

  for (int i = 0; i < n; ++i) {
    /* prodWithLabel: the first product carrying the label (model.cpp:227-240) */
    const gro_gene *prod = NULL;
    for (int g = 0; g < n_genes; ++g)
      if (genes[g].label == pg->labels[i]) {
        prod = &genes[g];
        break;
      }
    if (!prod) {
      free(or_l);
      free(and_l);
      return -4;
    }
    pg->lim[i] = prod->lim;
    pg->deg[i] = prod->deg;
    if (prod->dif > 0) {
      pg->dif_ind[pg->n_dif] = i;
      pg->dif_const[pg->n_dif] = prod->dif;
      pg->n_dif++;
    }

    /* linksToLabel keeps model order (model.cpp:391-399); split into OR and
     * AND lists; note whether any link is an activator
     * (modelsimulator.cpp:114-133). */
    int n_or = 0, n_and = 0, any_act = 0;
    for (int l = 0; l < n_links; ++l) {
      if (links[l].regulated != pg->labels[i]) continue;
      if (label_index(pg, links[l].regulator) >= 0) {
        if (links[l].is_and)
          and_l[n_and++] = l;
        else
          or_l[n_or++] = l;
      }
      if (links[l].hill_coef >= 0) any_act = 1;
    }

    if (any_act) {
      /* createProductOps (modelsimulator.cpp:156-192) */
      int temp_used = 0;
      for (int j = 0; j < n_or; ++j) {
        if (temp_used) {
          pg->ops[pg->n_ops++] = hill_op(pg, &links[or_l[j]], TO_TEMP);
          op_t o = {OP_OR, -1, i, 0, 0};
          pg->ops[pg->n_ops++] = o;
        } else {
          pg->ops[pg->n_ops++] = hill_op(pg, &links[or_l[j]], i);
          temp_used = 1;
        }
      }
      for (int j = 0; j < n_and; ++j) {
        if (temp_used) {
          pg->ops[pg->n_ops++] = hill_op(pg, &links[and_l[j]], TO_TEMP);
          op_t o = {OP_AND, -1, i, 0, 0};
          pg->ops[pg->n_ops++] = o;
        } else {
          pg->ops[pg->n_ops++] = hill_op(pg, &links[and_l[j]], i);
          temp_used = 1;
        }
      }
      if (!temp_used) {
        op_t o = {OP_ZERO, -1, i, 0, 0};
        pg->ops[pg->n_ops++] = o;
      }
    } else {
      op_t o = {OP_ZERO, -1, i, 0, 0};
      pg->ops[pg->n_ops++] = o;
    }
  }
  free(or_l);
  free(and_l);
  return 0;
}

/* ---- one Euler step (modelsimulator.cpp:205-280) ----------------------- */

static double simulate_step(const program_t *pg, const gro_params *p,
                            double dt, double *state, double *snap) {
  const int W = p->dim_x, H = p->dim_y, N = W * H;
  double old[64], ratios[64], temp = 0.0;
  double max_change = 0;

  /* Snapshot of the diffusing planes (modelsimulator.cpp:210-211). */
  for (int d = 0; d < pg->n_dif; ++d)
    memcpy(snap + (size_t)d * N, state + (size_t)pg->dif_ind[d] * N,
           sizeof(double) * (size_t)N);

  for (int j = 0; j < H; ++j) {
    for (int i = 0; i < W; ++i) {
      const size_t cell = (size_t)j * W + i;
      for (int k = 0; k < pg->n_products; ++k) old[k] = state[(size_t)k * N + cell];

      for (int o = 0; o < pg->n_ops; ++o) {
        const op_t *op = &pg->ops[o];
        double *to = (op->to == TO_TEMP) ? &temp : &ratios[op->to];
        switch (op->type) {
          case OP_ZERO:
            *to = 0;
            break;
          case OP_HILL_ACT: {
            double rcn = pow(old[op->from], op->hill);
            *to = rcn / (rcn + op->dis_n);
          } break;
          case OP_HILL_REP: {
            double rcn = pow(old[op->from], op->hill);
            *to = op->dis_n / (rcn + op->dis_n);
          } break;
          case OP_OR:
            *to = ma_max(temp, *to);
            break;
          case OP_AND:
            *to = ma_min(temp, *to);
            break;
        }
      }

      for (int k = 0; k < pg->n_products; ++k)
        ratios[k] = pg->lim[k] * ratios[k] - pg->deg[k] * old[k];

      for (int d = 0; d < pg->n_dif; ++d) {
        const double *prod = snap + (size_t)d * N;
        double diffusion = 0, total = 0;
        if (i > 0) { diffusion += prod[cell - 1]; ++total; }
        if (i < W - 1) { diffusion += prod[cell + 1]; ++total; }
        if (j > 0) { diffusion += prod[cell - W]; ++total; }
        if (j < H - 1) { diffusion += prod[cell + W]; ++total; }
        int k = pg->dif_ind[d];
        if (diffusion > 0 || (total > 0 && old[k] > 0))
          ratios[k] += pg->dif_const[d] * (diffusion - total * old[k]);
      }

      for (int k = 0; k < pg->n_products; ++k) {
        double ratio = ratios[k];
        if (ratio) {
          double c = old[k] + dt * ratio;
          state[(size_t)k * N + cell] = (c < p->min_conc) ? 0 : c;
          max_change = ma_max(max_change, fabs(ratio));
        }
      }
    }
  }
  return max_change;
}

/* ---- evaluate (evaluatormorphogens.cpp:48-76) -------------------------- */

static int evaluate_dt(const gro_params *p, double dt, int n_exp,
                       const double *init, const double *target, int n_genes,
                       const gro_gene *genes, int n_links,
                       const gro_link *links, double max_error,
                       gro_result *out, double *final_fields) {
  const int W = p->dim_x, H = p->dim_y, N = W * H;
  const int n_morph = p->n_in + p->n_tgt;
  program_t pg;
  memset(out, 0, sizeof(*out));
  int rc = compile_model(n_genes, genes, n_links, links, n_morph, &pg);
  if (rc) {
    free(pg.ops);
    return rc;
  }
  const int G = pg.n_products;
  double *state = (double *)malloc(sizeof(double) * (size_t)G * N);
  double *snap = (double *)malloc(sizeof(double) * (size_t)(pg.n_dif + 1) * N);
  if (!state || !snap) {
    free(state);
    free(snap);
    free(pg.ops);
    return -3;
  }

  double error = 0.0, change = 0.0;
  int e = 0, steps = 0;
  while (e < n_exp && error <= max_error) {
    /* Initial state: input planes for k < nIn, zero elsewhere
     * (simulator.cpp:35-52; simstate.cpp:64-90). */
    memset(state, 0, sizeof(double) * (size_t)G * N);
    for (int k = 0; k < p->n_in && k < G; ++k)
      memcpy(state + (size_t)k * N, init + ((size_t)e * p->n_in + k) * N,
             sizeof(double) * (size_t)N);

    /* Simulator::simulate (simulator.cpp:54-63) */
    double chg = 1;
    int s = 0;
    while (s < p->n_steps && chg > p->min_conc_change) {
      chg = simulate_step(&pg, p, dt, state, snap);
      ++s;
    }
    steps += s;
    change += chg;

    double exp_error = 0.0;
    const double *tgt = target + (size_t)e * p->n_tgt * N;
    if (p->fitness_kind == GRO_FITNESS_MSE) {
      /* Mean squared error over cells and target planes (no penalty). */
      double acc = 0.0;
      for (int k = 0; k < p->n_tgt; ++k)
        for (int c = 0; c < N; ++c) {
          double d = state[(size_t)(k + p->n_in) * N + c] - tgt[(size_t)k * N + c];
          acc += d * d;
        }
      exp_error = acc / ((double)N * p->n_tgt);
    } else {
      /* SimState::negativeLogLikelihood(other) (simstate.cpp:135-158):
       * loops i (x) outer, j (y), k inner. */
      double dist = 0;
      for (int i = 0; i < W; ++i)
        for (int j = 0; j < H; ++j)
          for (int k = 0; k < p->n_tgt; ++k) {
            double a = state[(size_t)(k + p->n_in) * N + (size_t)j * W + i];
            double b = tgt[(size_t)k * N + (size_t)j * W + i];
            double abs_sub = fabs(a - b) - p->dist_error_threshold;
            if (abs_sub > 0) dist += log(1 + abs_sub);
          }
      dist /= W * H;
      exp_error = dist;
      if (change > p->min_conc_change) exp_error += (change - p->min_conc_change);
    }
    error += exp_error / n_exp;
    ++e;
  }

  if (final_fields) memcpy(final_fields, state, sizeof(double) * (size_t)G * N);

  out->raw_error = error;
  out->change = change;
  out->n_products = G;
  out->steps = steps;
  out->experiments = e;
  if (error > max_error)
    out->fitness = 100;
  else if (p->error_precision > 0)
    out->fitness = gro_ceil_s(error, p->error_precision);
  else
    out->fitness = error;

  free(state);
  free(snap);
  free(pg.ops);
  return 0;
}

int gro_evaluate(const gro_params *p, int n_exp, const double *init,
                 const double *target, int n_genes, const gro_gene *genes,
                 int n_links, const gro_link *links, double max_error,
                 gro_result *out, double *final_fields) {
  return evaluate_dt(p, p->dt, n_exp, init, target, n_genes, genes, n_links,
                     links, max_error, out, final_fields);
}

/* ---- experiment images (simstate.cpp:71-90) ----------------------------- */

int gro_decode_rgb(int width, int height, const unsigned int *rgb, int n_planes,
                   double *out) {
  if (width < 1 || height < 1 || n_planes < 0 || n_planes > 3) return -1;
  /* SimState::loadMorphologyImage: scan line j of the image is row
   * height - 1 - j of the planes (Qt's y axis grows down), and plane k takes
   * byte 2 - k of the QRgb (k = 0 red, 1 green, 2 blue) divided by 255. */
  for (int j = 0; j < height; ++j) {
    const unsigned int *line = rgb + (size_t)j * width;
    for (int i = 0; i < width; ++i) {
      unsigned int color = line[i];
      for (int k = 0; k < n_planes; ++k)
        out[((size_t)k * height + (size_t)(height - j - 1)) * width + i] =
            ((double)((color >> (16 - (8 * k))) & 0xff)) / 255.0;
    }
  }
  return 0;
}

/* ---- model-parallel batch ---------------------------------------------- */

typedef struct {
  const gro_params *p;
  int n_exp;
  const double *init, *target;
  int n_models;
  const int *gene_off, *link_off;
  const gro_gene *genes;
  const gro_link *links;
  const double *max_error, *dt;
  gro_result *out;
  int *rc;
  int next;
  pthread_mutex_t mu;
} batch_t;

static void *batch_worker(void *arg) {
  batch_t *b = (batch_t *)arg;
  for (;;) {
    pthread_mutex_lock(&b->mu);
    int m = b->next++;
    pthread_mutex_unlock(&b->mu);
    if (m >= b->n_models) break;
    int g0 = b->gene_off[m], l0 = b->link_off[m];
    double dt = b->dt ? b->dt[m] : b->p->dt;
    int rc = evaluate_dt(b->p, dt, b->n_exp, b->init, b->target,
                         b->gene_off[m + 1] - g0, b->genes + g0,
                         b->link_off[m + 1] - l0, b->links + l0,
                         b->max_error ? b->max_error[m] : 1e300, &b->out[m],
                         NULL);
    if (rc) b->rc[m] = rc;
  }
  return NULL;
}

int gro_evaluate_batch_dt(const gro_params *p, int n_exp, const double *init,
                          const double *target, int n_models,
                          const int *gene_off, const gro_gene *genes,
                          const int *link_off, const gro_link *links,
                          const double *max_error, const double *dt,
                          int n_threads, gro_result *out) {
  batch_t b = {p, n_exp, init, target, n_models, gene_off, link_off, genes,
               links, max_error, dt, out, NULL, 0, PTHREAD_MUTEX_INITIALIZER};
  b.rc = (int *)calloc((size_t)(n_models > 0 ? n_models : 1), sizeof(int));
  if (!b.rc) return -3;
  if (n_threads < 1) n_threads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)n_threads);
  int started = 0;
  for (int t = 0; t < n_threads; ++t)
    if (pthread_create(&th[t], NULL, batch_worker, &b) == 0) started++;
  if (started == 0) batch_worker(&b);
  for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
  free(th);
  int rc = 0;
  for (int m = 0; m < n_models; ++m)
    if (b.rc[m]) rc = b.rc[m];
  free(b.rc);
  return rc;
}

int gro_evaluate_batch(const gro_params *p, int n_exp, const double *init,
                       const double *target, int n_models,
                       const int *gene_off, const gro_gene *genes,
                       const int *link_off, const gro_link *links,
                       const double *max_error, int n_threads,
                       gro_result *out) {
  return gro_evaluate_batch_dt(p, n_exp, init, target, n_models, gene_off,
                               genes, link_off, links, max_error, NULL,
                               n_threads, out);
}
