/* oracle_check.c -- the FP64 oracle (oracle/grn_oracle.c) under sanitizers:
 * single-model evaluations on small grids (both fitness kinds, several
 * experiments, a maxError cut, fields written back), the model-parallel
 * batch on several threads, the in-use closure with a too-small output, and
 * the image decode.  Test infrastructure, CPU only; built by the Makefile
 * next to it and run by tests/test_native_sanitizers.py. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "grn_oracle.h"

static int failures = 0;
#define CHECK(c, what)                          \
  do {                                          \
    if (!(c)) {                                 \
      fprintf(stderr, "FAILED: %s\n", what);    \
      ++failures;                               \
    }                                           \
  } while (0)

int main(void) {
  enum { W = 7, H = 5, N = W * H, NEXP = 3 };
  gro_gene genes[4] = {{0, 0, 0.3, 0.2, 0.05}, {1, 0, 0.8, 0.4, 0.02}, {2, 0, 0.5, 0.1, 0.0}, {9, 0, 0.1, 0.1, 0.1}};
  gro_link links[5] = {{0, 1, 0.4, 2.0, 0}, {2, 1, 0.3, -1.5, 1}, {1, 2, 0.6, 3.0, 1}, {0, 2, 0.2, 1.2, 0},
                       {9, 9, 0.5, 1.0, 0}};
  double init[NEXP * 1 * N], target[NEXP * 1 * N], fields[64 * N];
  for (int i = 0; i < NEXP * N; ++i) {
    init[i] = 0.5 + 0.4 * sin(0.3 * i);
    target[i] = 0.5 + 0.4 * cos(0.2 * i);
  }
  for (int kind = 0; kind < 2; ++kind) {
    gro_params p = {W, H, 1, 1, 200, 0.05, 1e-3, kind == 0 ? 1e-4 : -1.0, 0.02, 6, kind};
    gro_result r;
    CHECK(gro_evaluate(&p, NEXP, init, target, 4, genes, 5, links, 1e300, &r, fields) == 0, "evaluate");
    CHECK(r.n_products == 3 && r.experiments == NEXP && isfinite(r.raw_error), "result");
    gro_result cut;
    CHECK(gro_evaluate(&p, NEXP, init, target, 4, genes, 5, links, 0.5 * r.raw_error, &cut, NULL) == 0, "cut");
    CHECK(cut.experiments < NEXP || cut.fitness == 100, "maxError cut");
  }
  /* batch: 24 copies with varied time steps on 6 threads */
  enum { M = 24 };
  int goff[M + 1], loff[M + 1];
  gro_gene g[M * 4];
  gro_link l[M * 5];
  double dt[M];
  for (int m = 0; m < M; ++m) {
    goff[m] = 4 * m;
    loff[m] = 5 * m;
    for (int i = 0; i < 4; ++i) g[4 * m + i] = genes[i];
    for (int i = 0; i < 5; ++i) l[5 * m + i] = links[i];
    dt[m] = 0.01 + 0.002 * m;
  }
  goff[M] = 4 * M;
  loff[M] = 5 * M;
  gro_params p = {W, H, 1, 1, 100, 0.05, 0.0, -1.0, 0.0, 0, 1};
  gro_result res[M];
  CHECK(gro_evaluate_batch_dt(&p, NEXP, init, target, M, goff, g, loff, l, NULL, dt, 6, res) == 0, "batch");
  for (int m = 1; m < M; ++m) CHECK(res[m].raw_error != res[0].raw_error, "time steps differ");
  int labels[2];
  CHECK(gro_labels_in_use(4, genes, 5, links, 2, labels, 2) < 0, "too small an output is reported");
  unsigned int img[N];
  for (int i = 0; i < N; ++i) img[i] = 0xFF000000u | (unsigned)(i * 2654435761u >> 8);
  double planes[3 * N];
  CHECK(gro_decode_rgb(W, H, img, 3, planes) == 0, "decode");
  CHECK(gro_decode_rgb(W, H, img, 4, planes) != 0, "at most three planes");
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("oracle_check: all checks passed\n");
  return 0;
}
