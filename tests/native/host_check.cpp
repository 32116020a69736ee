// host_check.cpp -- the library's host-side C++ (model text format, model
// compiler and program validation, the island search and its C ABI) built
// with sanitizers and driven on the CPU.  Test infrastructure: the Makefile
// next to this file builds it twice (AddressSanitizer + UndefinedBehavior-
// Sanitizer, and ThreadSanitizer); tests/test_native_sanitizers.py runs both.
//
// The search's simulator is replaced by a deterministic CPU fitness of the
// model's parameters (grn_evaluate_batch below), so the search's own code --
// breeding on up to 16 host threads, selection, migration, the Pareto front --
// runs exactly as in the library.  The surrogate's entry points are stubs
// that this driver never reaches.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "grn.h"
#include "grn_internal.hpp"
#include "grn_model.hpp"
#include "grn_pool.hpp"
#include "grn_program.h"
#include "grn_search.hpp"
#include "grn_surrogate.hpp"

namespace {
thread_local std::string g_err;
int g_failures = 0;

void check(bool ok, const char *what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++g_failures;
  }
}
}  // namespace

// ---- stand-ins for the parts of the library that need a GPU -------------------

namespace grn {
int set_last_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
void encode_model(const Model &, int, float *out) { std::memset(out, 0, sizeof(float) * 468); }
uint16_t float_to_bf16(float) { return 0; }
void encode_model_bf16(const Model &, int, uint16_t *row) { std::memset(row, 0, 2 * 512); }
int Surrogate::train_async(int) { return GRN_E_INVALID; }
int Surrogate::predict_rows(int, const uint16_t *, double *) { return GRN_E_INVALID; }
uint16_t *Surrogate::staging_rows(size_t) { return nullptr; }
int Surrogate::add_rows(int, const uint16_t *, const double *) { return GRN_E_INVALID; }
int Surrogate::add_pairs(int, const uint16_t *, const uint16_t *, const double *, const double *) {
  return GRN_E_INVALID;
}
void Surrogate::encode_rows(const std::vector<const Model *> &, uint16_t *) const {}
}  // namespace grn

// A deterministic error of every parameter of the model, from the CPU.
static double fake_error(const grn_model &m, int dataset) {
  double v = 0;
  for (int i = 0; i < m.n_genes; ++i) v += m.genes[i].lim * (1 + m.genes[i].label % 7);
  v = std::fabs(v - 1.7) + 0.01 * m.n_links;
  double w = 0;
  for (int i = 0; i < m.n_links; ++i) w += m.links[i].dis_const * m.links[i].hill_coef;
  v += 0.1 * std::fabs(w - 0.3);
  return 1.0 + std::fmod(v, 1.0) * 0.999 + 1000.0 * dataset;
}

extern "C" int grn_evaluate_batch(void *, int n, const int *ds, const grn_model *models, const double *max_errors,
                                  const double *, int, grn_result *out) {
  for (int i = 0; i < n; ++i) {
    const double e = fake_error(models[i], ds ? ds[i] : 0);
    out[i] = grn_result{};
    out[i].raw_error = e;
    out[i].fitness = (max_errors && e > max_errors[i]) ? 100.0 : e;
  }
  return 0;
}

namespace grn {
int evaluate_batch_hooked(void *ctx, int n, const int *ds, const grn_model *models, const double *max_errors,
                          const double *dt, int path, grn_result *out, BatchHook hook, void *user) {
  grn_evaluate_batch(ctx, n, ds, models, max_errors, dt, path, out);
  std::string err;
  return hook ? hook(user, out, n, &err) : 0;
}
}  // namespace grn

// ---- checks ------------------------------------------------------------------

static void text_format(const std::string &kat) {
  grn::Model m;
  std::string err;
  check(grn::Model::parse(kat, &m, &err), "the known-answer model parses");
  check(m.genes.size() == 13 && m.links.size() == 19, "13 products and 19 links");
  grn::Model again;
  check(grn::Model::parse(m.to_string(17), &again, &err), "its text round-trips");
  check(again.to_string(17) == m.to_string(17), "identical after the round trip");
  const char *bad[] = {"", "(", "(0 0 0 0 0", "(0 0 0 0 0*1 2 3", "(x)", "(0 0 0 0 0|)"};
  for (const char *b : bad) check(!grn::Model::parse(b, &again, &err), "malformed text is rejected");
}

static void compile_and_validate() {
  grn::Rng rng(7);
  grn_search_params p;
  grn_search_default_params(&p);
  p.n_morphogens = 3;
  grn::Search s(p, nullptr);
  for (int k = 0; k < 200; ++k) {
    grn::Model m = s.create_random(rng, 3 + k % 10);
    GrnProgram prog;
    grn::ProgramInfo info;
    std::string err;
    if (grn::compile_model(m, 3, 0, 0.01f, &prog, &info, &err)) continue;
    check(grn::validate_program(prog, 1, &err) == 0, "a compiled program validates");
    GrnProgram bad = prog;
    bad.dataset = 1;
    check(grn::validate_program(bad, 1, &err) != 0, "dataset index out of range is caught");
    bad = prog;
    bad.G = GRN_PMAX_GENES + 1;
    check(grn::validate_program(bad, 1, &err) != 0, "gene count beyond capacity is caught");
    if (prog.n_links > 0) {
      bad = prog;
      bad.link[0].r = prog.G;
      check(grn::validate_program(bad, 1, &err) != 0, "regulator index beyond the genes is caught");
      bad = prog;
      bad.gene[0].link_begin = prog.n_links;
      bad.gene[0].n_or = 1;
      check(grn::validate_program(bad, 1, &err) != 0, "link range beyond the links is caught");
    }
    bad = prog;
    bad.dif_mask ^= 1;
    check(grn::validate_program(bad, 1, &err) != 0, "inconsistent masks are caught");
  }
}

struct Snapshot {
  std::vector<std::string> pending;
  grn_generation_stats st;
};

static std::vector<Snapshot> run_search(int oversample, int generations) {
  grn_search_params p;
  grn_search_default_params(&p);
  p.n_islands = 6;
  p.island_size = 40;
  p.n_morphogens = 3;
  p.migration = GRN_MIGRATION_RING;
  p.migration_period = 3;
  p.seed = 99;
  void *s = nullptr;
  std::vector<Snapshot> out;
  if (grn_search_create(&p, nullptr, &s)) {
    check(false, "search created");
    return out;
  }
  check(grn_search_set_oversample(s, oversample) == 0, "oversample accepted");
  int dummy_ctx = 0;
  for (int g = 0; g < generations; ++g) {
    if (oversample > 1 && g > 0) {
      int n = 0;
      const grn_model *cand = nullptr;
      check(grn_search_candidates(s, &n, &cand) == 0, "candidates listed");
      std::vector<double> scores(n);
      for (int i = 0; i < n; ++i) scores[i] = fake_error(cand[i], 0);
      check(grn_search_screen(s, scores.data()) == 0, "candidates screened");
    }
    Snapshot snap;
    int n = 0;
    const grn_model *models = nullptr;
    check(grn_search_pending(s, &n, &models, nullptr, nullptr, nullptr) == 0, "pending listed");
    for (int i = 0; i < n; ++i) {
      grn::Model m;
      grn::model_from_abi(&models[i], &m);
      snap.pending.push_back(m.to_string(17));
    }
    check(grn_search_run_generation(s, &dummy_ctx, GRN_PATH_AUTO, &snap.st) == 0, "generation ran");
    out.push_back(std::move(snap));
  }
  std::vector<grn_pareto_entry> front(256);
  int nf = 0;
  check(grn_search_pareto(s, front.data(), (int)front.size(), &nf) == 0 && nf >= 1, "Pareto front listed");
  std::vector<grn_gene> genes(256);
  std::vector<grn_link> links(1024);
  int ng = 0, nl = 0;
  check(grn_search_pareto_model(s, 0, genes.data(), 256, &ng, links.data(), 1024, &nl) == 0, "front model read");
  double err = 0;
  check(grn_search_best(s, genes.data(), 256, &ng, links.data(), 1024, &nl, &err) == 0, "best model read");
  grn_search_destroy(s);
  return out;
}

static void searches_replay(int oversample) {
  // breeding runs on up to 16 host threads; two runs must agree exactly
  const std::vector<Snapshot> a = run_search(oversample, 12), b = run_search(oversample, 12);
  check(a.size() == b.size(), "same number of generations");
  for (size_t g = 0; g < a.size() && g < b.size(); ++g) {
    check(a[g].pending == b[g].pending, "identical pending models");
    check(std::memcmp(&a[g].st, &b[g].st, sizeof(grn_generation_stats)) == 0, "identical statistics");
  }
}

// The worker pool: every index runs exactly once, also when loops start
// from several threads at once (the later ones run on their own thread) and
// from inside a loop (nested loops run inline).
static void pool_loops() {
  constexpr int kN = 1000;
  auto one_loop = [](std::vector<int> *hits, bool nested) {
    grn::parallel_for(kN, [&](int i) {
      (*hits)[i] += 1;
      if (nested && i % 100 == 0) {
        std::vector<int> inner(50, 0);
        grn::parallel_for(50, [&](int j) { inner[j] += j; });
        for (int j = 0; j < 50; ++j)
          if (inner[j] != j) (*hits)[i] = -1000;
      }
    });
  };
  for (int rep = 0; rep < 20; ++rep) {
    std::vector<std::vector<int>> hits(4, std::vector<int>(kN, 0));
    std::vector<std::thread> callers;
    for (int t = 0; t < 4; ++t) callers.emplace_back(one_loop, &hits[t], t % 2 == 1);
    for (std::thread &th : callers) th.join();
    for (const std::vector<int> &h : hits)
      for (int v : h) check(v == 1, "pool: every index of every loop runs exactly once");
  }
  std::vector<int> none;
  grn::parallel_for(0, [&](int) { none.push_back(1); });
  check(none.empty(), "pool: an empty loop runs nothing");
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: host_check <known-answer model file>\n");
    return 2;
  }
  std::ifstream f(argv[1]);
  std::stringstream ss;
  ss << f.rdbuf();
  std::string kat = ss.str();
  while (!kat.empty() && (kat.back() == '\n' || kat.back() == ' ')) kat.pop_back();
  text_format(kat);
  compile_and_validate();
  pool_loops();
  searches_replay(1);
  searches_replay(4);
  if (g_failures) {
    std::fprintf(stderr, "%d check(s) failed\n", g_failures);
    return 1;
  }
  std::printf("host_check: all checks passed\n");
  return 0;
}
