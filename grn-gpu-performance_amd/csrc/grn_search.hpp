// grn_search.hpp -- the evolutionary search that calls the simulation path:
// island populations evolved by deterministic crowding with the reference's
// crossover and mutation operators, plus migration between islands.
//
// Reference behaviour restated (paths relative to the reference tree):
//   random model ............ Src/Model/model.cpp:285-315, modelprod.cpp:14-24,
//                             modellink.cpp:10-22
//   crossover ............... model.cpp:21-185
//   mutation ................ model.cpp:317-333,472-547, modelprod.cpp:43-59,
//                             modellink.cpp:24-40
//   excess removal .......... model.cpp:549-574
//   deterministic crowding .. Src/Search/searchalgodetcrowd.cpp:54-147
//   pairwise migration ...... Src/Search/search.cpp:263-311
//   Pareto front ............ search.cpp:349-377, individual.cpp:139-143
//   generation statistics ... Src/Search/generation.cpp:42-72
// The reference draws from a per-translation-unit std::mt19937_64 seeded by
// std::random_device (mathalgo.h:222-230), so its runs cannot be replayed;
// here every island draws from its own seeded xoshiro256** stream.
#pragma once

#include <cstdint>
#include <limits>
#include <string>
#include <vector>

#include "grn.h"
#include "grn_model.hpp"

namespace grn {

// Sets the calling thread's grn_last_error() text and returns code (defined
// with the rest of the C ABI in grn_host.cpp).
int set_last_error(int code, const char *fmt, ...);

// Spearman rank correlation of two equally long samples (average ranks for
// ties); 0 when undefined.
double spearman(const std::vector<double> &a, const std::vector<double> &b);

// xoshiro256** (Blackman & Vigna) seeded through splitmix64.
class Rng {
 public:
  explicit Rng(uint64_t seed = 1);
  uint64_t next();
  double uniform01();                   // [0, 1)
  double uniform(double lo, double hi);  // [lo, hi)
  int below(int n);                     // [0, n), n > 0
  bool coin() { return (next() >> 63) != 0; }

 private:
  uint64_t s_[4];
};

struct Individual {
  Model model;
  double error = -1;           // -1 until evaluated
  double parent_error = 1000;  // maxError of its evaluation (individual.cpp:19-25)
  int complexity = 0;          // Model::calcComplexityInUse
  int dataset = 0;
  double dt = 0;               // per-model time step (0: the dataset's)
  uint64_t id = 0;             // (island bred on << 40) | serial
  int generation = -1;         // generation whose evaluation gave its error
  // ln of the simulated raw error where the caller recorded it (NaN if not):
  // the scale a surrogate predicts on, for comparing a child with its parent
  double log_raw = std::numeric_limits<double>::quiet_NaN();
  double parent_log_raw = std::numeric_limits<double>::quiet_NaN();  // its slot parent's
};

// Writes one model's feature row (row_width values) for a surrogate.
typedef void (*RowEncoder)(const Model &m, int n_morphogens, uint16_t *row);

class Search {
 public:
  Search(const grn_search_params &p, const int *island_datasets);

  // Models that need an error now (children or the initial population).
  int n_pending() const { return (int)pending_.size(); }
  const std::vector<Individual *> &pending() const { return pending_; }
  // Receives the errors of pending(), in order, then selects, migrates and
  // breeds: with oversample 1 the children go straight to pending(), with k > 1
  // each slot gets k candidates (candidates()) until screen() picks one.
  void advance(const double *errors);

  // Candidates per offspring slot for the next breeding (1 = plain search).
  void set_oversample(int k) { oversample_ = k < 1 ? 1 : k; }
  int oversample() const { return oversample_; }
  // Candidates awaiting screen(): island-major, then repetition, then slot.
  const std::vector<Individual *> &candidates() const { return cand_ptrs_; }
  // Keeps, per slot, the candidate with the lowest score (ties: the first
  // repetition) as the child, and fills pending().  Candidates identical to
  // the slot's parent are passed over unless every candidate is one: a clone
  // ties with its parent and changes nothing.  chosen (optional) receives
  // the kept candidates' indices, in pending order.
  void screen(const double *scores, std::vector<size_t> *chosen = nullptr);
  // Fraction of each island's slots that screen() sends to the simulator
  // (1 = all): the slots whose kept candidate promises the largest gain over
  // its parent (score minus the parent's log_raw; parents without one rank
  // last).  A slot left out keeps its parent, as a child that lost would.
  void set_screen_fraction(double f) { screen_fraction_ = f; }
  double screen_fraction() const { return screen_fraction_; }
  // For the next advance() only: while breeding k > 1 candidates per slot,
  // each island's thread also writes every candidate's feature row to
  // rows + i * row_width (i in candidates() order), if they fit in cap rows.
  void set_candidate_rows(uint16_t *rows, size_t cap, int row_width, RowEncoder enc) {
    rows_ = rows, rows_cap_ = cap, row_width_ = row_width, encoder_ = enc;
  }
  // True when the last advance() wrote those rows.
  bool candidate_rows_written() const { return rows_written_; }
  // What the candidate rows hold (GRN_SCREEN_*): the candidate's features, or
  // its features minus its slot parent's with the parent's ln raw error / 8
  // in the first padding column (a score is then a predicted change of ln
  // error, and screen() ranks slots by the score alone).
  void set_screen_policy(int policy) { policy_ = policy; }
  int screen_policy() const { return policy_; }
  // Writes every candidate's row as breeding would (rows: candidates() order).
  void encode_candidates(uint16_t *rows, int row_width);

  // The parent a pending child competes with (its slot's parent), or null
  // for a member of the initial population.  Valid until advance().
  const Individual *slot_parent(const Individual *child) const;
  // The island whose population, children or candidates hold x (-1: none).
  int island_of(const Individual *x) const;
  // Puts every island member back into pending() the way an initial
  // population is: the children or candidates bred for the next generation
  // are dropped, the Pareto front and the perfect-model end are cleared, and
  // the next advance() records the members' errors without selection.  For a
  // change of fidelity: members scored on a coarse grid scored again on the
  // full one.
  void rescore();

  const grn_search_params &params() const { return p_; }
  int generation() const { return generation_; }
  const std::vector<std::vector<Individual>> &islands() const { return islands_; }
  const std::vector<Individual> &pareto() const { return pareto_; }
  grn_generation_stats stats() const { return stats_; }

  // Operators, public for tests.
  Model create_random(Rng &rng, int n_products);
  void cross(Rng &rng, const Model &p1, const Model &p2, Model *c1, Model *c2);
  // Names for duplicated products: island i hands out 11 + i + n_islands * k
  // for k = first, first + stride, first + 2 stride, ...
  struct LabelStream {
    int island;
    long first;
    int stride;
    int used = 0;
  };
  // island: whose label stream names duplicated products
  void mutate(Rng &rng, Model *m, int island = 0);
  void mutate(Rng &rng, Model *m, LabelStream *labels);
  void remove_excess(Rng &rng, Model *m);

 private:
  Gene random_gene(Rng &rng, int label);
  Link random_link(Rng &rng, int regulator, int regulated);
  void mutate_gene_params(Rng &rng, Gene *g);
  void mutate_link_params(Rng &rng, Link *l);
  // Model::createNewLabel (model.cpp:373-379) hands out 11, 12, ... from one
  // counter; here island i draws 11 + i + n_islands * k, so labels stay
  // unique across islands while islands breed concurrently.  The k of one
  // island come from label_next_[i] on: in order when the island breeds its
  // children (oversample 1); when it breeds k > 1 candidate rounds at once,
  // round r takes every k-th value from label_next_[i] + r, so the rounds
  // never share a label either.
  int new_label(LabelStream *ls) {
    return (int)(11 + ls->island + (long)p_.n_islands * (ls->first + (long)ls->stride * ls->used++));
  }
  void prepare(Individual *ind, int island, uint64_t serial);
  // Shuffles the island's parents; with oversample 1 breeds its children.
  void reproduce(int island);
  // Candidate round r of the island (oversample > 1): its own random stream
  // and label stream; rows (optional) receive the round's feature rows.
  void breed_round(int island, int r, uint16_t *rows);
  void encode_round(int island, int r, uint16_t *rows) const;
  void breed_pair(Rng &rng, const Individual &a, const Individual &b, Individual *k1, Individual *k2,
                  LabelStream *labels, uint64_t serial1, uint64_t serial2);
  void choose_next_generation(int island, std::vector<Individual *> *winners);
  void migrate();
  void update_pareto(const std::vector<Individual *> &fresh);
  void update_stats();

  grn_search_params p_;
  int n_morph_ = 0;
  int generation_ = 0;  // 0 = initial population pending
  std::vector<int> label_next_;     // per island
  std::vector<uint64_t> id_next_;   // per island
  std::vector<Rng> rngs_;  // one stream per island, plus one for migration
  std::vector<std::vector<Individual>> islands_;
  std::vector<std::vector<Individual>> children_;
  std::vector<std::vector<Individual>> cand_;  // oversample x island size per island
  std::vector<Individual *> cand_ptrs_;
  int oversample_ = 1;
  double screen_fraction_ = 1.0;
  int policy_ = GRN_SCREEN_ABSOLUTE;
  uint16_t *rows_ = nullptr;  // candidate feature rows (set_candidate_rows)
  size_t rows_cap_ = 0;
  int row_width_ = 0;
  RowEncoder encoder_ = nullptr;
  bool rows_written_ = false;
  std::vector<std::vector<int>> perm_;  // shuffled parent order per island
  std::vector<std::vector<uint64_t>> round_seed_;  // per island and candidate round
  std::vector<std::vector<int>> round_labels_;     // labels used per island and round
  std::vector<Individual *> pending_;
  std::vector<Individual> pareto_;
  grn_generation_stats stats_{};
  // Search::calcParalEvolution's end of the run (search.cpp:128-167)
  int last_generation_ = -1;   // -1: no limit
  int best_perfect_ = -1;      // complexity of the least complex error-0 model on the front
  bool rescoring_ = false;     // pending() holds the members (rescore())
};

}  // namespace grn
