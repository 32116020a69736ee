// grn_search.cpp -- island search by deterministic crowding (the caller of
// the simulation path); see grn_search.hpp for the reference map.
#include "grn_search.hpp"

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <iterator>
#include <string>
#include <thread>

#include "grn_internal.hpp"
#include "grn_pool.hpp"
#include "grn_surrogate.hpp"

namespace grn {

// Spearman rank correlation (average ranks for ties); 0 for fewer than 2 points.
double spearman(const std::vector<double> &a, const std::vector<double> &b) {
  const size_t n = a.size();
  if (n < 2) return 0;
  auto ranks = [n](const std::vector<double> &v) {
    std::vector<int> o(n);
    std::iota(o.begin(), o.end(), 0);
    std::stable_sort(o.begin(), o.end(), [&](int x, int y) { return v[x] < v[y]; });
    std::vector<double> r(n);
    for (size_t i = 0; i < n;) {
      size_t j = i;
      while (j + 1 < n && v[o[j + 1]] == v[o[i]]) ++j;
      for (size_t k = i; k <= j; ++k) r[o[k]] = 0.5 * (double)(i + j);
      i = j + 1;
    }
    return r;
  };
  const std::vector<double> ra = ranks(a), rb = ranks(b);
  double ma = 0, mb = 0;
  for (size_t i = 0; i < n; ++i) {
    ma += ra[i];
    mb += rb[i];
  }
  ma /= n;
  mb /= n;
  double sab = 0, saa = 0, sbb = 0;
  for (size_t i = 0; i < n; ++i) {
    sab += (ra[i] - ma) * (rb[i] - mb);
    saa += (ra[i] - ma) * (ra[i] - ma);
    sbb += (rb[i] - mb) * (rb[i] - mb);
  }
  return saa > 0 && sbb > 0 ? sab / std::sqrt(saa * sbb) : 0;
}

// ---- RNG ---------------------------------------------------------------------

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

Rng::Rng(uint64_t seed) {
  uint64_t x = seed;
  for (uint64_t &w : s_) w = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

uint64_t Rng::next() {
  const uint64_t r = rotl(s_[1] * 5, 7) * 9;
  const uint64_t t = s_[1] << 17;
  s_[2] ^= s_[0];
  s_[3] ^= s_[1];
  s_[1] ^= s_[2];
  s_[0] ^= s_[3];
  s_[2] ^= t;
  s_[3] = rotl(s_[3], 45);
  return r;
}

double Rng::uniform01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
double Rng::uniform(double lo, double hi) { return lo + (hi - lo) * uniform01(); }
int Rng::below(int n) { return (int)(((unsigned __int128)next() * (unsigned)n) >> 64); }

// ---- operators on models -----------------------------------------------------

// Label sets are small sorted vectors (a model holds tens of labels).
typedef std::vector<int> LabelSet;

static bool has_label(const LabelSet &s, int l) { return std::binary_search(s.begin(), s.end(), l); }

static LabelSet labels_of(const Model &m) {
  LabelSet s;
  s.reserve(m.genes.size());
  for (const Gene &g : m.genes) s.push_back(g.label);
  std::sort(s.begin(), s.end());
  s.erase(std::unique(s.begin(), s.end()), s.end());
  return s;
}

static LabelSet unite(const LabelSet &a, const LabelSet &b) {
  LabelSet u;
  u.reserve(a.size() + b.size());
  std::set_union(a.begin(), a.end(), b.begin(), b.end(), std::back_inserter(u));
  return u;
}

static int find_link(const Model &m, int regulator, int regulated) {
  for (size_t i = 0; i < m.links.size(); ++i)
    if (m.links[i].regulator == regulator && m.links[i].regulated == regulated) return (int)i;
  return -1;
}

// Model::removeLink(regulator, regulated): the first matching link, if any.
static void remove_link(Model *m, int regulator, int regulated) {
  const int i = find_link(*m, regulator, regulated);
  if (i >= 0) m->links.erase(m->links.begin() + i);
}

// Model::removeProduct / removeProductWithLabel: the product and every link
// touching its label.
static void remove_product_at(Model *m, int i) {
  const int label = m->genes[i].label;
  m->genes.erase(m->genes.begin() + i);
  for (int j = (int)m->links.size() - 1; j >= 0; --j)
    if (m->links[j].regulator == label || m->links[j].regulated == label) m->links.erase(m->links.begin() + j);
}

Gene Search::random_gene(Rng &rng, int label) {
  // ModelProd(nMorphogens, label) (modelprod.cpp:14-24); zeroInit tags the
  // red/green/blue morphogens with -1/-2/-3 (modelprod.cpp:61-70)
  Gene g;
  g.label = label;
  g.init = (label < 3 && label < n_morph_) ? -(double)(label + 1) : 0.0;
  g.lim = rng.uniform(p_.lim_lo, p_.lim_hi);
  g.deg = rng.uniform(p_.deg_lo, p_.deg_hi);
  g.dif = rng.uniform01() < p_.dif_prob ? rng.uniform(p_.dif_lo, p_.dif_hi) : 0.0;
  return g;
}

Link Search::random_link(Rng &rng, int regulator, int regulated) {
  // ModelLink(rtor, rted, isActivation = false) (modellink.cpp:10-22)
  Link l;
  l.regulator = regulator;
  l.regulated = regulated;
  l.dis_const = rng.uniform(p_.k_lo, p_.k_hi);
  const double n = rng.uniform(p_.n_lo, p_.n_hi);
  l.hill_coef = rng.coin() ? n : -n;
  l.is_and = l.hill_coef < 0 || rng.coin();
  return l;
}

void Search::mutate_gene_params(Rng &rng, Gene *g) {
  // ModelProd::mutateParams(1, nMorphogens) (modelprod.cpp:43-59)
  if (rng.below(100) < 1) g->lim = rng.uniform(p_.lim_lo, p_.lim_hi);
  if (rng.below(100) < 1) g->deg = rng.uniform(p_.deg_lo, p_.deg_hi);
  if (rng.below(100) < 1) g->dif = rng.coin() ? rng.uniform(p_.dif_lo, p_.dif_hi) : 0.0;
}

void Search::mutate_link_params(Rng &rng, Link *l) {
  // ModelLink::mutateParams(1) (modellink.cpp:24-40)
  if (rng.below(100) < 1) l->dis_const = rng.uniform(p_.k_lo, p_.k_hi);
  if (rng.below(100) < 1) {
    const double n = rng.uniform(p_.n_lo, p_.n_hi);
    l->hill_coef = rng.coin() ? n : -n;
  }
  if (l->hill_coef < 0)
    l->is_and = true;
  else if (rng.below(100) < 1)
    l->is_and = rng.coin();
}

Model Search::create_random(Rng &rng, int n_products) {
  // Model::createRandom (model.cpp:285-315)
  Model m;
  for (int i = 0; i < n_products; ++i) m.genes.push_back(random_gene(rng, i));
  for (int i = 0; i < n_products; ++i) {
    const int n_reg = rng.below(1 + n_products);
    for (int j = 0; j < n_reg; ++j) {
      const int regulated = rng.below(n_products);
      remove_link(&m, i, regulated);  // addOrReplaceRandomLink (model.cpp:449-453)
      m.links.push_back(random_link(rng, i, regulated));
    }
  }
  return m;
}

void Search::remove_excess(Rng &rng, Model *m) {
  // Model::removeExcess(nMinProducts, nMaxProducts, nMaxLinks) (model.cpp:550-574)
  const int n_min = n_morph_;
  if (p_.n_max_products < 0 && p_.n_max_links < 0) return;
  LabelSet in_use = m->labels_in_use(n_min);  // sorted; recomputed only after removals
  if (p_.n_max_products > -1) {
    int to_remove = (int)in_use.size() - p_.n_max_products;
    if (to_remove > 0) {
      in_use.erase(std::remove_if(in_use.begin(), in_use.end(), [&](int l) { return l < n_min; }), in_use.end());
      for (int i = 0; i < to_remove && !in_use.empty(); ++i) {
        const int k = rng.below((int)in_use.size());
        const int label = in_use[k];
        for (size_t g = 0; g < m->genes.size(); ++g)
          if (m->genes[g].label == label) {
            remove_product_at(m, (int)g);
            break;
          }
        in_use.erase(in_use.begin() + k);
      }
      in_use = m->labels_in_use(n_min);
    }
  }
  if (p_.n_max_links > -1) {
    // calcLinksInUse: links whose regulator and regulated are both in use
    std::vector<int> links_in_use;
    for (size_t i = 0; i < m->links.size(); ++i)
      if (has_label(in_use, m->links[i].regulator) && has_label(in_use, m->links[i].regulated))
        links_in_use.push_back((int)i);
    int to_remove = (int)links_in_use.size() - p_.n_max_links;
    if (to_remove <= 0) return;
    std::vector<bool> drop(m->links.size(), false);
    for (int i = 0; i < to_remove; ++i) {
      const int k = rng.below((int)links_in_use.size());
      drop[links_in_use[k]] = true;
      links_in_use.erase(links_in_use.begin() + k);
    }
    std::vector<Link> kept;
    for (size_t i = 0; i < m->links.size(); ++i)
      if (!drop[i]) kept.push_back(m->links[i]);
    m->links.swap(kept);
  }
}

void Search::cross(Rng &rng, const Model &p1, const Model &p2, Model *c1, Model *c2) {
  // Model::cross (model.cpp:21-85): product-based uniform crossover
  const LabelSet prods1 = labels_of(p1), prods2 = labels_of(p2);
  LabelSet common, only1, only2;
  for (int l : prods1) (has_label(prods2, l) ? common : only1).push_back(l);
  for (int l : prods2)
    if (!has_label(prods1, l)) only2.push_back(l);
  // distributeProducts: each label goes to the second set with probability 1/2
  // (labels visited in ascending order, so every set stays sorted)
  auto distribute = [&](const LabelSet &from, LabelSet *to1, LabelSet *to2) {
    for (int l : from) (rng.below(100) < 50 ? to2 : to1)->push_back(l);
  };
  LabelSet copied, swapped, ex1to1, ex1to2, ex2to2, ex2to1;
  distribute(common, &copied, &swapped);
  distribute(only1, &ex1to1, &ex1to2);
  distribute(only2, &ex2to2, &ex2to1);
  const int n_morph = n_morph_;
  // copyProductsAndLinks (model.cpp:101-184)
  auto build = [&](const Model &from1, const LabelSet &products1, const Model &from2, const LabelSet &products2) {
    Model to;
    to.genes.reserve(products1.size() + products2.size() + 2);
    to.links.reserve(from1.links.size() + from2.links.size() + 2);
    for (const Gene &g : from1.genes)
      if (has_label(products1, g.label)) to.genes.push_back(g);
    for (const Gene &g : from2.genes)
      if (has_label(products2, g.label)) to.genes.push_back(g);
    // candidate substitute regulators: the other side's non-morphogen labels
    std::vector<int> regs1, regs2;
    for (int l : products1)
      if (l >= n_morph) regs1.push_back(l);
    for (int l : products2)
      if (l >= n_morph) regs2.push_back(l);
    auto copy_side = [&](const Model &from, const LabelSet &mine, const LabelSet &other,
                         const std::vector<int> &subst) {
      for (const Link &l : from.links) {
        if (!has_label(mine, l.regulated)) continue;
        if (has_label(mine, l.regulator) || has_label(other, l.regulator)) {
          to.links.push_back(l);
        } else if (!subst.empty()) {
          const int r = subst[rng.below((int)subst.size())];
          if (find_link(from, r, l.regulated) < 0) {
            Link nl = l;
            nl.regulator = r;
            to.links.push_back(nl);
          }
        }
      }
    };
    copy_side(from1, products1, products2, regs2);
    copy_side(from2, products2, products1, regs1);
    return to;
  };
  *c1 = build(p1, unite(copied, ex1to1), p2, unite(swapped, ex2to1));
  *c2 = build(p2, unite(copied, ex2to2), p1, unite(swapped, ex1to2));
  if (copied.size() + ex1to1.size() < swapped.size() + ex2to1.size()) std::swap(*c1, *c2);
  remove_excess(rng, c1);
  remove_excess(rng, c2);
}

void Search::mutate(Rng &rng, Model *m, int island) {
  LabelStream ls{island, label_next_[island], 1};
  mutate(rng, m, &ls);
  label_next_[island] += ls.used;
}

void Search::mutate(Rng &rng, Model *m, LabelStream *labels) {
  // Model::mutate (model.cpp:499-547)
  const int n_min = n_morph_;
  for (int i = (int)m->genes.size() - 1; i >= 0; --i) {
    const int label = m->genes[i].label;
    if (rng.below(100) < 1) {
      // duplicateProduct (model.cpp:317-333)
      Gene g = m->genes[i];
      g.init = 0;
      g.label = new_label(labels);
      m->genes.push_back(g);
      const int regulated = m->genes[rng.below((int)m->genes.size())].label;
      remove_link(m, g.label, regulated);
      m->links.push_back(random_link(rng, g.label, regulated));
      std::vector<int> labels;
      for (const Gene &x : m->genes) labels.push_back(x.label);
      std::sort(labels.begin(), labels.end());
      const int span = (int)labels.size() - n_min;
      if (span > 0) {
        const int regulator = labels[n_min + rng.below(span)];
        remove_link(m, regulator, g.label);
        m->links.push_back(random_link(rng, regulator, g.label));
      }
    }
    if (label >= n_min && rng.below(1000) < 15) remove_product_at(m, i);
  }
  for (int i = (int)m->links.size() - 1; i >= 0; --i) {
    if (rng.below(100) < 1 && !m->genes.empty()) {
      // duplicateLink (model.cpp:472-483)
      Link nl = m->links[i];
      const int regulator = m->genes[rng.below((int)m->genes.size())].label;
      const int regulated = m->genes[rng.below((int)m->genes.size())].label;
      remove_link(m, regulator, regulated);
      nl.regulator = regulator;
      nl.regulated = regulated;
      m->links.push_back(nl);
    }
    if (rng.below(1000) < 15 && i < (int)m->links.size()) m->links.erase(m->links.begin() + i);
  }
  remove_excess(rng, m);
  for (Gene &g : m->genes) mutate_gene_params(rng, &g);
  for (Link &l : m->links) mutate_link_params(rng, &l);
}

// ---- the search ---------------------------------------------------------------

// splitmix64's finaliser: a well-mixed 64-bit hash of x.
static uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

Search::Search(const grn_search_params &p, const int *island_datasets) : p_(p) {
  last_generation_ = p.n_generations > 0 ? p.n_generations : -1;
  if (p_.island_size % 2) p_.island_size++;  // searchalgodetcrowd.cpp:33-35
  n_morph_ = p_.n_morphogens;
  for (int i = 0; i <= p_.n_islands; ++i) rngs_.emplace_back(p_.seed * 0x100000001B3ULL + 0x9E37u * (uint64_t)(i + 1));
  islands_.resize(p_.n_islands);
  children_.resize(p_.n_islands);
  cand_.resize(p_.n_islands);
  label_next_.assign(p_.n_islands, 0);
  id_next_.assign(p_.n_islands, 0);
  perm_.resize(p_.n_islands);
  round_seed_.resize(p_.n_islands);
  round_labels_.resize(p_.n_islands);
  for (int is = 0; is < p_.n_islands; ++is) {
    Rng &rng = rngs_[is];
    perm_[is].resize(p_.island_size);
    std::iota(perm_[is].begin(), perm_[is].end(), 0);
    for (int k = 0; k < p_.island_size; ++k) {
      // SearchAlgoDetCrowd::calcInitialPopulation (searchalgodetcrowd.cpp:54-61)
      const int lo = std::max(p_.init_products_min, n_morph_);
      const int hi = std::max(p_.init_products_max, lo);
      Individual ind;
      ind.model = create_random(rng, lo + rng.below(hi - lo + 1));
      ind.dataset = island_datasets ? island_datasets[is] : 0;
      prepare(&ind, is, ++id_next_[is]);
      ind.complexity = ind.model.complexity_in_use(n_morph_);
      islands_[is].push_back(std::move(ind));
    }
  }
  for (auto &isl : islands_)
    for (Individual &ind : isl) pending_.push_back(&ind);
}

void Search::prepare(Individual *ind, int island, uint64_t serial) {
  ind->id = ((uint64_t)island << 40) | serial;
  ind->error = -1;
  // the in-use complexity is computed once an individual is pending
  // (a screened-out candidate never needs it)
  ind->complexity = -1;
  if (p_.dt_rule == 1) {
    double dmax = 0;
    for (const Gene &g : ind->model.genes) dmax = std::max(dmax, g.dif);
    ind->dt = dmax > 0 ? std::min(p_.dt_scale / dmax, p_.dt_cap) : p_.dt_cap;
  } else {
    ind->dt = 0;
  }
}

// A copy with room for the genes and links a mutation may add.
static void copy_model(const Model &from, Model *to) {
  to->genes.reserve(from.genes.size() + 2);
  to->links.reserve(from.links.size() + 4);
  to->genes.assign(from.genes.begin(), from.genes.end());
  to->links.assign(from.links.begin(), from.links.end());
}

void Search::breed_pair(Rng &rng, const Individual &a, const Individual &b, Individual *k1, Individual *k2,
                        LabelStream *labels, uint64_t serial1, uint64_t serial2) {
  // one iteration of SearchAlgoDetCrowd::reproduce (searchalgodetcrowd.cpp:70-97)
  if (rng.below(100) < (int)std::lround(p_.crossover_rate * 100)) {
    cross(rng, a.model, b.model, &k1->model, &k2->model);
  } else {
    copy_model(a.model, &k1->model);
    copy_model(b.model, &k2->model);
  }
  mutate(rng, &k1->model, labels);
  mutate(rng, &k2->model, labels);
  k1->parent_error = a.error;
  k2->parent_error = b.error;
  k1->parent_log_raw = a.log_raw;
  k2->parent_log_raw = b.log_raw;
  k1->dataset = a.dataset;
  k2->dataset = b.dataset;
  prepare(k1, labels->island, serial1);
  prepare(k2, labels->island, serial2);
}

void Search::reproduce(int is) {
  // SearchAlgoDetCrowd::reproduce (searchalgodetcrowd.cpp:63-101): shuffle
  // the parents, breed consecutive pairs; child i belongs to parent perm[i]
  Rng &rng = rngs_[is];
  std::vector<int> &perm = perm_[is];
  for (int i = (int)perm.size() - 1; i > 0; --i) std::swap(perm[i], perm[rng.below(i + 1)]);  // MathAlgo::shuffle
  const std::vector<Individual> &pop = islands_[is];
  const int n = (int)pop.size();
  std::vector<Individual> &kids = children_[is];
  kids.assign(n, Individual());
  if (oversample_ <= 1) {
    LabelStream ls{is, label_next_[is], 1};
    for (int i = 0; i + 1 < n; i += 2) {
      const uint64_t s1 = ++id_next_[is], s2 = ++id_next_[is];
      breed_pair(rng, pop[perm[i]], pop[perm[i + 1]], &kids[i], &kids[i + 1], &ls, s1, s2);
    }
    label_next_[is] += ls.used;
    for (Individual &k : kids) k.complexity = k.model.complexity_in_use(n_morph_);
    return;
  }
  // k candidate rounds, all with the same parent pairs; round r draws from a
  // stream seeded by the island's r-th draw here, so the rounds can breed
  // concurrently (breed_round) with results independent of the schedule
  round_seed_[is].resize(oversample_);
  for (int r = 0; r < oversample_; ++r) round_seed_[is][r] = rng.next();
  round_labels_[is].assign(oversample_, 0);
  cand_[is].assign((size_t)oversample_ * n, Individual());
}

void Search::breed_round(int is, int r, uint16_t *rows) {
  Rng rng(round_seed_[is][r]);
  const std::vector<int> &perm = perm_[is];
  const std::vector<Individual> &pop = islands_[is];
  const int n = (int)pop.size();
  std::vector<Individual> &cand = cand_[is];
  LabelStream ls{is, label_next_[is] + r, oversample_};
  const uint64_t base = id_next_[is] + (uint64_t)r * n;
  for (int i = 0; i + 1 < n; i += 2)
    breed_pair(rng, pop[perm[i]], pop[perm[i + 1]], &cand[(size_t)r * n + i], &cand[(size_t)r * n + i + 1], &ls,
               base + i + 1, base + i + 2);
  round_labels_[is][r] = ls.used;
  if (rows) encode_round(is, r, rows);
}

void Search::encode_round(int is, int r, uint16_t *rows) const {
  const std::vector<Individual> &cand = cand_[is];
  const int n = (int)islands_[is].size();
  if (policy_ != GRN_SCREEN_DELTA) {
    for (size_t c = (size_t)r * n; c < (size_t)(r + 1) * n; ++c)
      encoder_(cand[c].model, n_morph_, rows + c * (size_t)row_width_);
    return;
  }
  float fp[MLP_WIDTH], fc[MLP_WIDTH];
  for (int i = 0; i < n; ++i) {
    const Individual &parent = islands_[is][perm_[is][i]];
    encode_model(parent.model, n_morph_, fp);
    const size_t c = (size_t)r * n + i;
    encode_model(cand[c].model, n_morph_, fc);
    uint16_t *row = rows + c * (size_t)row_width_;
    for (int k = 0; k < MLP_WIDTH; ++k) row[k] = float_to_bf16(fc[k] - fp[k]);
    const double lr = std::isnan(parent.log_raw) ? 0.0 : parent.log_raw;
    row[MLP_FEATURES] = float_to_bf16((float)(lr / 8.0));
  }
}

void Search::encode_candidates(uint16_t *rows, int row_width) {
  if (cand_ptrs_.empty()) return;
  std::vector<uint16_t *> base(p_.n_islands);
  size_t at = 0;
  for (int is = 0; is < p_.n_islands; ++is) {
    base[is] = rows + at * (size_t)row_width;
    at += cand_[is].size();
  }
  row_width_ = row_width;
  encoder_ = &encode_model_bf16;
  const int K = oversample_;
  parallel_for(p_.n_islands * K, [this, K, &base](int t) { encode_round(t / K, t % K, base[t / K]); });
}

static bool same_model(const Model &a, const Model &b) {
  if (a.genes.size() != b.genes.size() || a.links.size() != b.links.size()) return false;
  for (size_t i = 0; i < a.genes.size(); ++i) {
    const Gene &x = a.genes[i], &y = b.genes[i];
    if (x.label != y.label || x.init != y.init || x.lim != y.lim || x.deg != y.deg || x.dif != y.dif) return false;
  }
  for (size_t i = 0; i < a.links.size(); ++i) {
    const Link &x = a.links[i], &y = b.links[i];
    if (x.regulator != y.regulator || x.regulated != y.regulated || x.dis_const != y.dis_const ||
        x.hill_coef != y.hill_coef || x.is_and != y.is_and)
      return false;
  }
  return true;
}

void Search::screen(const double *scores, std::vector<size_t> *chosen) {
  // candidate offsets per island (island-major order)
  std::vector<size_t> off(p_.n_islands + 1, 0);
  for (int is = 0; is < p_.n_islands; ++is) off[is + 1] = off[is] + cand_[is].size();
  std::vector<std::vector<size_t>> kept(p_.n_islands);  // per island, in slot order
  parallel_for(p_.n_islands, [&](int is) {
    std::vector<Individual> &cand = cand_[is];
    std::vector<Individual> &kids = children_[is];
    const std::vector<Individual> &pop = islands_[is];
    const std::vector<int> &perm = perm_[is];
    const double *sc = scores + off[is];
    const int n = (int)kids.size();
    std::vector<int> pick(n);
    std::vector<double> gain(n);
    std::vector<int> fresh;  // the slot's candidates that are not clones of the parent
    for (int i = 0; i < n; ++i) {
      const Individual &parent = pop[perm[i]];
      int best = -1, best_any = 0;
      fresh.clear();
      for (int r = 0; r < oversample_; ++r) {
        const double v = sc[(size_t)r * n + i];
        if (v < sc[(size_t)best_any * n + i]) best_any = r;
        if (same_model(cand[(size_t)r * n + i].model, parent.model)) continue;
        fresh.push_back(r);
        if (best < 0 || v < sc[(size_t)best * n + i]) best = r;
      }
      pick[i] = best < 0 ? best_any : best;
      if (policy_ == GRN_SCREEN_HALF && fresh.size() > 1) {
        // a uniform pick among the better half of the fresh candidates
        std::stable_sort(fresh.begin(), fresh.end(),
                         [&](int a, int b) { return sc[(size_t)a * n + i] < sc[(size_t)b * n + i]; });
        const uint64_t h = mix64(p_.seed ^ mix64(((uint64_t)generation_ << 32) ^ ((uint64_t)is << 20) ^ (uint64_t)i));
        pick[i] = fresh[h % ((fresh.size() + 1) / 2)];
      }
      const double v = sc[(size_t)pick[i] * n + i];
      const double g = policy_ == GRN_SCREEN_DELTA ? v : v - parent.log_raw;
      gain[i] = std::isnan(g) ? std::numeric_limits<double>::infinity() : g;
    }
    // the slots to simulate, in slot order
    int keep = n;
    if (screen_fraction_ < 1.0) keep = std::max(1, std::min(n, (int)std::ceil(screen_fraction_ * n)));
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    if (keep < n) {
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return gain[a] < gain[b]; });
      order.resize(keep);
      std::sort(order.begin(), order.end());
    }
    size_t next = 0;
    for (int i = 0; i < n; ++i) {
      if (next < order.size() && order[next] == i) {
        ++next;
        const size_t c = (size_t)pick[i] * n + i;
        kept[is].push_back(c);
        kids[i] = std::move(cand[c]);
        kids[i].complexity = kids[i].model.complexity_in_use(n_morph_);
      } else {
        // not simulated: cannot replace its parent
        kids[i].error = std::numeric_limits<double>::infinity();
      }
    }
    cand.clear();
  });
  pending_.clear();
  if (chosen) chosen->clear();
  for (int is = 0; is < p_.n_islands; ++is) {
    const size_t n = children_[is].size();
    for (size_t c : kept[is]) {
      if (chosen) chosen->push_back(off[is] + c);
      pending_.push_back(&children_[is][c % n]);
    }
  }
  cand_ptrs_.clear();
  rows_written_ = false;
}

const Individual *Search::slot_parent(const Individual *child) const {
  for (int is = 0; is < p_.n_islands; ++is) {
    const std::vector<Individual> &kids = children_[is];
    if (kids.empty() || child < kids.data() || child >= kids.data() + kids.size()) continue;
    const size_t i = (size_t)(child - kids.data());
    return i < perm_[is].size() ? &islands_[is][perm_[is][i]] : nullptr;
  }
  return nullptr;
}

int Search::island_of(const Individual *x) const {
  for (int is = 0; is < p_.n_islands; ++is)
    for (const std::vector<Individual> *v : {&islands_[is], &children_[is], &cand_[is]})
      if (!v->empty() && x >= v->data() && x < v->data() + v->size()) return is;
  return -1;
}

void Search::rescore() {
  pending_.clear();
  cand_ptrs_.clear();
  for (std::vector<Individual> &isl : islands_)
    for (Individual &ind : isl) {
      ind.error = -1;
      ind.parent_error = Individual().parent_error;  // no maxError cut, as for the initial population
      ind.generation = -1;
      ind.log_raw = ind.parent_log_raw = std::numeric_limits<double>::quiet_NaN();
      pending_.push_back(&ind);
    }
  pareto_.clear();
  best_perfect_ = -1;
  last_generation_ = p_.n_generations > 0 ? p_.n_generations : -1;
  rescoring_ = true;
}

void Search::choose_next_generation(int is, std::vector<Individual *> *winners) {
  // SearchAlgoDetCrowd::chooseNextGeneration (searchalgodetcrowd.cpp:103-147):
  // each child competes with its own parent and wins ties
  std::vector<Individual> &pop = islands_[is];
  std::vector<Individual> &kids = children_[is];
  const std::vector<int> &perm = perm_[is];
  for (size_t i = 0; i < kids.size(); ++i) {
    Individual &parent = pop[perm[i]];
    if (kids[i].error <= parent.error) {
      parent = std::move(kids[i]);
      winners->push_back(&parent);
    }
  }
  kids.clear();
}

void Search::migrate() {
  Rng &rng = rngs_.back();
  const int n = p_.n_islands;
  if (n < 2 || p_.migration == GRN_MIGRATION_NONE) return;
  if (p_.migration == GRN_MIGRATION_PAIRWISE) {
    // createMigrationPartners + migrateIndividuals (search.cpp:263-311)
    std::vector<int> ids(n);
    std::iota(ids.begin(), ids.end(), 0);
    for (int i = n - 1; i > 0; --i) std::swap(ids[i], ids[rng.below(i + 1)]);
    for (int i = 0; i + 1 < n; i += 2) {
      std::vector<Individual> &g1 = islands_[ids[i]], &g2 = islands_[ids[i + 1]];
      const int ds1 = g1.empty() ? 0 : g1[0].dataset, ds2 = g2.empty() ? 0 : g2[0].dataset;
      std::vector<Individual> all;
      for (Individual &x : g1) all.push_back(std::move(x));
      for (Individual &x : g2) all.push_back(std::move(x));
      std::vector<int> r(all.size());
      std::iota(r.begin(), r.end(), 0);
      for (int k = (int)r.size() - 1; k > 0; --k) std::swap(r[k], r[rng.below(k + 1)]);
      const size_t half = g1.size();
      std::vector<Individual> n1, n2;
      for (size_t k = 0; k < r.size(); ++k) (k < half ? n1 : n2).push_back(std::move(all[r[k]]));
      // an individual that changes island is scored on its new island's
      // data from its next child on; its own error stays as measured
      for (Individual &x : n1) x.dataset = ds1;
      for (Individual &x : n2) x.dataset = ds2;
      g1.swap(n1);
      g2.swap(n2);
    }
    return;
  }
  // ring: the best fraction of island i replaces the worst of island i+1
  const int k = std::max(1, (int)std::lround(p_.migration_fraction * p_.island_size));
  std::vector<std::vector<Individual>> emigrants(n);
  auto better = [](const Individual &a, const Individual &b) {
    return a.error < b.error || (a.error == b.error && a.complexity < b.complexity);
  };
  for (int i = 0; i < n; ++i) {
    std::vector<int> order(islands_[i].size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return better(islands_[i][a], islands_[i][b]); });
    for (int j = 0; j < k && j < (int)order.size(); ++j) emigrants[i].push_back(islands_[i][order[j]]);
  }
  for (int i = 0; i < n; ++i) {
    std::vector<Individual> &dst = islands_[(i + 1) % n];
    std::vector<int> order(dst.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return better(dst[b], dst[a]); });
    const int ds = dst.empty() ? 0 : dst[0].dataset;
    for (int j = 0; j < (int)emigrants[i].size() && j < (int)order.size(); ++j) {
      dst[order[j]] = emigrants[i][j];
      dst[order[j]].dataset = ds;
    }
  }
}

void Search::update_pareto(const std::vector<Individual *> &fresh) {
  // Search::recalculateParetoFront (search.cpp:349-377): sort by (error,
  // complexity), keep strictly decreasing complexity, duplicates included
  for (Individual *x : fresh) pareto_.push_back(*x);
  std::stable_sort(pareto_.begin(), pareto_.end(), [](const Individual &a, const Individual &b) {
    return a.error < b.error || (a.error == b.error && a.complexity < b.complexity);
  });
  std::vector<Individual> front;
  for (Individual &x : pareto_) {
    if (front.empty() || x.complexity < front.back().complexity ||
        (x.complexity == front.back().complexity && x.error == front.back().error))
      front.push_back(std::move(x));
  }
  pareto_.swap(front);
}

void Search::update_stats() {
  // Generation::calcPopulSta (generation.cpp:42-72), over all islands
  grn_generation_stats s{};
  s.generation = generation_;
  s.min_fit = std::numeric_limits<double>::infinity();
  s.max_fit = -std::numeric_limits<double>::infinity();
  s.best_complexity = 0;
  double sum = 0;
  int n = 0;
  for (int is = 0; is < (int)islands_.size(); ++is)
    for (const Individual &x : islands_[is]) {
      sum += x.error;
      ++n;
      if (x.error < s.min_fit || (x.error == s.min_fit && x.complexity < s.best_complexity)) {
        s.min_fit = x.error;
        s.best_complexity = x.complexity;
        s.best_island = is;
      }
      s.max_fit = std::max(s.max_fit, x.error);
    }
  s.mean_fit = n ? sum / n : 0;
  s.pareto_size = (int)pareto_.size();
  s.last_generation = last_generation_;
  s.best_perfect_complexity = best_perfect_;
  // the generation just evaluated is generation_ - 1; the reference goes on
  // while its mean generation is below lastGeneration (search.cpp:158)
  s.done = (last_generation_ >= 0 && generation_ - 1 >= last_generation_) ? 1 : 0;
  stats_ = s;
}

void Search::advance(const double *errors) {
  for (size_t i = 0; i < pending_.size(); ++i) {
    pending_[i]->error = errors[i];
    pending_[i]->generation = generation_;
  }
  const int n_evaluated = (int)pending_.size();
  // the Pareto front sees the initial population and every selected child
  // (Search::addNewIndividual, search.cpp:323-326)
  std::vector<Individual *> winners;
  if (generation_ == 0 || rescoring_)
    winners = pending_;
  else
    for (int is = 0; is < p_.n_islands; ++is) choose_next_generation(is, &winners);
  rescoring_ = false;
  update_pareto(winners);
  // Early end once a perfect model is found (search.cpp:157-167): a new least
  // complex error-0 model at the head of the front moves the last generation
  // to 250 past the current one (never beyond nGenerations).
  if (!pareto_.empty() && pareto_.front().error == 0 &&
      (best_perfect_ < 0 || pareto_.front().complexity < best_perfect_)) {
    best_perfect_ = pareto_.front().complexity;
    const int cut = 250 + generation_;
    last_generation_ = p_.n_generations > 0 ? std::min(cut, p_.n_generations) : cut;
  }
  ++generation_;
  update_stats();
  stats_.n_evaluated = n_evaluated;
  if (p_.migration_period > 0 && generation_ % p_.migration_period == 0) migrate();
  pending_.clear();
  cand_ptrs_.clear();
  // candidate feature rows: island is's block starts at row_off[is]
  std::vector<uint16_t *> rows(p_.n_islands, nullptr);
  rows_written_ = false;
  if (rows_ && encoder_ && oversample_ > 1) {
    size_t total = 0;
    for (int is = 0; is < p_.n_islands; ++is) total += (size_t)oversample_ * islands_[is].size();
    if (total <= rows_cap_) {
      size_t at = 0;
      for (int is = 0; is < p_.n_islands; ++is) {
        rows[is] = rows_ + at * (size_t)row_width_;
        at += (size_t)oversample_ * islands_[is].size();
      }
      rows_written_ = true;
    }
  }
  rows_ = nullptr;  // the buffer is lent for this call only
  // islands breed independently (own random stream, label stream and
  // individuals), so they breed concurrently with identical results; so do
  // the candidate rounds of one island
  parallel_for(p_.n_islands, [this](int is) { reproduce(is); });
  if (oversample_ > 1) {
    const int K = oversample_;
    parallel_for(p_.n_islands * K, [this, K, &rows](int t) { breed_round(t / K, t % K, rows[t / K]); });
    for (int is = 0; is < p_.n_islands; ++is) {
      label_next_[is] += K * *std::max_element(round_labels_[is].begin(), round_labels_[is].end());
      id_next_[is] += (uint64_t)K * islands_[is].size();
    }
  }
  for (int is = 0; is < p_.n_islands; ++is) {
    if (oversample_ <= 1)
      for (Individual &k : children_[is]) pending_.push_back(&k);
    else
      for (Individual &k : cand_[is]) cand_ptrs_.push_back(&k);
  }
}

}  // namespace grn

// ---- C ABI --------------------------------------------------------------------

namespace {

// Flat C views of a list of individuals.
struct FlatModels {
  std::vector<grn_model> models;
  std::vector<grn_gene> genes;
  std::vector<grn_link> links;
  std::vector<int> datasets;
  std::vector<double> max_errors, dts;

  void build(const std::vector<grn::Individual *> &list) {
    size_t ng = 0, nl = 0;
    for (const grn::Individual *x : list) {
      ng += x->model.genes.size();
      nl += x->model.links.size();
    }
    genes.resize(ng);
    links.resize(nl);
    models.resize(list.size());
    datasets.resize(list.size());
    max_errors.resize(list.size());
    dts.resize(list.size());
    size_t gi = 0, li = 0;
    for (size_t i = 0; i < list.size(); ++i) {
      const grn::Individual &x = *list[i];
      grn_model &m = models[i];
      m.n_genes = (int)x.model.genes.size();
      m.genes = genes.data() + gi;
      m.n_links = (int)x.model.links.size();
      m.links = links.data() + li;
      for (const grn::Gene &g : x.model.genes) genes[gi++] = grn_gene{g.label, g.init, g.lim, g.deg, g.dif};
      for (const grn::Link &l : x.model.links)
        links[li++] = grn_link{l.regulator, l.regulated, l.dis_const, l.hill_coef, l.is_and ? 1 : 0};
      datasets[i] = x.dataset;
      max_errors[i] = x.parent_error;
      dts[i] = x.dt;
    }
  }
};

// A search plus the flat arrays grn_search_pending / _candidates hand out.
struct SearchHandle {
  std::unique_ptr<grn::Search> s;
  FlatModels pend, cand;
  std::vector<grn_result> results;
  std::vector<double> errors;

  bool cand_built = false;
  const void *rows_lender = nullptr;  // the surrogate whose staging rows the last breeding wrote
  // the coarse screen's arrays (kept, with their capacity, across generations)
  std::vector<grn_gene> coarse_genes;
  std::vector<grn_model> coarse_models;
  std::vector<int> coarse_ds;
  std::vector<grn_result> coarse_res;
  std::vector<double> coarse_scores;

  // pending() is flattened at once; candidates() only when asked for
  void flatten() {
    pend.build(s->pending());
    cand_built = false;  // (its arrays keep their capacity for the next candidates)
  }
  const FlatModels &candidates() {
    if (!cand_built) cand.build(s->candidates());
    cand_built = true;
    return cand;
  }
};

SearchHandle *as_search(void *p) { return static_cast<SearchHandle *>(p); }

}  // namespace

extern "C" {

void grn_search_default_params(grn_search_params *p) {
  if (!p) return;
  std::memset(p, 0, sizeof *p);
  p->n_islands = 1;
  p->island_size = 64;
  p->n_morphogens = 3;
  p->n_max_products = -1;
  p->n_max_links = -1;
  p->init_products_min = 3;
  p->init_products_max = 3;
  p->migration = GRN_MIGRATION_PAIRWISE;
  p->migration_period = 10;
  p->migration_fraction = 0.1;
  p->crossover_rate = 0.75;
  p->lim_lo = 0, p->lim_hi = 1;
  p->deg_lo = 0, p->deg_hi = 1;
  p->dif_prob = 0.1, p->dif_lo = 0, p->dif_hi = 1;
  p->k_lo = 0, p->k_hi = 1;
  p->n_lo = 1, p->n_hi = 5;
  p->dt_rule = 0;
  p->dt_scale = 0.2;
  p->dt_cap = 1.0;
  p->seed = 1;
}

int grn_search_create(const grn_search_params *p, const int *island_datasets, void **out) {
  if (!p || !out) return grn::set_last_error(GRN_E_INVALID, "null argument");
  *out = nullptr;
  grn_search_params q = *p;
  if (q.n_islands < 1 || q.island_size < 2 || q.n_morphogens < 1)
    return grn::set_last_error(GRN_E_INVALID, "need n_islands >= 1, island_size >= 2, n_morphogens >= 1");
  // -1 (no limit) becomes the capacity of the kernels' program tables
  if (q.n_max_products < 0) q.n_max_products = GRN_PMAX_GENES;
  if (q.n_max_links < 0) q.n_max_links = GRN_PMAX_LINKS;
  if (q.n_max_products < q.n_morphogens || q.n_max_products > GRN_PMAX_GENES || q.n_max_links > GRN_PMAX_LINKS)
    return grn::set_last_error(GRN_E_CAPACITY, "n_max_products must lie in [n_morphogens, %d], n_max_links in [0, %d]",
                               GRN_PMAX_GENES, GRN_PMAX_LINKS);
  if (q.init_products_max > q.n_max_products || q.init_products_min > q.init_products_max)
    return grn::set_last_error(GRN_E_INVALID, "initial product range [%d, %d] outside [n_morphogens, n_max_products]",
                               q.init_products_min, q.init_products_max);
  if (q.migration < GRN_MIGRATION_PAIRWISE || q.migration > GRN_MIGRATION_NONE)
    return grn::set_last_error(GRN_E_INVALID, "unknown migration %d", q.migration);
  if (q.crossover_rate < 0 || q.crossover_rate > 1 || q.dif_prob < 0 || q.dif_prob > 1)
    return grn::set_last_error(GRN_E_INVALID, "probabilities must lie in [0, 1]");
  if (q.dt_rule == 1 && !(q.dt_scale > 0 && q.dt_cap > 0))
    return grn::set_last_error(GRN_E_INVALID, "dt_rule 1 needs dt_scale > 0 and dt_cap > 0");
  std::unique_ptr<SearchHandle> h(new SearchHandle);
  h->s.reset(new grn::Search(q, island_datasets));
  h->flatten();
  *out = h.release();
  return 0;
}

void grn_search_destroy(void *search) { delete as_search(search); }

int grn_search_pending(void *search, int *n, const grn_model **models, const int **datasets,
                       const double **max_errors, const double **dt) {
  SearchHandle *h = as_search(search);
  if (!h || !n) return grn::set_last_error(GRN_E_INVALID, "null argument");
  *n = (int)h->pend.models.size();
  if (models) *models = h->pend.models.data();
  if (datasets) *datasets = h->pend.datasets.data();
  if (max_errors) *max_errors = h->pend.max_errors.data();
  if (dt) *dt = h->pend.dts.data();
  return 0;
}

int grn_search_advance(void *search, const double *errors, grn_generation_stats *stats) {
  SearchHandle *h = as_search(search);
  if (!h || (!errors && !h->pend.models.empty())) return grn::set_last_error(GRN_E_INVALID, "null argument");
  if (h->pend.models.empty() && !h->s->candidates().empty())
    return grn::set_last_error(GRN_E_INVALID, "candidates await grn_search_screen");
  for (size_t i = 0; i < h->pend.models.size(); ++i)
    if (!(errors[i] >= 0) || std::isinf(errors[i]))
      return grn::set_last_error(GRN_E_INVALID, "error %zu is %g; errors must be finite and >= 0", i, errors[i]);
  h->s->advance(errors);
  h->flatten();
  if (stats) *stats = h->s->stats();
  return 0;
}

// The host stage of a generation: selection (deterministic crowding), the
// Pareto front, the statistics, migration when due and breeding of the next
// children, from the simulated errors.  It runs in the batch's stream order
// -- a host node of the captured graph when the context launches graphs --
// so one generation is one graph launch.
static int generation_stage(void *user, const grn_result *res, int n, std::string *err) {
  SearchHandle *h = static_cast<SearchHandle *>(user);
  h->errors.resize(n);
  for (int i = 0; i < n; ++i) {
    const double e = res[i].fitness;
    if (!(e >= 0) || std::isinf(e)) {
      *err = "error " + std::to_string(i) + " is not finite and >= 0";
      return GRN_E_INVALID;
    }
    h->errors[i] = e;
  }
  h->s->advance(h->errors.data());
  h->flatten();
  return 0;
}

int grn_search_run_generation(void *search, void *ctx, int path, grn_generation_stats *stats) {
  SearchHandle *h = as_search(search);
  if (!h || !ctx) return grn::set_last_error(GRN_E_INVALID, "null argument");
  if (h->pend.models.empty() && !h->s->candidates().empty())
    return grn::set_last_error(GRN_E_INVALID, "candidates await grn_search_screen");
  const int n = (int)h->pend.models.size();
  if (n == 0) return grn_search_advance(search, nullptr, stats);
  h->results.resize(n);
  const int rc = grn::evaluate_batch_hooked(ctx, n, h->pend.datasets.data(), h->pend.models.data(),
                                            h->pend.max_errors.data(), h->pend.dts.data(), path, h->results.data(),
                                            generation_stage, h);
  if (rc) return rc;
  if (stats) *stats = h->s->stats();
  return 0;
}

int grn_search_set_oversample(void *search, int k) {
  SearchHandle *h = as_search(search);
  if (!h) return grn::set_last_error(GRN_E_INVALID, "null search");
  if (k < 1 || k > 64) return grn::set_last_error(GRN_E_INVALID, "oversample must lie in [1, 64]");
  h->s->set_oversample(k);
  return 0;
}

int grn_search_candidates(void *search, int *n, const grn_model **models) {
  SearchHandle *h = as_search(search);
  if (!h || !n) return grn::set_last_error(GRN_E_INVALID, "null argument");
  const FlatModels &c = h->candidates();
  *n = (int)c.models.size();
  if (models) *models = c.models.data();
  return 0;
}

int grn_search_screen(void *search, const double *scores) {
  SearchHandle *h = as_search(search);
  if (!h) return grn::set_last_error(GRN_E_INVALID, "null search");
  const size_t nc = h->s->candidates().size();
  if (!nc) return grn::set_last_error(GRN_E_INVALID, "no candidates to screen");
  if (!scores) return grn::set_last_error(GRN_E_INVALID, "null scores");
  for (size_t i = 0; i < nc; ++i)
    if (std::isnan(scores[i])) return grn::set_last_error(GRN_E_INVALID, "score %zu is NaN", i);
  h->s->screen(scores);
  h->flatten();
  return 0;
}

int grn_search_set_screen_fraction(void *search, double fraction) {
  SearchHandle *h = as_search(search);
  if (!h) return grn::set_last_error(GRN_E_INVALID, "null search");
  if (!(fraction > 0 && fraction <= 1)) return grn::set_last_error(GRN_E_INVALID, "fraction must lie in (0, 1]");
  h->s->set_screen_fraction(fraction);
  return 0;
}

int grn_search_set_screen_policy(void *search, int policy) {
  SearchHandle *h = as_search(search);
  if (!h) return grn::set_last_error(GRN_E_INVALID, "null search");
  if (policy != GRN_SCREEN_ABSOLUTE && policy != GRN_SCREEN_DELTA && policy != GRN_SCREEN_HALF)
    return grn::set_last_error(GRN_E_INVALID, "unknown screen policy %d", policy);
  h->s->set_screen_policy(policy);
  return 0;
}

int grn_search_record_raw(void *search, const double *raw_errors) {
  SearchHandle *h = as_search(search);
  if (!h || (!raw_errors && h->s->n_pending())) return grn::set_last_error(GRN_E_INVALID, "null argument");
  const std::vector<grn::Individual *> &pend = h->s->pending();
  for (size_t i = 0; i < pend.size(); ++i)
    pend[i]->log_raw = raw_errors[i] > 0 && std::isfinite(raw_errors[i]) ? std::log(raw_errors[i])
                                                                           : std::numeric_limits<double>::quiet_NaN();
  return 0;
}

}  // extern "C"

// Copies of the flattened models f (of the individuals list) with every
// diffusion constant times dif_scale, each on island_datasets[its island],
// in the handle's coarse arrays.
static int scaled_copies(SearchHandle *h, const std::vector<grn::Individual *> &list, const FlatModels &f,
                         const int *island_datasets, double dif_scale) {
  const int n = (int)f.models.size();
  std::vector<grn_gene> &genes = h->coarse_genes;
  std::vector<grn_model> &models = h->coarse_models;
  std::vector<int> &ds = h->coarse_ds;
  genes.assign(f.genes.begin(), f.genes.end());
  for (grn_gene &g : genes) g.dif *= dif_scale;
  models.assign(f.models.begin(), f.models.end());
  ds.resize(n);
  for (int i = 0; i < n; ++i) {
    models[i].genes = genes.data() + (f.models[i].genes - f.genes.data());
    const int island = h->s->island_of(list[i]);
    if (island < 0) return grn::set_last_error(GRN_E_INVALID, "model %d: no island", i);
    ds[i] = island_datasets[island];
  }
  return 0;
}

static int check_scale(void *search, void *ctx, const int *datasets, double dif_scale) {
  if (!search || !ctx || !datasets) return grn::set_last_error(GRN_E_INVALID, "null argument");
  if (!(dif_scale >= 0) || std::isinf(dif_scale))
    return grn::set_last_error(GRN_E_INVALID, "dif_scale must be finite and >= 0");
  return 0;
}

extern "C" {

int grn_search_run_generation_coarse(void *search, void *ctx, const int *coarse_datasets, double dif_scale,
                                     int path, grn_generation_stats *stats, double *coarse_ms) {
  if (int rc = check_scale(search, ctx, coarse_datasets, dif_scale)) return rc;
  SearchHandle *h = as_search(search);
  const auto t0 = std::chrono::steady_clock::now();
  if (h->pend.models.empty() && !h->s->candidates().empty()) {
    const FlatModels &c = h->candidates();
    const int n = (int)c.models.size();
    // the candidates with every diffusion constant scaled to the coarse grid
    if (int rc = scaled_copies(h, h->s->candidates(), c, coarse_datasets, dif_scale)) return rc;
    std::vector<grn_result> &r = h->coarse_res;
    r.resize(n);
    // each candidate keeps the time step of its full-size simulation
    const int rc = grn_evaluate_batch(ctx, n, h->coarse_ds.data(), h->coarse_models.data(), nullptr, c.dts.data(),
                                      path, r.data());
    if (rc) return rc;
    std::vector<double> &sc = h->coarse_scores;
    sc.resize(n);
    for (int i = 0; i < n; ++i) {
      const double e = r[i].raw_error;
      sc[i] = std::isnan(e) ? HUGE_VAL : std::log(std::max(e, 1e-300));
    }
    h->s->screen(sc.data());
    h->flatten();
  }
  if (coarse_ms) *coarse_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return grn_search_run_generation(search, ctx, path, stats);
}

int grn_search_run_generation_scaled(void *search, void *ctx, const int *island_datasets, double dif_scale,
                                     int path, grn_generation_stats *stats) {
  if (int rc = check_scale(search, ctx, island_datasets, dif_scale)) return rc;
  SearchHandle *h = as_search(search);
  if (h->pend.models.empty() && !h->s->candidates().empty())
    return grn::set_last_error(GRN_E_INVALID, "candidates await grn_search_screen");
  const int n = (int)h->pend.models.size();
  if (n == 0) return grn_search_advance(search, nullptr, stats);
  if (int rc = scaled_copies(h, h->s->pending(), h->pend, island_datasets, dif_scale)) return rc;
  h->results.resize(n);
  // each model keeps its own time step and its slot parent's error as maxError
  const int rc = grn::evaluate_batch_hooked(ctx, n, h->coarse_ds.data(), h->coarse_models.data(),
                                            h->pend.max_errors.data(), h->pend.dts.data(), path, h->results.data(),
                                            generation_stage, h);
  if (rc) return rc;
  if (stats) *stats = h->s->stats();
  return 0;
}

int grn_search_rescore(void *search) {
  SearchHandle *h = as_search(search);
  if (!h) return grn::set_last_error(GRN_E_INVALID, "null search");
  h->s->rescore();
  h->flatten();
  return 0;
}

int grn_search_run_generation_screened(void *search, void *ctx, void *sur, int train_steps, int path,
                                       grn_generation_stats *stats, grn_screen_stats *screen) {
  SearchHandle *h = as_search(search);
  grn::Surrogate *S = static_cast<grn::Surrogate *>(sur);
  if (!h || !ctx || !S) return grn::set_last_error(GRN_E_INVALID, "null argument");
  using clock = std::chrono::steady_clock;
  auto ms_since = [](clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(clock::now() - t0).count();
  };
  grn_screen_stats sc{};
  std::vector<double> predicted;  // for the children that get simulated
  std::vector<size_t> chosen;     // their candidate indices
  const uint16_t *cand_rows = nullptr;
  int rc;
  const int nc = (int)h->s->candidates().size();
  if (nc) {
    // score every candidate, keep the best per slot
    const auto t0 = clock::now();
    if (h->s->candidate_rows_written() && h->rows_lender == S) {
      cand_rows = S->staging_rows(0);  // written while breeding
    } else {
      uint16_t *buf = S->staging_rows(nc);
      if (!buf) return GRN_E_NOMEM;
      h->s->encode_candidates(buf, MLP_WIDTH);
      cand_rows = buf;
    }
    std::vector<double> scores(nc);
    if ((rc = S->predict_rows(nc, cand_rows, scores.data()))) return rc;
    h->s->screen(scores.data(), &chosen);
    h->flatten();
    for (size_t c : chosen) predicted.push_back(scores[c]);  // in pending order
    sc.n_candidates = nc;
    sc.predict_ms = ms_since(t0);
  }
  // simulate the children
  const auto t1 = clock::now();
  const int n = (int)h->pend.models.size();
  h->results.resize(n);
  h->errors.resize(n);
  rc = grn_evaluate_batch(ctx, n, h->pend.datasets.data(), h->pend.models.data(), h->pend.max_errors.data(),
                          h->pend.dts.data(), path, h->results.data());
  if (rc) return rc;
  sc.simulate_ms = ms_since(t1);
  sc.n_simulated = n;
  const auto t2 = clock::now();
  std::vector<double> raw(n);
  for (int i = 0; i < n; ++i) {
    h->errors[i] = h->results[i].fitness;
    raw[i] = h->results[i].raw_error;
  }
  if ((rc = grn_search_record_raw(search, raw.data()))) return rc;
  // the surrogate's target per simulated child (NaN: not a sample)
  const bool delta = h->s->screen_policy() == GRN_SCREEN_DELTA;
  std::vector<double> target(n, std::numeric_limits<double>::quiet_NaN());
  for (int i = 0; i < n; ++i) {
    if (!(raw[i] > 0 && std::isfinite(raw[i]))) continue;
    const double base = delta ? h->s->pending()[i]->parent_log_raw : 0.0;
    if (!std::isnan(base)) target[i] = std::log(raw[i]) - base;
  }
  if (!predicted.empty()) {
    std::vector<double> pv, av;
    for (int i = 0; i < n; ++i)
      if (!std::isnan(target[i])) {
        pv.push_back(predicted[i]);
        av.push_back(target[i]);
      }
    sc.spearman = grn::spearman(pv, av);
    double mae = 0;
    for (size_t i = 0; i < pv.size(); ++i) mae += std::fabs(pv[i] - av[i]);
    sc.mae = pv.empty() ? 0 : mae / pv.size();
  }
  // learn from the simulated children (their rows are candidate rows
  // already encoded, except for the initial population), then select and breed
  if (S->pair_weight() > 0 && !delta) {
    // pair mode: every child with its slot parent (the initial population
    // as pairs of a model with itself)
    std::vector<int> idx;
    std::vector<double> tc, tp;
    std::vector<const grn::Model *> cm, pm;
    for (int i = 0; i < n; ++i) {
      if (std::isnan(target[i])) continue;
      const grn::Individual *child = h->s->pending()[i];
      const grn::Individual *par = h->s->slot_parent(child);  // null in the initial population
      const bool paired = par && !std::isnan(par->log_raw);
      idx.push_back(i);
      tc.push_back(target[i]);
      tp.push_back(paired ? par->log_raw : target[i]);
      cm.push_back(&child->model);
      pm.push_back(paired ? &par->model : &child->model);
    }
    std::vector<uint16_t> crow(idx.size() * (size_t)MLP_WIDTH), prow(idx.size() * (size_t)MLP_WIDTH);
    if (cand_rows) {
      for (size_t j = 0; j < idx.size(); ++j)
        std::memcpy(crow.data() + j * MLP_WIDTH, cand_rows + chosen[idx[j]] * MLP_WIDTH,
                    sizeof(uint16_t) * MLP_WIDTH);
    } else {
      S->encode_rows(cm, crow.data());
    }
    S->encode_rows(pm, prow.data());
    if ((rc = S->add_pairs((int)idx.size(), crow.data(), prow.data(), tc.data(), tp.data()))) return rc;
  } else {
    std::vector<const grn::Model *> ms;
    std::vector<int> idx;
    std::vector<double> logs;
    for (int i = 0; i < n; ++i)
      if (!std::isnan(target[i]) && (cand_rows || !delta)) {
        idx.push_back(i);
        logs.push_back(target[i]);
      }
    std::vector<uint16_t> rows(idx.size() * (size_t)MLP_WIDTH);
    if (cand_rows) {
      for (size_t j = 0; j < idx.size(); ++j)
        std::memcpy(rows.data() + j * MLP_WIDTH, cand_rows + chosen[idx[j]] * MLP_WIDTH,
                    sizeof(uint16_t) * MLP_WIDTH);
    } else {
      for (int i : idx) ms.push_back(&h->s->pending()[i]->model);
      S->encode_rows(ms, rows.data());
    }
    if ((rc = S->add_rows((int)idx.size(), rows.data(), logs.data()))) return rc;
  }
  // the training steps run on the device while the host selects and breeds;
  // the next prediction (or any other surrogate call) waits for them
  if (train_steps > 0 && S->stats().samples > 0 && (rc = S->train_async(train_steps))) return rc;
  sc.train_ms = S->stats().train_ms;  // device time of the last completed training call
  sc.loss = S->stats().last_loss;     // and its loss
  sc.samples = S->stats().samples;
  sc.record_ms = ms_since(t2);
  const auto t3 = clock::now();
  // the next candidates are encoded into the surrogate's pinned staging rows
  // by the breeding threads
  if (h->s->oversample() > 1) {
    size_t total = 0;
    for (const auto &isl : h->s->islands()) total += isl.size();
    total *= (size_t)h->s->oversample();
    if (uint16_t *buf = S->staging_rows(total)) {
      h->s->set_candidate_rows(buf, total, MLP_WIDTH, &grn::encode_model_bf16);
      h->rows_lender = S;
    }
  }
  if ((rc = grn_search_advance(search, h->errors.data(), stats))) return rc;
  sc.breed_ms = ms_since(t3);
  if (screen) *screen = sc;
  return 0;
}

static int write_model(const grn::Model &m, grn_gene *genes, int gene_cap, int *n_genes, grn_link *links,
                       int link_cap, int *n_links) {
  *n_genes = (int)m.genes.size();
  *n_links = (int)m.links.size();
  if (*n_genes > gene_cap || *n_links > link_cap || (*n_genes && !genes) || (*n_links && !links))
    return grn::set_last_error(GRN_E_CAPACITY, "model has %d genes and %d links", *n_genes, *n_links);
  for (int i = 0; i < *n_genes; ++i) {
    const grn::Gene &g = m.genes[i];
    genes[i] = grn_gene{g.label, g.init, g.lim, g.deg, g.dif};
  }
  for (int i = 0; i < *n_links; ++i) {
    const grn::Link &l = m.links[i];
    links[i] = grn_link{l.regulator, l.regulated, l.dis_const, l.hill_coef, l.is_and ? 1 : 0};
  }
  return 0;
}

int grn_search_pareto(void *search, grn_pareto_entry *entries, int cap, int *n) {
  SearchHandle *h = as_search(search);
  if (!h || !n || cap < 0 || (cap && !entries)) return grn::set_last_error(GRN_E_INVALID, "invalid arguments");
  const std::vector<grn::Individual> &front = h->s->pareto();
  *n = (int)front.size();
  for (int i = 0; i < *n && i < cap; ++i) {
    const grn::Individual &x = front[i];
    entries[i] = grn_pareto_entry{x.error, x.complexity, x.generation, (int)(x.id >> 40),
                                  (int)x.model.genes.size(), (int)x.model.links.size(), 0};
  }
  return 0;
}

int grn_search_pareto_model(void *search, int i, grn_gene *genes, int gene_cap, int *n_genes, grn_link *links,
                            int link_cap, int *n_links) {
  SearchHandle *h = as_search(search);
  if (!h || !n_genes || !n_links) return grn::set_last_error(GRN_E_INVALID, "null argument");
  const std::vector<grn::Individual> &front = h->s->pareto();
  if (i < 0 || i >= (int)front.size())
    return grn::set_last_error(GRN_E_INVALID, "front member %d of %zu", i, front.size());
  return write_model(front[i].model, genes, gene_cap, n_genes, links, link_cap, n_links);
}

int grn_search_best(void *search, grn_gene *genes, int gene_cap, int *n_genes, grn_link *links, int link_cap,
                    int *n_links, double *error) {
  SearchHandle *h = as_search(search);
  if (!h || !n_genes || !n_links) return grn::set_last_error(GRN_E_INVALID, "null argument");
  const grn::Individual *best = nullptr;
  for (const auto &isl : h->s->islands())
    for (const grn::Individual &x : isl) {
      if (x.error < 0) continue;
      if (!best || x.error < best->error || (x.error == best->error && x.complexity < best->complexity)) best = &x;
    }
  if (!best) return grn::set_last_error(GRN_E_INVALID, "no individual has been evaluated yet");
  *n_genes = (int)best->model.genes.size();
  *n_links = (int)best->model.links.size();
  if (error) *error = best->error;
  if (*n_genes > gene_cap || *n_links > link_cap || (*n_genes && !genes) || (*n_links && !links))
    return grn::set_last_error(GRN_E_CAPACITY, "best model has %d genes and %d links", *n_genes, *n_links);
  for (int i = 0; i < *n_genes; ++i) {
    const grn::Gene &g = best->model.genes[i];
    genes[i] = grn_gene{g.label, g.init, g.lim, g.deg, g.dif};
  }
  for (int i = 0; i < *n_links; ++i) {
    const grn::Link &l = best->model.links[i];
    links[i] = grn_link{l.regulator, l.regulated, l.dis_const, l.hill_coef, l.is_and ? 1 : 0};
  }
  return 0;
}

}  // extern "C"
