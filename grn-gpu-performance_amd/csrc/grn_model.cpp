// grn_model.cpp -- GRN model closure, text format and program compiler.
#include "grn_model.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <set>

#include "grn.h"

namespace grn {

std::vector<int> Model::labels_in_use(int n_morphogens) const {
  // The labels reachable from the structural labels 0..nMorphogens-1 over
  // regulated -> regulator edges (model.cpp:243-270).  A model names tens of
  // labels: numbered in a table of at most 64, the closure is a fixpoint on
  // a 64-bit set over the links' (regulated, regulator) pairs; a model with
  // more distinct labels takes the breadth-first walk below.
  constexpr int kTab = 64, kEdges = 256;
  if (n_morphogens >= 0 && n_morphogens <= kTab && links.size() <= (size_t)kEdges) {
    int tab[kTab];
    int nt = n_morphogens;
    for (int i = 0; i < nt; ++i) tab[i] = i;
    auto id = [&](int lab) {
      for (int i = 0; i < nt; ++i)
        if (tab[i] == lab) return i;
      if (nt == kTab) return -1;
      tab[nt] = lab;
      return nt++;
    };
    uint8_t from[kEdges], to[kEdges];
    bool ok = true;
    const int ne = (int)links.size();
    for (int e = 0; e < ne && ok; ++e) {
      const int a = id(links[e].regulated), b = id(links[e].regulator);
      ok = a >= 0 && b >= 0;
      from[e] = (uint8_t)a;
      to[e] = (uint8_t)b;
    }
    if (ok) {
      uint64_t set = n_morphogens == 64 ? ~0ull : (1ull << n_morphogens) - 1;
      for (bool grew = true; grew;) {
        grew = false;
        for (int e = 0; e < ne; ++e)
          if (((set >> from[e]) & 1) && !((set >> to[e]) & 1)) {
            set |= 1ull << to[e];
            grew = true;
          }
      }
      std::vector<int> out;
      out.reserve(__builtin_popcountll(set));
      for (int i = 0; i < nt; ++i)
        if ((set >> i) & 1) out.push_back(tab[i]);
      std::sort(out.begin(), out.end());
      return out;
    }
  }
  // (a linear membership test beats a tree at these sizes)
  std::vector<int> order;
  order.reserve(n_morphogens + links.size());
  for (int i = 0; i < n_morphogens; ++i) order.push_back(i);
  for (size_t head = 0; head < order.size(); ++head) {
    const int lab = order[head];
    for (const Link &l : links)
      if (l.regulated == lab && std::find(order.begin(), order.end(), l.regulator) == order.end())
        order.push_back(l.regulator);
  }
  std::sort(order.begin(), order.end());
  return order;
}

std::vector<int> Model::links_to_label(int label) const {
  std::vector<int> out;
  for (size_t i = 0; i < links.size(); ++i)
    if (links[i].regulated == label) out.push_back((int)i);
  return out;
}

const Gene *Model::gene_with_label(int label) const {
  for (const Gene &g : genes)
    if (g.label == label) return &g;
  return nullptr;
}

int Model::complexity_in_use(int n_morphogens) const {
  const std::vector<int> in_use = labels_in_use(n_morphogens);  // sorted
  auto used = [&](int label) { return std::binary_search(in_use.begin(), in_use.end(), label); };
  int c = 0;
  for (const Gene &g : genes)
    if (used(g.label)) c += 1 + (g.dif > 0);  // ModelProd::complexity
  for (const Link &l : links)
    if (used(l.regulator) && used(l.regulated)) c += 1;  // ModelLink::complexity
  return c;
}

// ---- text format -----------------------------------------------------------

static void put_num(std::string *s, double v, int precision) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*g", precision, v);
  *s += buf;
}

std::string Model::to_string(int precision) const {
  std::string s = "(";
  for (size_t i = 0; i < genes.size(); ++i) {
    if (i) s += '|';
    const Gene &g = genes[i];
    s += std::to_string(g.label);
    s += ' ';
    put_num(&s, g.init, precision);
    s += ' ';
    put_num(&s, g.lim, precision);
    s += ' ';
    put_num(&s, g.deg, precision);
    s += ' ';
    put_num(&s, g.dif, precision);
  }
  s += '*';
  for (size_t i = 0; i < links.size(); ++i) {
    if (i) s += '|';
    const Link &l = links[i];
    s += std::to_string(l.regulator);
    s += ' ';
    s += std::to_string(l.regulated);
    s += ' ';
    put_num(&s, l.dis_const, precision);
    s += ' ';
    put_num(&s, l.hill_coef, precision);
    s += ' ';
    s += l.is_and ? '1' : '0';
  }
  s += ')';
  return s;
}

namespace {

struct Cursor {
  const char *p, *end;
  void skip_ws() {
    while (p < end && (*p == '\n' || *p == '\r' || *p == '\t')) ++p;
  }
  bool eat(char c) {
    skip_ws();
    if (p < end && *p == c) {
      ++p;
      return true;
    }
    return false;
  }
  bool peek(char c) {
    skip_ws();
    return p < end && *p == c;
  }
  // Model::parseDouble (model.cpp:690-706): characters up to ' ', '|', '*'
  // or ')'.
  bool number(double *v) {
    skip_ws();
    const char *s = p;
    while (p < end && *p != ' ' && *p != '|' && *p != '*' && *p != ')') ++p;
    if (p == s) return false;
    std::string tok(s, p);
    char *e = nullptr;
    *v = std::strtod(tok.c_str(), &e);
    return e && *e == '\0';
  }
  bool integer(int *v) {
    double d;
    if (!number(&d) || d != std::floor(d)) return false;
    *v = (int)d;
    return true;
  }
};

}  // namespace

bool Model::parse(const std::string &text, Model *out, std::string *err) {
  Model m;
  Cursor c{text.data(), text.data() + text.size()};
  auto fail = [&](const char *what) {
    if (err) *err = std::string("model text: ") + what;
    return false;
  };
  if (!c.eat('(')) return fail("expected '('");
  for (;;) {  // products until '*'
    Gene g;
    double dl;
    if (!c.number(&dl) || dl != std::floor(dl)) return fail("bad product label");
    g.label = (int)dl;
    if (!c.eat(' ') || !c.number(&g.init)) return fail("bad product init");
    if (!c.eat(' ') || !c.number(&g.lim)) return fail("bad product lim");
    if (!c.eat(' ') || !c.number(&g.deg)) return fail("bad product deg");
    if (!c.eat(' ') || !c.number(&g.dif)) return fail("bad product dif");
    m.genes.push_back(g);
    if (c.eat('|')) continue;
    if (c.eat('*')) break;
    return fail("expected '|' or '*' after a product");
  }
  if (!c.eat(')')) {
    for (;;) {  // links until ')'
      Link l;
      int and_flag = 0;
      if (!c.integer(&l.regulator)) return fail("bad link regulator");
      if (!c.eat(' ') || !c.integer(&l.regulated)) return fail("bad link regulated");
      if (!c.eat(' ') || !c.number(&l.dis_const)) return fail("bad link K");
      if (!c.eat(' ') || !c.number(&l.hill_coef)) return fail("bad link n");
      if (!c.eat(' ') || !c.integer(&and_flag)) return fail("bad link isAnd");
      l.is_and = and_flag != 0;
      m.links.push_back(l);
      if (c.eat('|')) continue;
      if (c.eat(')')) break;
      return fail("expected '|' or ')' after a link");
    }
  }
  if (m.genes.empty()) return fail("model without products");
  *out = std::move(m);
  return true;
}

// ---- compiler ---------------------------------------------------------------

int compile_model(const Model &m, int n_morphogens, int dataset, float dt, GrnProgram *out,
                  ProgramInfo *info, std::string *err, GrnProgram64 *out64, double dt64) {
  const std::vector<int> labels = m.labels_in_use(n_morphogens);
  const int G = (int)labels.size();
  if (G > GRN_PMAX_GENES) {
    if (err) *err = "model has " + std::to_string(G) + " genes in use; the kernels hold at most " +
                    std::to_string(GRN_PMAX_GENES);
    return GRN_E_CAPACITY;
  }
  std::memset(out, 0, sizeof(*out));
  if (out64) {
    std::memset(out64, 0, sizeof(*out64));
    out64->dt = dt64;
  }
  out->G = G;
  out->dataset = dataset;
  out->dt = dt;
  ProgramInfo inf;
  inf.G = G;
  auto index_of = [&](int label) {
    return (int)(std::lower_bound(labels.begin(), labels.end(), label) - labels.begin());
  };
  const float inf_f = std::numeric_limits<float>::infinity();
  int nl = 0;
  for (int i = 0; i < G; ++i) {
    const Gene *g = m.gene_with_label(labels[i]);
    if (!g) {
      if (err) *err = "label " + std::to_string(labels[i]) + " is in use but has no product";
      return GRN_E_MODEL;
    }
    GrnGene &gp = out->gene[i];
    gp.lim = (float)g->lim;
    gp.deg = (float)g->deg;
    gp.dif = g->dif > 0 ? (float)g->dif : 0.f;  // only dif > 0 diffuses (modelsimulator.cpp:106)
    inf.D += g->dif > 0;
    if (out64) out64->gene[i] = GrnGene64{g->lim, g->deg, g->dif > 0 ? g->dif : 0.0};

    // Links into this gene in model order, split OR / AND; h = 0 unless one
    // of them activates (modelsimulator.cpp:114-138).
    std::vector<int> or_l, and_l;
    bool any_act = false;
    for (int li : m.links_to_label(labels[i])) {
      const Link &l = m.links[li];
      if (std::binary_search(labels.begin(), labels.end(), l.regulator)) (l.is_and ? and_l : or_l).push_back(li);
      if (l.hill_coef >= 0) any_act = true;
    }
    gp.link_begin = nl;
    if (!any_act) {
      gp.n_or = gp.n_and = 0;
      gp.tor0 = -inf_f;
      gp.tand0 = inf_f;
      continue;
    }
    if ((int)(nl + or_l.size() + and_l.size()) > GRN_PMAX_LINKS) {
      if (err) *err = "model has more than " + std::to_string(GRN_PMAX_LINKS) + " Hill links";
      return GRN_E_CAPACITY;
    }
    gp.n_or = (int)or_l.size();
    gp.n_and = (int)and_l.size();
    gp.tor0 = gp.n_or > 0 ? inf_f : -inf_f;
    gp.tand0 = -inf_f;
    inf.G_h += 1;
    for (int pass = 0; pass < 2; ++pass) {
      for (int li : (pass == 0 ? or_l : and_l)) {
        const Link &l = m.links[li];
        // Exponent form of createHillOpForLink (modelsimulator.cpp:194-203):
        // activation x^n/(x^n+K^n) = 1/(1+2^t), t = -n log2 x + n log2 K;
        // repression  K^n/(x^n+K^n) = 1/(1+2^t), t = +n log2 x - n log2 K.
        const bool act = l.hill_coef >= 0;
        const double n = act ? l.hill_coef : -l.hill_coef;
        const double nlk = n == 0 ? 0.0 : n * std::log2(l.dis_const);
        if (out64) out64->link[nl] = GrnLink64{act ? -n : n, act ? nlk : -nlk};
        GrnLink &gl = out->link[nl++];
        gl.r = index_of(l.regulator);
        gl.a = (float)(act ? -n : n);
        gl.b = (float)(act ? nlk : -nlk);
        gl.pad = 0;
      }
    }
  }
  out->n_links = nl;
  for (int i = 0; i < G; ++i) {
    const GrnGene &gp = out->gene[i];
    if (gp.n_or + gp.n_and > 0) out->hill_mask |= 1 << i;
    if (gp.dif != 0.f) out->dif_mask |= 1 << i;
  }
  for (int l = 0; l < nl; ++l) out->reg_mask |= 1 << out->link[l].r;
  // The fast step drops the minConc clamp and the log2 floor.  Exact update
  // c' = c (1 - dt deg - dt dif n) + dt (lim h + dif sum(neighbours)) stays
  // >= 0 from a nonnegative state when lim, deg, dif >= 0 and
  // dt (deg + 4 dif) <= 0.99 (the margin covers the rounding of dt*R); the
  // caller adds the dataset's conditions (nonnegative inputs, minConc <= 0).
  // log2(0) = -inf is harmless once no link has a zero Hill exponent
  // (a*(-inf) + b = +-inf, and h = 0 or 1 as x^n/(x^n+K^n) gives).
  bool fast = true;
  for (int i = 0; i < G; ++i) {
    const GrnGene &gp = out->gene[i];
    if (!(gp.lim >= 0.f && gp.deg >= 0.f && gp.dif >= 0.f)) fast = false;
    if (!((double)dt * ((double)gp.deg + 4.0 * (double)gp.dif) <= 0.99)) fast = false;
  }
  for (int l = 0; l < nl; ++l)
    if (out->link[l].a == 0.f) fast = false;
  out->fast = fast ? 1 : 0;
  inf.L = nl;
  inf.flops_per_cell_step = 5.0 * inf.L + (inf.L - inf.G_h) + 8.0 * inf.G + 7.0 * inf.D;
  if (info) *info = inf;
  return 0;
}

int validate_program(const GrnProgram &p, int n_datasets, std::string *err) {
  auto bad = [&](const std::string &what) {
    if (err) *err = "compiled program failed validation: " + what;
    return GRN_E_MODEL;
  };
  if (p.G < 1 || p.G > GRN_PMAX_GENES) return bad("gene count " + std::to_string(p.G));
  if (p.n_links < 0 || p.n_links > GRN_PMAX_LINKS) return bad("link count " + std::to_string(p.n_links));
  if (p.dataset < 0 || p.dataset >= n_datasets) return bad("dataset " + std::to_string(p.dataset));
  if (!std::isfinite(p.dt)) return bad("time step");
  const int all = (int)((1u << p.G) - 1u);
  if ((p.reg_mask | p.hill_mask | p.dif_mask) & ~all) return bad("gene masks beyond the genes in use");
  int hill = 0, dif = 0, reg = 0;
  for (int g = 0; g < p.G; ++g) {
    const GrnGene &q = p.gene[g];
    if (q.n_or < 0 || q.n_and < 0 || q.link_begin < 0 || q.link_begin + q.n_or + q.n_and > p.n_links)
      return bad("link range of gene " + std::to_string(g));
    if (q.n_or + q.n_and > 0) hill |= 1 << g;
    if (q.dif != 0.f) dif |= 1 << g;
  }
  for (int l = 0; l < p.n_links; ++l) {
    const int r = p.link[l].r;
    if (r < 0 || r >= p.G) return bad("regulator index of link " + std::to_string(l));
    reg |= 1 << r;
  }
  if (hill != p.hill_mask || dif != p.dif_mask || reg != p.reg_mask) return bad("gene masks");
  return 0;
}

bool model_from_abi(const grn_model *gm, Model *m) {
  if (!gm || gm->n_genes < 0 || gm->n_links < 0 || (gm->n_genes && !gm->genes) || (gm->n_links && !gm->links))
    return false;
  m->genes.resize(gm->n_genes);
  for (int i = 0; i < gm->n_genes; ++i) {
    const grn_gene &g = gm->genes[i];
    m->genes[i] = Gene{g.label, g.init, g.lim, g.deg, g.dif};
  }
  m->links.resize(gm->n_links);
  for (int i = 0; i < gm->n_links; ++i) {
    const grn_link &l = gm->links[i];
    m->links[i] = Link{l.regulator, l.regulated, l.dis_const, l.hill_coef, l.is_and != 0};
  }
  return true;
}

}  // namespace grn
