// Host planner: matching order, symmetry breaking and the flat plan descriptor.
//
// Restates, from scratch, the planning steps the reference runs before the join
// (main.cpp:64-69):
//  * symmetry breaking (PatternGraph::computeCandidateRules graph.cpp:910,
//    HyperTree::selectSymmetry/setSymmetry decomposition.cpp:621-868): here the
//    textbook stabiliser-chain construction along the matching order, so every
//    subgraph instance satisfies exactly one image ordering;
//  * the clamps (largerAttrs/smallerAttrs -> setBeginPos/setEndPos,
//    join.cpp:137-178, utils.cpp:581-612) become per-level lower/upper masks;
//  * injectivity (visited[], join.cpp:369-373, and attributesToCheck in
//    TrieLevel::numResults, relation.cpp:301-338) becomes the per-level check mask;
//  * candidate reuse (HyperNode::candidatesBefore) becomes reuse[];
//  * the IEP last-level group (buildTraverseUnlabeled symmLastLevel,
//    decomposition.cpp:1022-1047, counted by choosec in join.cpp:1532-1537)
//    becomes `tail`;
//  * the single-bag embedding-count convention (decomposition.cpp:983-986)
//    becomes `multiplier` = |Aut|.
// Order choice restates the idea of optCostOrder/computeOrderCost
// (optimizer.cpp:319-577): estimated partial-match cardinality per level times
// the list work of the next level, with data-graph degree moments and a sampled
// wedge-closure rate standing in for the reference's sampled cardinalities.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <iterator>
#include <numeric>
#include <map>
#include <mutex>
#include <set>

#include "common.h"

namespace sge {

static GraphSummary summarize_uncached(const HostGraph *g);

// Degree moments and the sampled closure rate are computed once per graph
// (keyed by address and a content fingerprint) and reused for every order.
GraphSummary summarize(const HostGraph *g) {
    static std::mutex mu;
    static std::map<const HostGraph *, std::pair<uint64_t, GraphSummary>> cache;
    if (!g || g->n == 0) return summarize_uncached(g);
    uint64_t fp = ((uint64_t)g->n << 32) ^ g->nnz ^ ((uint64_t)g->off[g->n / 2] << 17) ^
                  (g->nnz ? (uint64_t)g->nbr[g->nnz / 2] << 40 : 0);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(g);
        if (it != cache.end() && it->second.first == fp) return it->second.second;
    }
    GraphSummary s = summarize_uncached(g);
    std::lock_guard<std::mutex> lk(mu);
    if (cache.size() > 64) cache.clear();
    cache[g] = {fp, s};
    return s;
}

static GraphSummary summarize_uncached(const HostGraph *g) {
    GraphSummary s;
    if (!g || g->n == 0) {
        s.n = 1000;
        s.m = 10000;
        s.avg_deg = 20;
        s.d2 = 40;
        s.tri_per_wedge = 0.05;
        s.valid = false;
        return s;
    }
    s.n = g->n;
    s.m = g->nnz / 2.0;
    double sd = 0, sd2 = 0;
    for (uint32_t v = 0; v < g->n; ++v) {
        double d = g->deg(v);
        sd += d;
        sd2 += d * d;
    }
    s.avg_deg = sd / std::max(1.0, (double)g->n);
    s.d2 = sd > 0 ? sd2 / sd : 0;
    // wedge closure: sample wedges (a - c - b) centred on a vertex chosen with
    // probability ~ deg^2 (via a random edge endpoint), test edge a-b
    Rng rng(12345);
    uint64_t closed = 0, total = 0;
    for (int it = 0; it < 20000 && g->nnz > 0; ++it) {
        uint64_t e = rng.below(g->nnz);
        uint32_t c = (uint32_t)(std::upper_bound(g->off.begin(), g->off.end(), (uint32_t)e) - g->off.begin()) - 1;
        uint32_t d = g->deg(c);
        if (d < 2) continue;
        uint32_t a = g->nbr[g->off[c] + rng.below(d)], b = g->nbr[g->off[c] + rng.below(d)];
        if (a == b) continue;
        ++total;
        const uint32_t *row = g->nbr.data() + g->off[a];
        if (std::binary_search(row, row + g->deg(a), b)) ++closed;
    }
    s.tri_per_wedge = total ? std::max(1e-4, (double)closed / total) : 0.01;
    s.valid = true;
    return s;
}

// Stabiliser chain along `order`: for each level i, every other vertex in the
// orbit of order[i] under the pointwise stabiliser of order[0..i-1] must map to
// a larger data vertex.  Returns lower masks (bit i set in lower[level(w)]).
static void symmetry_masks(const Pattern &p, const int *order, const std::vector<std::vector<int>> &aut,
                           uint32_t *lower) {
    int k = p.k;
    int level_of[SGE_MAX_K];
    for (int i = 0; i < k; ++i) level_of[order[i]] = i;
    std::vector<const std::vector<int> *> G;
    for (auto &g : aut) G.push_back(&g);
    for (int i = 0; i < k; ++i) lower[i] = 0;
    for (int i = 0; i < k; ++i) {
        int u = order[i];
        uint32_t orbit = 0;
        for (auto *g : G) orbit |= 1u << (*g)[u];
        for (int w = 0; w < k; ++w)
            if (w != u && ((orbit >> w) & 1u)) lower[level_of[w]] |= 1u << i;
        std::vector<const std::vector<int> *> H;
        for (auto *g : G)
            if ((*g)[u] == u) H.push_back(g);
        G.swap(H);
    }
}

std::vector<std::vector<int>> connected_orders(const Pattern &p, const std::vector<std::vector<int>> &aut) {
    std::vector<std::vector<int>> out;
    std::set<std::vector<int>> seen;
    std::vector<int> o(p.k);
    std::iota(o.begin(), o.end(), 0);
    do {
        bool ok = true;
        uint32_t placed = 1u << o[0];
        for (int i = 1; i < p.k && ok; ++i) {
            ok = (p.adj[o[i]] & placed) != 0;
            placed |= 1u << o[i];
        }
        if (!ok) continue;
        std::vector<int> best = o;
        for (auto &g : aut) {
            std::vector<int> t(p.k);
            for (int i = 0; i < p.k; ++i) t[i] = g[o[i]];
            best = std::min(best, t);
        }
        if (seen.insert(best).second) out.push_back(best);
    } while (std::next_permutation(o.begin(), o.end()));
    return out;
}

// Trailing symmetric group (kind 1): the last r levels are pairwise
// non-adjacent twins that share every parent, bound and check below the group
// and are chained among themselves; choosec(|C|, r) counts them
// (decomposition.cpp:1022-1047, join.cpp:1532-1537).
static bool valid_sym_tail(const sge_plan &pl, int r) {
    int k = pl.k, b = k - r;
    if (r < 2 || b < 2) return false;  // the root edge (levels 0, 1) is never part of the group
    uint32_t below = (1u << b) - 1;
    for (int i = b; i < k; ++i) {
        if ((pl.parent[i] & ~below) != 0) return false;  // group members non-adjacent
        if ((pl.parent[i] & below) != (pl.parent[b] & below)) return false;
        if ((pl.lower[i] & below) != (pl.lower[b] & below)) return false;
        if ((pl.upper[i] & below) != (pl.upper[b] & below)) return false;
        if ((pl.check[i] & below) != (pl.check[b] & below)) return false;
        uint32_t chain = ((1u << i) - 1) & ~below;  // all earlier group members
        if ((pl.lower[i] & ~below) != chain) return false;
        if ((pl.upper[i] & ~below) != 0) return false;
    }
    return true;
}

// Independent tail (kind 2): the last r levels are pairwise non-adjacent and
// carry no symmetry bound among themselves, so for fixed images of the levels
// below, the number of injective assignments is an inclusion-exclusion sum
// over set partitions of the tail of products of intersection sizes.
static bool valid_ie_tail(const sge_plan &pl, int r) {
    int k = pl.k, b = k - r;
    if (r < 2 || r > 3 || b < 2) return false;
    uint32_t below = (1u << b) - 1;
    for (int i = b; i < k; ++i) {
        if ((pl.parent[i] | pl.lower[i] | pl.upper[i]) & ~below) return false;
    }
    return true;
}

static void choose_tail(sge_plan &pl) {
    pl.tail = 0;
    pl.tail_kind = 0;
    if (pl.mode != SGE_MODE_IEP) return;
    int k = pl.k;
    int sym = 0, ie = 0;
    for (int r = k - 1; r >= 2 && !sym; --r)
        if (valid_sym_tail(pl, r)) sym = r;
    for (int r = 3; r >= 2 && !ie; --r)
        if (valid_ie_tail(pl, r)) ie = r;
    if (sym && sym >= ie) {
        pl.tail = sym;
        pl.tail_kind = 1;
    } else if (ie) {
        pl.tail = ie;
        pl.tail_kind = 2;
    }
}

static void fill_masks(const Pattern &p, const int *order, const std::vector<std::vector<int>> &aut, uint32_t mode,
                       sge_plan &pl) {
    int k = p.k;
    pl.k = k;
    pl.mode = mode;
    for (int i = 0; i < SGE_MAX_K; ++i) {
        pl.order[i] = i < k ? order[i] : 0;
        pl.parent[i] = pl.lower[i] = pl.upper[i] = pl.check[i] = 0;
        pl.reuse[i] = -1;
        pl.ready[i] = 0;
    }
    symmetry_masks(p, order, aut, pl.lower);
    for (int i = 0; i < k; ++i) {
        for (int j = 0; j < i; ++j)
            if (p.edge(order[i], order[j])) pl.parent[i] |= 1u << j;
        uint32_t earlier = (1u << i) - 1;
        pl.check[i] = earlier & ~(pl.parent[i] | pl.lower[i] | pl.upper[i]);
        pl.ready[i] = pl.parent[i] ? (31 - __builtin_clz(pl.parent[i])) : 0;
    }
    // candidate reuse: an earlier level whose filters are all implied by this
    // level's filters and whose parents are a subset of this level's parents
    for (int j = 3; j < k; ++j) {
        int best = -1, bestp = 1;
        for (int i = 2; i < j; ++i) {
            if ((pl.parent[i] & pl.parent[j]) != pl.parent[i]) continue;
            int np = __builtin_popcount(pl.parent[i]);
            if (np < 2) continue;
            if ((pl.lower[i] & ~pl.lower[j]) || (pl.upper[i] & ~pl.upper[j])) continue;
            uint32_t excl_j = pl.check[j] | pl.lower[j] | pl.upper[j] | pl.parent[j];
            if (pl.check[i] & ~excl_j) continue;
            if (np > bestp) {
                best = i;
                bestp = np;
            }
        }
        pl.reuse[j] = best;
    }
    choose_tail(pl);
}

// Analytic cost of a plan, following the executor's structure: root tasks are
// directed edges; levels 2..s are materialised per parent node (work = the
// clamped parent lists); the closed-form level(s) are counted for all
// siblings of level s at once, their sibling-independent parts formed once per
// node of level s-1 (one per inclusion-exclusion term).  Sizes use degree
// moments (a list reached through an edge has the size-biased degree
// E[d^2]/E[d]) and a sampled wedge-closure rate q per extra list; a symmetry
// bound halves a set.
static double plan_cost(const sge_plan &pl, const GraphSummary &s) {
    int k = pl.k;
    if (k <= 2) return 1.0;
    double q = std::min(0.9, std::max(1e-4, s.tri_per_wedge));
    double d2 = std::max(1.0, s.d2);
    auto set_size = [&](uint32_t par, bool clamped) {
        int np = __builtin_popcount(par);
        double sz = d2 * std::pow(q, std::max(np - 1, 0));
        if (clamped) sz *= 0.5;
        return std::max(sz, 1e-6);
    };
    int r = pl.tail >= 2 ? (int)pl.tail : 1;
    int first = k - r;  // first closed-form level
    int sib = first - 1;
    double card = 2.0 * s.m * ((pl.lower[1] & 1u) ? 0.5 : 1.0);  // nodes at level 1
    double cost = card * 2.0;
    for (int i = 2; i <= sib; ++i) {
        int np = __builtin_popcount(pl.parent[i]);
        cost += card * (np * d2 * (pl.lower[i] ? 0.6 : 1.0) + 8.0);
        card *= set_size(pl.parent[i], pl.lower[i] != 0);
    }
    double parent_nodes = card / set_size(pl.parent[sib], pl.lower[sib] != 0);
    // terms: one per non-empty subset of an independent tail, else one
    int nterms = (pl.tail_kind == 2) ? (1 << r) - 1 : 1;
    for (int t = 1; t <= nterms; ++t) {
        uint32_t par = 0;
        if (pl.tail_kind == 2) {
            for (int j = 0; j < r; ++j)
                if ((t >> j) & 1) par |= pl.parent[first + j];
        } else {
            par = pl.parent[first];
        }
        uint32_t fpar = par & ~(1u << sib);
        bool s_par = (par >> sib) & 1u;
        int nf = __builtin_popcount(fpar);
        double fsize = nf ? d2 * std::pow(q, nf - 1) : 0.0;
        if (nf) cost += parent_nodes * nf * d2;
        double per_sib = s_par ? (nf ? std::min(fsize, d2) * std::log2(std::max(fsize, d2) + 2.0)
                                     : std::log2(d2 + 2.0))
                               : std::log2(fsize + 2.0);
        cost += card * (per_sib + 4.0);
    }
    return cost;
}

double order_cost(const Pattern &p, const GraphSummary &s, const int *order, const std::vector<std::vector<int>> &aut,
                  uint32_t mode) {
    for (int i = 1; i < p.k; ++i) {
        uint32_t placed = 0;
        for (int j = 0; j < i; ++j) placed |= 1u << order[j];
        if (!(p.adj[order[i]] & placed)) return 1e300;
    }
    sge_plan pl;
    std::memset(&pl, 0, sizeof(pl));
    fill_masks(p, order, aut, mode, pl);
    return plan_cost(pl, s);
}

int build_plan(const Pattern &p, const HostGraph *g, uint32_t mode, const uint32_t *order_in, sge_plan &out) {
    if (p.k < 1 || p.k > SGE_MAX_K) {
        set_error("pattern must have 1..8 vertices");
        return SGE_EINVAL;
    }
    if (!p.connected()) {
        set_error("pattern must be connected");
        return SGE_EINVAL;
    }
    if (mode > SGE_MODE_IEP) {
        set_error("unknown mode");
        return SGE_EINVAL;
    }
    std::memset(&out, 0, sizeof(out));
    auto aut = automorphisms(p);
    Decomposition d = decompose(p, aut);
    GraphSummary s = summarize(g);
    int order[SGE_MAX_K];
    double cost = 0;
    if (order_in) {
        uint32_t seen = 0;
        for (int i = 0; i < p.k; ++i) {
            if (order_in[i] >= (uint32_t)p.k || ((seen >> order_in[i]) & 1u)) {
                set_error("order is not a permutation of the pattern vertices");
                return SGE_EINVAL;
            }
            seen |= 1u << order_in[i];
            order[i] = (int)order_in[i];
        }
        for (int i = 1; i < p.k; ++i) {
            uint32_t placed = 0;
            for (int j = 0; j < i; ++j) placed |= 1u << order[j];
            if (!(p.adj[order[i]] & placed)) {
                set_error("order is not connected (a level has no earlier neighbour)");
                return SGE_EINVAL;
            }
        }
        cost = order_cost(p, s, order, aut, mode);
    } else {
        auto orders = connected_orders(p, aut);
        double best = 1e308;
        int besti = 0, best_tail = -1;
        for (size_t i = 0; i < orders.size(); ++i) {
            sge_plan trial;
            std::memset(&trial, 0, sizeof(trial));
            fill_masks(p, orders[i].data(), aut, mode, trial);
            double c = plan_cost(trial, s);
            // IEP: a closed-form tail removes whole enumeration levels, so the
            // longest tail wins and the cost model breaks ties
            int tl = (int)trial.tail;
            if (tl > best_tail || (tl == best_tail && c < best)) {
                best = c;
                besti = (int)i;
                best_tail = tl;
            }
        }
        for (int i = 0; i < p.k; ++i) order[i] = orders[besti][i];
        cost = best;
    }
    fill_masks(p, order, aut, mode, out);
    out.num_bags = d.num_bags;
    out.aut_size = aut.size();
    out.multiplier = d.single_bag ? aut.size() : 1;
    out.cost = cost;
    return SGE_OK;
}

// ---- materialised bags
// The level at which the closed-form count starts (every level below it is
// enumerated).
static uint32_t first_closed_level(const sge_plan &pl) {
    return (pl.mode == SGE_MODE_IEP && pl.tail >= 2) ? pl.k - pl.tail : pl.k - 1;
}

int bag_check(const sge_plan &pl, uint32_t B) {
    if (B == 0) return SGE_OK;
    if (__builtin_popcount(B) < 3) {
        set_error("bag: a materialised bag needs at least 3 vertices");
        return SGE_EINVAL;
    }
    if (pl.k < 4 || (B >> first_closed_level(pl))) {
        set_error("bag: every bag level must lie below the closed-form level(s)");
        return SGE_EINVAL;
    }
    const int b0 = __builtin_ctz(B);
    const uint32_t rest = B & (B - 1);
    const int b1 = __builtin_ctz(rest);
    if (!((pl.parent[b1] >> b0) & 1u)) {
        set_error("bag: the first two bag levels must be adjacent");
        return SGE_EINVAL;
    }
    for (uint32_t m = rest & (rest - 1); m; m &= m - 1) {
        const int l = __builtin_ctz(m);
        if (!(pl.parent[l] & B)) {
            set_error("bag: every bag level after the second needs an earlier neighbour in the bag");
            return SGE_EINVAL;
        }
    }
    return SGE_OK;
}

// Every bag of a plan (level masks in bits 0-7 and 8-15 of pl.bag): each a
// valid bag, the second only beside a first and different from it, nothing
// above bit 15.
int bags_check(const sge_plan &pl) {
    if (pl.bag >> (8 * SGE_MAX_BAGS)) {
        set_error("bag: at most two bags per plan (level masks in bits 0-7 and 8-15)");
        return SGE_EINVAL;
    }
    const uint32_t b0 = pl.bag & 0xffu, b1 = (pl.bag >> 8) & 0xffu;
    if (b1 && (!b0 || b0 == b1)) {
        set_error("bag: a second bag needs a first one and must differ from it");
        return SGE_EINVAL;
    }
    if (int rc = bag_check(pl, b0)) return rc;
    return bag_check(pl, b1);
}

// The bag relation's own plan: the levels of bag mask B in plan order, masks
// restricted to the bag and renumbered, no closed-form tail, output column i =
// bag level i.
void bag_subplan(const sge_plan &pl, uint32_t B, sge_plan &sub) {
    std::memset(&sub, 0, sizeof(sub));
    int lev[SGE_MAX_K], p = 0;
    for (uint32_t m = B; m; m &= m - 1) lev[p++] = __builtin_ctz(m);
    auto remap = [&](uint32_t mask) {
        uint32_t r = 0;
        for (int i = 0; i < p; ++i)
            if ((mask >> lev[i]) & 1u) r |= 1u << i;
        return r;
    };
    sub.k = (uint32_t)p;
    sub.mode = SGE_MODE_WCOJ;
    for (int i = 0; i < SGE_MAX_K; ++i) sub.reuse[i] = -1;
    for (int i = 0; i < p; ++i) {
        sub.order[i] = (uint32_t)i;
        sub.parent[i] = remap(pl.parent[lev[i]]);
        sub.lower[i] = remap(pl.lower[lev[i]]);
        sub.upper[i] = remap(pl.upper[lev[i]]);
        sub.check[i] = remap(pl.check[lev[i]]);
    }
    sub.num_bags = 1;
    sub.multiplier = 1;
    sub.aut_size = 1;
}

}  // namespace sge

using namespace sge;

extern "C" int sge_plan_set_bag(sge_plan *plan, uint32_t vertex_mask) {
    SGE_API_BEGIN
    if (!plan || plan->k < 1 || plan->k > SGE_MAX_K) {
        set_error("sge_plan_set_bag: bad plan");
        return SGE_EINVAL;
    }
    if (vertex_mask >> plan->k) {
        set_error("sge_plan_set_bag: vertex mask names vertices beyond k");
        return SGE_EINVAL;
    }
    uint32_t B = 0;
    for (uint32_t i = 0; i < plan->k; ++i)
        if ((vertex_mask >> plan->order[i]) & 1u) B |= 1u << i;
    int rc = bag_check(*plan, B);
    if (rc) return rc;
    plan->bag = B;
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_plan_add_bag(sge_plan *plan, uint32_t vertex_mask) {
    SGE_API_BEGIN
    if (!plan || plan->k < 1 || plan->k > SGE_MAX_K) {
        set_error("sge_plan_add_bag: bad plan");
        return SGE_EINVAL;
    }
    if (vertex_mask >> plan->k) {
        set_error("sge_plan_add_bag: vertex mask names vertices beyond k");
        return SGE_EINVAL;
    }
    if (!(plan->bag & 0xffu) || (plan->bag >> 8)) {
        set_error("sge_plan_add_bag: the plan needs exactly one bag already (sge_plan_set_bag)");
        return SGE_EINVAL;
    }
    uint32_t B = 0;
    for (uint32_t i = 0; i < plan->k; ++i)
        if ((vertex_mask >> plan->order[i]) & 1u) B |= 1u << i;
    sge_plan trial = *plan;
    trial.bag |= B << 8;
    int rc = bags_check(trial);
    if (rc) return rc;
    plan->bag = trial.bag;
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_plan_bags(const sge_pattern *cp, const sge_plan *plan, uint32_t *masks, uint32_t cap,
                             uint32_t *n) {
    SGE_API_BEGIN
    if (!plan || !n || (cap && !masks)) {
        set_error("sge_plan_bags: null argument");
        return SGE_EINVAL;
    }
    Pattern p;
    if (int rc = pattern_arg(cp, "sge_plan_bags", p)) return rc;
    if ((uint32_t)p.k != plan->k) {
        set_error("sge_plan_bags: plan and pattern sizes differ");
        return SGE_EINVAL;
    }
    Decomposition d = decompose(p, automorphisms(p));
    uint32_t cnt = 0;
    for (uint32_t vm : d.bags) {
        sge_plan trial = *plan;
        if (sge_plan_set_bag(&trial, vm) != SGE_OK) continue;
        if (cnt < cap) masks[cnt] = vm;
        ++cnt;
    }
    set_error("");
    *n = cnt;
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_plan_build(const sge_pattern *p, const sge_graph *g, uint32_t mode, const uint32_t *order,
                              sge_plan *out) {
    SGE_API_BEGIN
    if (!out) {
        set_error("sge_plan_build: null argument");
        return SGE_EINVAL;
    }
    Pattern pat;
    if (int rc = pattern_arg(p, "sge_plan_build", pat)) return rc;
    return build_plan(pat, g ? &g->h : nullptr, mode, order, *out);
    SGE_API_END
}

extern "C" int sge_plan_cost(const sge_pattern *cp, const sge_graph *g, uint32_t mode, const uint32_t *order,
                             double *cost) {
    SGE_API_BEGIN
    if (!cost) {
        set_error("sge_plan_cost: null argument");
        return SGE_EINVAL;
    }
    if (mode > SGE_MODE_IEP) {
        set_error("sge_plan_cost: unknown mode");
        return SGE_EINVAL;
    }
    Pattern p;
    if (int rc = pattern_arg(cp, "sge_plan_cost", p)) return rc;
    if (int rc = order_arg(order, p.k, "sge_plan_cost")) return rc;
    auto aut = automorphisms(p);
    int o[SGE_MAX_K];
    for (int i = 0; i < p.k; ++i) o[i] = (int)order[i];
    *cost = order_cost(p, summarize(g ? &g->h : nullptr), o, aut, mode);
    return SGE_OK;
    SGE_API_END
}

// Every connected order ranked the way the planner ranks them for `mode`:
// longest closed-form tail first (IEP), then lowest analytic cost.
extern "C" int sge_plan_rank_orders(const sge_pattern *cp, const sge_graph *g, uint32_t mode, uint32_t *orders,
                                    double *costs, uint32_t *tails, uint32_t cap, uint32_t *n) {
    SGE_API_BEGIN
    if (!n || mode > SGE_MODE_IEP) {
        set_error("sge_plan_rank_orders: bad arguments");
        return SGE_EINVAL;
    }
    Pattern p;
    if (int rc = pattern_arg(cp, "sge_plan_rank_orders", p)) return rc;
    if (!p.connected() || p.k < 2) {
        set_error("sge_plan_rank_orders: pattern must be connected with >= 2 vertices");
        return SGE_EINVAL;
    }
    auto aut = automorphisms(p);
    GraphSummary s = summarize(g ? &g->h : nullptr);
    auto os = connected_orders(p, aut);
    struct R { int tail; double cost; size_t i; };
    std::vector<R> rs;
    for (size_t i = 0; i < os.size(); ++i) {
        sge_plan trial;
        std::memset(&trial, 0, sizeof(trial));
        fill_masks(p, os[i].data(), aut, mode, trial);
        rs.push_back({(int)trial.tail, plan_cost(trial, s), i});
    }
    std::stable_sort(rs.begin(), rs.end(), [](const R &a, const R &b) {
        return a.tail != b.tail ? a.tail > b.tail : a.cost < b.cost;
    });
    *n = (uint32_t)rs.size();
    for (uint32_t i = 0; i < std::min<uint32_t>(cap, *n); ++i) {
        if (orders)
            for (int j = 0; j < p.k; ++j) orders[i * p.k + j] = os[rs[i].i][j];
        if (costs) costs[i] = rs[i].cost;
        if (tails) tails[i] = rs[i].tail;
    }
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_plan_orders(const sge_pattern *cp, uint32_t *orders, uint32_t cap, uint32_t *n) {
    SGE_API_BEGIN
    if (!n) {
        set_error("sge_plan_orders: null argument");
        return SGE_EINVAL;
    }
    Pattern p;
    if (int rc = pattern_arg(cp, "sge_plan_orders", p)) return rc;
    auto aut = automorphisms(p);
    auto os = connected_orders(p, aut);
    *n = (uint32_t)os.size();
    if (orders)
        for (uint32_t i = 0; i < std::min<uint32_t>(cap, *n); ++i)
            for (int j = 0; j < p.k; ++j) orders[i * p.k + j] = os[i][j];
    return SGE_OK;
    SGE_API_END
}

// Plan text in the reference's -r / stdout style (decomposition.cpp:537-553
// prints "num bags" and bags; symmetry rules are printed as "u < a,b,").
extern "C" int sge_plan_text(const sge_plan *pl, char *buf, size_t cap) {
    SGE_API_BEGIN
    if (!pl || !buf || cap == 0) {
        set_error("sge_plan_text: bad arguments");
        return SGE_EINVAL;
    }
    std::string s = "num bags: " + std::to_string(pl->num_bags) + "\n";
    s += "order:";
    for (uint32_t i = 0; i < pl->k; ++i) s += " " + std::to_string(pl->order[i]);
    s += "\n";
    for (uint32_t i = 0; i < pl->k; ++i) {
        std::string larger;
        for (uint32_t j = 0; j < pl->k; ++j)
            if ((pl->lower[j] >> i) & 1u) larger += std::to_string(pl->order[j]) + ",";
        if (!larger.empty()) s += std::to_string(pl->order[i]) + " < " + larger + "\n";
    }
    for (int b = 0; b < SGE_MAX_BAGS; ++b) {
        const uint32_t B = (pl->bag >> (8 * b)) & 0xffu;
        if (!B) continue;
        s += "materialised bag:";
        for (uint32_t i = 0; i < pl->k; ++i)
            if ((B >> i) & 1u) s += " " + std::to_string(pl->order[i]);
        s += "\n";
    }
    if (s.size() + 1 > cap) {
        set_error("sge_plan_text: buffer too small");
        return SGE_ERANGE;
    }
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return SGE_OK;
    SGE_API_END
}

// ---------------------------------------------------------------------------
// Features of one (pattern, order) for the learned cost model.
//
// 64 floats: pattern shape (k, m, |Aut|, fhw, bag count, degree sequence),
// per-level plan structure (parents, clamps, checks, reuse, analytic size),
// graph moments, the analytic cost, and sampled statistics: a Knuth random-walk
// estimate over `nsamples` random root edges of the number of search-tree nodes
// and of the list work at every level (a seeded restatement of the reference's
// sampled cardinalities, plan/estimator.cpp:562-661).
namespace sge {

// Root edges are drawn from the CSR positions [rb, re) (the task set a run
// covers: the whole graph, or a root-edge range), each weighted by re - rb.
static void knuth_estimate(const HostGraph &g, const sge_plan &pl, uint32_t nsamples, uint64_t seed, uint64_t rb,
                           uint64_t re, double *card, double *work) {
    int k = pl.k;
    for (int i = 0; i < SGE_MAX_K; ++i) card[i] = work[i] = 0;
    if (g.nnz == 0 || k < 3 || re <= rb) return;
    Rng rng(seed);
    std::vector<uint32_t> cur, nxt;
    cur.reserve(g.max_degree + 1);
    nxt.reserve(g.max_degree + 1);
    for (uint32_t s = 0; s < nsamples; ++s) {
        uint64_t e = rb + rng.below(re - rb);
        uint32_t v0 = (uint32_t)(std::upper_bound(g.off.begin(), g.off.end(), (uint32_t)e) - g.off.begin()) - 1;
        uint32_t v1 = g.nbr[e];
        if ((pl.lower[1] & 1u) && !(v1 > v0)) continue;
        uint32_t phi[SGE_MAX_K] = {v0, v1};
        double w = (double)(re - rb);
        card[1] += w;
        for (int d = 2; d < k; ++d) {
            uint32_t lo = 0, hi = 0xffffffffu;
            bool has_lo = false, has_hi = false;
            for (int j = 0; j < d; ++j) {
                if ((pl.lower[d] >> j) & 1u) { lo = std::max(lo, phi[j]); has_lo = true; }
                if ((pl.upper[d] >> j) & 1u) { hi = std::min(hi, phi[j]); has_hi = true; }
            }
            bool first = true;
            double wk = 0;
            for (int j = 0; j < d; ++j) {
                if (!((pl.parent[d] >> j) & 1u)) continue;
                const uint32_t *row = g.nbr.data() + g.off[phi[j]];
                uint32_t n = g.deg(phi[j]);
                const uint32_t *b = has_lo ? std::upper_bound(row, row + n, lo) : row;
                const uint32_t *en = has_hi ? std::lower_bound(b, row + n, hi) : row + n;
                if (en < b) en = b;
                wk += (double)(en - b);
                if (first) {
                    cur.assign(b, en);
                    first = false;
                } else {
                    nxt.clear();
                    std::set_intersection(cur.begin(), cur.end(), b, en, std::back_inserter(nxt));
                    cur.swap(nxt);
                }
            }
            // injectivity
            size_t wpos = 0;
            for (size_t i = 0; i < cur.size(); ++i) {
                bool bad = false;
                for (int j = 0; j < d; ++j)
                    if (((pl.check[d] >> j) & 1u) && phi[j] == cur[i]) bad = true;
                if (!bad) cur[wpos++] = cur[i];
            }
            cur.resize(wpos);
            work[d] += w * wk;
            card[d] += w * (double)cur.size();
            if (cur.empty() || d == k - 1) break;
            w *= (double)cur.size();
            phi[d] = cur[rng.below(cur.size())];
        }
    }
    for (int i = 0; i < SGE_MAX_K; ++i) {
        card[i] /= std::max(1u, nsamples);
        work[i] /= std::max(1u, nsamples);
    }
}

int order_features(const Pattern &p, const HostGraph *g, const uint32_t *order, uint32_t mode, uint32_t nsamples,
                   uint64_t seed, uint64_t rb, uint64_t re, float *f) {
    sge_plan pl;
    int rc = build_plan(p, g, mode, order, pl);
    if (rc) return rc;
    for (int i = 0; i < 64; ++i) f[i] = 0.f;
    auto aut = automorphisms(p);
    Decomposition d = decompose(p, aut);
    GraphSummary s = summarize(g);
    int k = p.k;
    f[0] = k / 8.f;
    f[1] = p.num_edges() / 28.f;
    f[2] = (float)(std::log2((double)aut.size()) / 16.0);
    f[3] = (float)(d.fhw / 4.0);
    f[4] = d.num_bags / 8.f;
    std::vector<int> deg;
    for (int v = 0; v < k; ++v) deg.push_back(__builtin_popcount(p.adj[v]));
    std::sort(deg.rbegin(), deg.rend());
    for (int i = 0; i < k && i < 7; ++i) f[5 + i] = deg[i] / 7.f;
    double q = std::min(0.9, std::max(1e-4, s.tri_per_wedge));
    for (int i = 1; i < k && i < 8; ++i) {
        int base = 12 + (i - 1) * 5;
        int np = __builtin_popcount(pl.parent[i]);
        f[base + 0] = np / 7.f;
        f[base + 1] = pl.lower[i] ? 1.f : 0.f;
        f[base + 2] = __builtin_popcount(pl.check[i]) / 7.f;
        f[base + 3] = pl.reuse[i] >= 0 ? 1.f : 0.f;
        double sz = (i == 1 ? s.avg_deg : s.d2) * std::pow(q, np - 1) * (pl.lower[i] ? 0.5 : 1.0);
        f[base + 4] = (float)(std::log2(std::max(sz, 1e-6)) / 20.0);
    }
    double card[SGE_MAX_K], work[SGE_MAX_K];
    if (g) knuth_estimate(*g, pl, nsamples, seed, rb, re ? std::min<uint64_t>(re, g->nnz) : g->nnz, card, work);
    else for (int i = 0; i < SGE_MAX_K; ++i) card[i] = work[i] = 0;
    double total = 0;
    for (int i = 1; i < 8; ++i) {
        f[46 + i] = (float)(std::log2(1.0 + card[i]) / 48.0);
        f[53 + i] = (float)(std::log2(1.0 + work[i]) / 48.0);
        total += work[i] + card[i];
    }
    f[61] = (float)(std::log2(1.0 + total) / 48.0);
    f[62] = (float)(std::log2(1.0 + std::max(0.0, pl.cost)) / 64.0);
    f[63] = (float)(std::log2(1.0 + s.m) / 32.0);
    f[46] = pl.tail / 8.f;
    return SGE_OK;
}

}  // namespace sge

extern "C" int sge_order_features(const sge_pattern *cp, const sge_graph *g, const uint32_t *order, uint32_t mode,
                                  uint32_t nsamples, uint64_t seed, float *feat) {
    SGE_API_BEGIN
    if (!order || !feat) {
        set_error("sge_order_features: null argument");
        return SGE_EINVAL;
    }
    Pattern p;
    if (int rc = pattern_arg(cp, "sge_order_features", p)) return rc;
    return order_features(p, g ? &g->h : nullptr, order, mode, nsamples, seed, 0, 0, feat);
    SGE_API_END
}

extern "C" int sge_order_features_range(const sge_pattern *cp, const sge_graph *g, const uint32_t *order,
                                        uint32_t mode, uint32_t nsamples, uint64_t seed, uint64_t begin, uint64_t end,
                                        float *feat) {
    SGE_API_BEGIN
    if (!order || !feat) {
        set_error("sge_order_features_range: null argument");
        return SGE_EINVAL;
    }
    if (g && (begin >= end || end > g->h.nnz)) {
        set_error("sge_order_features_range: root range must satisfy begin < end <= number of directed edges");
        return SGE_EINVAL;
    }
    Pattern p;
    if (int rc = pattern_arg(cp, "sge_order_features_range", p)) return rc;
    return order_features(p, g ? &g->h : nullptr, order, mode, nsamples, seed, begin, end, feat);
    SGE_API_END
}
