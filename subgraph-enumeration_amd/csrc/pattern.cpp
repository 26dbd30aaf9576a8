// Pattern analysis: automorphism group, canonical code, 2-core, minimal
// fractional hypertree decompositions (bag count / single-bag rule) and the
// seeded pattern sampler.
//
// The reference obtains automorphisms and orbits from nauty 2.8
// (relation/graph.cpp:791-908) and FHDs from GLPK 4.55 LPs
// (plan/all_td.cpp:483-691, utility/td/SCsolver.cpp:23-76).  Patterns have at
// most 8 vertices here, so both are computed exactly without those libraries:
// automorphisms by backtracking over vertex maps, and each bag's fractional edge
// cover number through LP duality as the maximum fractional vertex packing,
// whose optimum is half-integral for graphs (enumerated over {0, 1/2, 1}).
#include <algorithm>
#include <functional>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <unordered_set>

#include "common.h"

namespace sge {

int Pattern::num_edges() const {
    int m = 0;
    for (int i = 0; i < k; ++i) m += __builtin_popcount(adj[i]);
    return m / 2;
}

static bool mask_connected(const Pattern &p, uint32_t mask) {
    if (!mask) return true;
    uint32_t seen = mask & (~mask + 1), frontier = seen;
    while (frontier) {
        uint32_t nxt = 0;
        for (int v = 0; v < p.k; ++v)
            if ((frontier >> v) & 1u) nxt |= p.adj[v];
        nxt &= mask & ~seen;
        seen |= nxt;
        frontier = nxt;
    }
    return seen == mask;
}

bool Pattern::connected() const { return mask_connected(*this, k >= 32 ? ~0u : ((1u << k) - 1)); }

Pattern from_c(const sge_pattern &c) {
    Pattern p;
    p.k = (int)c.k;
    for (int i = 0; i < SGE_MAX_K; ++i) p.adj[i] = (i < p.k) ? (c.adj[i] & ((1u << p.k) - 1)) : 0;
    return p;
}

int pattern_arg(const sge_pattern *cp, const char *fn, Pattern &out) {
    if (!cp || cp->k == 0 || cp->k > SGE_MAX_K) {
        set_error(std::string(fn) + (cp ? ": pattern must have 1..8 vertices" : ": null pattern"));
        return SGE_EINVAL;
    }
    out = from_c(*cp);
    return SGE_OK;
}

int order_arg(const uint32_t *order, int k, const char *fn) {
    uint32_t seen = 0;
    for (int i = 0; order && i < k; ++i) {
        if (order[i] >= (uint32_t)k || ((seen >> order[i]) & 1u)) break;
        seen |= 1u << order[i];
    }
    if (!order || seen != (1u << k) - 1) {
        set_error(std::string(fn) + ": order is not a permutation of the pattern vertices");
        return SGE_EINVAL;
    }
    return SGE_OK;
}

sge_pattern to_c(const Pattern &p) {
    sge_pattern c;
    std::memset(&c, 0, sizeof(c));
    c.k = (uint32_t)p.k;
    for (int i = 0; i < p.k; ++i) c.adj[i] = p.adj[i];
    return c;
}

// All permutations g with (u,v) edge <=> (g u, g v) edge; backtracking that
// checks adjacency to already-mapped vertices.
std::vector<std::vector<int>> automorphisms(const Pattern &p) {
    std::vector<std::vector<int>> out;
    int k = p.k;
    std::vector<int> img(k, -1);
    uint32_t used = 0;
    std::vector<int> deg(k);
    for (int i = 0; i < k; ++i) deg[i] = __builtin_popcount(p.adj[i]);
    std::function<void(int)> rec = [&](int i) {
        if (i == k) {
            out.push_back(img);
            return;
        }
        for (int c = 0; c < k; ++c) {
            if ((used >> c) & 1u) continue;
            if (deg[c] != deg[i]) continue;
            bool ok = true;
            for (int j = 0; j < i && ok; ++j) ok = (p.edge(i, j) == p.edge(c, img[j]));
            if (!ok) continue;
            img[i] = c;
            used |= 1u << c;
            rec(i + 1);
            used &= ~(1u << c);
        }
    };
    rec(0);
    return out;
}

// Canonical code: minimum, over labellings that list vertices by descending
// degree, of the upper-triangle adjacency word read row-major (bit set = edge).
// Two patterns are isomorphic iff their codes are equal.
uint64_t canonical_code(const Pattern &p) {
    int k = p.k;
    std::vector<int> deg(k), verts(k);
    for (int i = 0; i < k; ++i) {
        deg[i] = __builtin_popcount(p.adj[i]);
        verts[i] = i;
    }
    std::sort(verts.begin(), verts.end(), [&](int a, int b) { return deg[a] > deg[b] || (deg[a] == deg[b] && a < b); });
    // classes of equal degree, permuted independently
    std::vector<int> cls_begin;
    for (int i = 0; i < k; ++i)
        if (i == 0 || deg[verts[i]] != deg[verts[i - 1]]) cls_begin.push_back(i);
    cls_begin.push_back(k);
    uint64_t best = ~0ULL;
    std::vector<int> lab = verts;
    std::function<void(size_t)> rec = [&](size_t c) {
        if (c + 1 == cls_begin.size()) {
            uint64_t code = 0;
            for (int i = 0; i < k; ++i)
                for (int j = i + 1; j < k; ++j) code = (code << 1) | (p.edge(lab[i], lab[j]) ? 0u : 1u);
            // complement bits so that "more edges early" is smaller; prefix by k
            code |= (uint64_t)k << 56;
            best = std::min(best, code);
            return;
        }
        int b = cls_begin[c], e = cls_begin[c + 1];
        std::sort(lab.begin() + b, lab.begin() + e);
        do {
            rec(c + 1);
        } while (std::next_permutation(lab.begin() + b, lab.begin() + e));
    };
    rec(0);
    return best;
}

// Fractional edge cover number of bag B (mask) w.r.t. all pattern edges, in
// halves (returns 2*rho).  LP duality: rho*(B) = max sum_{v in B} y_v subject to
// y_u + y_v <= 1 for edges inside B and 0 <= y <= 1; extreme points are
// half-integral, so enumerate y in {0,1,2}/2.
static int cover_halves(const Pattern &p, uint32_t B) {
    int vs[SGE_MAX_K], nb = 0;
    for (int v = 0; v < p.k; ++v)
        if ((B >> v) & 1u) vs[nb++] = v;
    int best = 0;
    int y[SGE_MAX_K] = {0};
    std::function<void(int, int)> rec = [&](int i, int sum) {
        if (sum + 2 * (nb - i) <= best) return;
        if (i == nb) {
            best = std::max(best, sum);
            return;
        }
        for (int val = 2; val >= 0; --val) {
            bool ok = true;
            for (int j = 0; j < i && ok; ++j)
                if (p.edge(vs[i], vs[j]) && y[j] + val > 2) ok = false;
            if (!ok) continue;
            y[i] = val;
            rec(i + 1, sum + val);
        }
    };
    rec(0, 0);
    return best;
}

// 2-core by repeated removal of vertices of degree < 2 (relation/graph.cpp:1231-1300).
static uint32_t two_core(const Pattern &p) {
    uint32_t alive = (1u << p.k) - 1;
    bool changed = true;
    while (changed) {
        changed = false;
        for (int v = 0; v < p.k; ++v) {
            if (!((alive >> v) & 1u)) continue;
            if (__builtin_popcount(p.adj[v] & alive) < 2) {
                alive &= ~(1u << v);
                changed = true;
            }
        }
    }
    return alive;
}

// FHD of the core from one elimination order (utility/td/td.cpp:12-97
// restated): bag_i = {o_i} + later vertices sharing an (original or fill)
// hyperedge with o_i; tree edge i -> first later bag containing the fill edge;
// bags that are subsets of a tree neighbour are merged away.
struct Tree {
    std::vector<uint32_t> bags;
};

static Tree fhd_from_order(const Pattern &p, const std::vector<int> &o) {
    int n = (int)o.size();
    std::vector<uint32_t> hyper;  // original edges + fill hyperedges
    for (int u = 0; u < p.k; ++u)
        for (int v = u + 1; v < p.k; ++v)
            if (p.edge(u, v)) hyper.push_back((1u << u) | (1u << v));
    std::vector<uint32_t> X(n), fill(n);
    for (int i = 0; i < n; ++i) {
        uint32_t ei = 0;
        for (int j = i + 1; j < n; ++j)
            for (uint32_t h : hyper)
                if (((h >> o[i]) & 1u) && ((h >> o[j]) & 1u)) {
                    ei |= 1u << o[j];
                    break;
                }
        fill[i] = ei;
        hyper.push_back(ei);
        X[i] = ei | (1u << o[i]);
    }
    std::vector<std::pair<int, int>> eg;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if ((X[j] & fill[i]) == fill[i]) {
                eg.emplace_back(i, j);
                break;
            }
    bool flag = true;
    while (flag) {
        flag = false;
        for (auto ed : eg) {
            int a = ed.first, b = ed.second;
            if (!X[a] || !X[b] || a == b) continue;
            int keep = -1, drop = -1;
            if ((X[a] & X[b]) == X[a]) { keep = a; drop = b; }       // X[a] subset of X[b]: merge b into a
            else if ((X[a] & X[b]) == X[b]) { keep = b; drop = a; }  // X[b] subset of X[a]
            if (keep < 0) continue;
            // the reference keeps the first-listed index's slot but the union is the superset
            uint32_t uni = X[a] | X[b];
            X[keep] = uni;
            X[drop] = 0;
            for (auto &ej : eg) {
                if (ej.first == drop) ej.first = keep;
                if (ej.second == drop) ej.second = keep;
            }
            flag = true;
            break;
        }
    }
    Tree t;
    for (uint32_t x : X)
        if (x) t.bags.push_back(x);
    return t;
}

static Decomposition decompose_uncached(const Pattern &p, const std::vector<std::vector<int>> &aut);

// The decomposition depends on the pattern only; cache it per adjacency.
Decomposition decompose(const Pattern &p, const std::vector<std::vector<int>> &aut) {
    static std::mutex mu;
    static std::map<std::vector<uint32_t>, Decomposition> cache;
    std::vector<uint32_t> key(p.adj, p.adj + SGE_MAX_K);
    key.push_back((uint32_t)p.k);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    Decomposition d = decompose_uncached(p, aut);
    std::lock_guard<std::mutex> lk(mu);
    if (cache.size() > 4096) cache.clear();
    cache[key] = d;
    return d;
}

static Decomposition decompose_uncached(const Pattern &p, const std::vector<std::vector<int>> &aut) {
    Decomposition d;
    uint32_t core = two_core(p);
    d.core_size = __builtin_popcount(core);
    int peripheral = p.k - d.core_size;
    std::vector<int> orbit(p.k);
    for (int v = 0; v < p.k; ++v) {
        int m = v;
        for (auto &g : aut) m = std::min(m, g[v]);
        orbit[v] = m;
    }
    std::vector<int> verts;
    for (int v = 0; v < p.k; ++v)
        if ((core >> v) & 1u) verts.push_back(v);
    std::map<uint32_t, int> cover_cache;
    auto cover = [&](uint32_t B) {
        auto it = cover_cache.find(B);
        if (it != cover_cache.end()) return it->second;
        int c = cover_halves(p, B);
        cover_cache[B] = c;
        return c;
    };
    int min_width = 1 << 30;
    std::vector<Tree> optimal;
    std::set<std::vector<int>> seen_orbit_seq;
    std::set<std::multiset<uint64_t>> seen_ids;
    if (!verts.empty()) {
        do {
            std::vector<int> seq;
            for (int v : verts) seq.push_back(orbit[v]);
            if (!seen_orbit_seq.insert(seq).second) continue;
            Tree t = fhd_from_order(p, verts);
            // width over the (unmerged) elimination bags equals the max over merged bags
            int w = 0;
            for (uint32_t b : t.bags) w = std::max(w, cover(b));
            if (w > min_width) continue;
            bool conn = true;
            for (uint32_t b : t.bags) conn = conn && mask_connected(p, b);
            if (w < min_width) {
                min_width = w;
                optimal.clear();
                seen_ids.clear();
                if (!conn) continue;  // plan/all_td.cpp:527-540 clears, then skips a disconnected tree
            } else if (!conn) {
                continue;
            }
            std::multiset<uint64_t> ids;
            for (uint32_t b : t.bags) {
                Pattern sub;
                sub.k = 0;
                int map_[SGE_MAX_K];
                for (int v = 0; v < p.k; ++v)
                    if ((b >> v) & 1u) map_[sub.k++] = v;
                for (int i = 0; i < sub.k; ++i)
                    for (int j = 0; j < sub.k; ++j)
                        if (p.edge(map_[i], map_[j])) sub.adj[i] |= 1u << j;
                ids.insert(canonical_code(sub));
            }
            if (!seen_ids.insert(ids).second) continue;
            optimal.push_back(t);
        } while (std::next_permutation(verts.begin(), verts.end()));
    } else {
        min_width = 0;
    }
    d.fhw = min_width >= (1 << 30) ? 0.0 : min_width / 2.0;
    // filter by tree width (max bag size) as plan/all_td.cpp:620-650
    int min_tw = 1 << 30;
    for (auto &t : optimal) {
        int mx = 0;
        for (uint32_t b : t.bags) mx = std::max(mx, __builtin_popcount(b));
        min_tw = std::min(min_tw, mx);
    }
    std::vector<Tree> kept;
    for (auto &t : optimal) {
        int mx = 0;
        for (uint32_t b : t.bags) mx = std::max(mx, __builtin_popcount(b));
        if (mx == min_tw) kept.push_back(t);
    }
    bool all_single = !kept.empty();
    for (auto &t : kept) all_single = all_single && t.bags.size() == 1;
    int core_bags = kept.empty() ? 0 : (int)kept.front().bags.size();
    if (!kept.empty()) d.bags = kept.front().bags;
    // peripheral edges become two-vertex bags (plan/decomposition.cpp addPeripheral)
    for (int v = 0; v < p.k; ++v) {
        if ((core >> v) & 1u) continue;
        for (int u = 0; u < p.k; ++u)
            if (p.edge(u, v) && (((core >> u) & 1u) || u < v)) {
                d.bags.push_back((1u << u) | (1u << v));
                break;
            }
    }
    d.num_bags = core_bags + peripheral;
    d.single_bag = all_single && peripheral == 0;
    if (d.core_size == 0) {
        // a tree: every edge becomes its own bag (k-1 bags); a single edge is
        // one bag, so it reports embeddings like any single-bag plan
        d.num_bags = p.k - 1;
        d.single_bag = (p.k == 2);
    }
    return d;
}

}  // namespace sge

using namespace sge;

extern "C" int sge_pattern_from_edges(uint32_t k, const uint32_t *uv, uint32_t m, sge_pattern *out) {
    SGE_API_BEGIN
    if (!out || k == 0 || k > SGE_MAX_K || (m && !uv)) {
        set_error("sge_pattern_from_edges: bad arguments (k must be 1..8)");
        return SGE_EINVAL;
    }
    std::memset(out, 0, sizeof(*out));
    out->k = k;
    for (uint32_t i = 0; i < m; ++i) {
        uint32_t a = uv[2 * i], b = uv[2 * i + 1];
        if (a >= k || b >= k || a == b) {
            set_error("sge_pattern_from_edges: edge endpoint out of range or self-loop");
            return SGE_EINVAL;
        }
        out->adj[a] |= 1u << b;
        out->adj[b] |= 1u << a;
    }
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_pattern_load(const char *path, sge_pattern *out) {
    SGE_API_BEGIN
    if (!path || !out) {
        set_error("sge_pattern_load: null argument");
        return SGE_EINVAL;
    }
    uint32_t n = 0;
    EdgeVec edges;
    // pattern files share the text format of data graphs; the optional orbit
    // line after the edges only relabels vertices in the reference and does not
    // change counts (HyperTree::divideFactor is always 1), so it is ignored.
    int rc = read_edge_file(path, n, edges);
    if (rc) return rc;
    if (n == 0 || n > SGE_MAX_K) {
        set_error("pattern must have 1..8 vertices");
        return SGE_EINVAL;
    }
    std::vector<uint32_t> uv;
    for (uint64_t e : edges) {
        uv.push_back((uint32_t)(e >> 32));
        uv.push_back((uint32_t)e);
    }
    return sge_pattern_from_edges(n, uv.data(), (uint32_t)edges.size(), out);
    SGE_API_END
}

extern "C" int sge_pattern_analyze(const sge_pattern *cp, sge_pattern_info *out) {
    SGE_API_BEGIN
    if (!cp || !out || cp->k == 0 || cp->k > SGE_MAX_K) {
        set_error("sge_pattern_analyze: bad arguments");
        return SGE_EINVAL;
    }
    Pattern p = from_c(*cp);
    std::memset(out, 0, sizeof(*out));
    auto aut = automorphisms(p);
    out->k = p.k;
    out->m = p.num_edges();
    out->aut_size = aut.size();
    out->canon = canonical_code(p);
    for (int v = 0; v < p.k; ++v) {
        int m = v;
        for (auto &g : aut) m = std::min(m, g[v]);
        out->orbit[v] = m;
    }
    out->connected = p.connected();
    Decomposition d = decompose(p, aut);
    out->core_size = d.core_size;
    out->num_bags = d.num_bags;
    out->single_bag = d.single_bag;
    out->fhw = d.fhw;
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_pattern_sample(uint32_t k, uint32_t emin, uint32_t emax, uint32_t want, uint64_t seed,
                                  const uint64_t *exclude, uint32_t n_exclude, sge_pattern *out, uint32_t *got) {
    SGE_API_BEGIN
    if (!out || !got || k < 2 || k > SGE_MAX_K || emin > emax || emax > k * (k - 1) / 2) {
        set_error("sge_pattern_sample: bad arguments");
        return SGE_EINVAL;
    }
    std::vector<std::pair<int, int>> slots;
    for (uint32_t i = 0; i < k; ++i)
        for (uint32_t j = i + 1; j < k; ++j) slots.emplace_back(i, j);
    std::unordered_set<uint64_t> seen(exclude, exclude + n_exclude);
    Rng rng(seed);
    uint32_t n = 0;
    uint64_t attempts = 0, limit = 2000000ULL + 20000ULL * want;
    while (n < want && attempts++ < limit) {
        uint32_t e = emin + (uint32_t)rng.below(emax - emin + 1);
        // uniform e-subset: partial Fisher-Yates over the slot list
        std::vector<std::pair<int, int>> s = slots;
        Pattern p;
        p.k = (int)k;
        for (uint32_t i = 0; i < e; ++i) {
            uint32_t j = i + (uint32_t)rng.below(s.size() - i);
            std::swap(s[i], s[j]);
            p.adj[s[i].first] |= 1u << s[i].second;
            p.adj[s[i].second] |= 1u << s[i].first;
        }
        if (!p.connected()) continue;
        uint64_t c = canonical_code(p);
        if (!seen.insert(c).second) continue;
        out[n++] = to_c(p);
    }
    *got = n;
    return SGE_OK;
    SGE_API_END
}

// Canonical embedding rows: each row (images per pattern vertex) is replaced by
// its lexicographically smallest relabelling over Aut(P); rows are then sorted.
extern "C" int sge_canonicalize_embeddings(const sge_pattern *cp, uint32_t *rows, uint64_t n) {
    SGE_API_BEGIN
    if (n && !rows) {
        set_error("sge_canonicalize_embeddings: null rows");
        return SGE_EINVAL;
    }
    Pattern p;
    if (int rc = pattern_arg(cp, "sge_canonicalize_embeddings", p)) return rc;
    auto aut = automorphisms(p);
    int k = p.k;
    std::vector<uint32_t> best(k), cand(k);
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t *r = rows + i * k;
        for (int j = 0; j < k; ++j) best[j] = r[j];
        for (auto &g : aut) {
            for (int j = 0; j < k; ++j) cand[j] = r[g[j]];
            if (cand < best) best = cand;
        }
        for (int j = 0; j < k; ++j) r[j] = best[j];
    }
    std::vector<uint64_t> idx(n);
    for (uint64_t i = 0; i < n; ++i) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](uint64_t a, uint64_t b) {
        return std::lexicographical_compare(rows + a * k, rows + a * k + k, rows + b * k, rows + b * k + k);
    });
    std::vector<uint32_t> tmp((size_t)n * k);
    for (uint64_t i = 0; i < n; ++i) std::memcpy(tmp.data() + i * k, rows + idx[i] * k, 4ull * k);
    if (n) std::memcpy(rows, tmp.data(), 4ull * n * k);
    return SGE_OK;
    SGE_API_END
}
