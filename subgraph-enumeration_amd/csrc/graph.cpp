// Data-graph input, synthetic generators and CSR construction (host side).
//
// Input formats follow the reference's DataGraph loaders: text "n m" + m lines
// "u v" (relation/graph.cpp:255-314) and binary "BINGRPH" + u32 n + u32 m +
// m x (u32 u, u32 v) (relation/graph.cpp:316-419).  The CSR matches the
// reference's Graph layout (offsets[n+1], ascending neighbour lists,
// relation/graph.cpp:143-173) with one deliberate difference: duplicate edges and
// self-loops are dropped here, where the reference keeps duplicates (which
// double-counts).  Optional relabelling by ascending degree is this build's.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numeric>
#include <thread>
#include <unordered_set>

#include "common.h"

namespace sge {

static unsigned worker_count() {
    unsigned h = std::thread::hardware_concurrency();
    if (h == 0) h = 4;
    return std::min(h, 16u);
}

template <class F>
static void parallel_for(uint64_t n, F f) {
    unsigned w = worker_count();
    if (n < 4096 || w <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> ts;
    uint64_t step = (n + w - 1) / w;
    for (unsigned i = 0; i < w; ++i) {
        uint64_t b = i * step, e = std::min(n, b + step);
        if (b >= e) break;
        ts.emplace_back([=] { f(b, e); });
    }
    for (auto &t : ts) t.join();
}

int read_edge_file(const char *path, uint32_t &n, EdgeVec &edges) {
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        set_error(std::string("Can not open file ") + path + "!");
        return SGE_EIO;
    }
    char magic[8] = {0};
    size_t got = std::fread(magic, 1, 7, f);
    if (got == 7 && std::memcmp(magic, "BINGRPH", 7) == 0) {
        uint32_t hdr[2];
        if (std::fread(hdr, 4, 2, f) != 2) {
            std::fclose(f);
            set_error("truncated BINGRPH header");
            return SGE_EIO;
        }
        n = hdr[0];
        uint64_t m = hdr[1];
        // the header's m is checked against the file length before anything is sized by it
        long here = std::ftell(f);
        if (here < 0 || std::fseek(f, 0, SEEK_END) != 0) {
            std::fclose(f);
            set_error(std::string("can not seek in ") + path);
            return SGE_EIO;
        }
        long end = std::ftell(f);
        if (end < here || (uint64_t)(end - here) < 8 * m || std::fseek(f, here, SEEK_SET) != 0) {
            std::fclose(f);
            set_error("truncated BINGRPH edge list");
            return SGE_EIO;
        }
        std::vector<uint32_t> raw(2 * m);
        if (m && std::fread(raw.data(), 8, m, f) != m) {
            std::fclose(f);
            set_error("truncated BINGRPH edge list");
            return SGE_EIO;
        }
        std::fclose(f);
        edges.resize(m);
        for (uint64_t i = 0; i < m; ++i) {
            uint32_t u = raw[2 * i], v = raw[2 * i + 1];
            if (u >= n || v >= n) {
                set_error("edge endpoint out of range in BINGRPH file");
                return SGE_EIO;
            }
            edges[i] = ((uint64_t)u << 32) | v;
        }
        return SGE_OK;
    }
    std::fclose(f);
    std::ifstream in(path);
    uint64_t m = 0;
    uint64_t nn = 0;
    if (!(in >> nn >> m)) {
        set_error(std::string("bad header in ") + path);
        return SGE_EIO;
    }
    if (nn > 0xffffffffull) {
        set_error("vertex count above 2^32 - 1 in " + std::string(path));
        return SGE_EIO;
    }
    n = (uint32_t)nn;
    // grown as lines arrive: a header's m alone sizes nothing
    edges.clear();
    edges.reserve(std::min<uint64_t>(m, 1ull << 24));
    for (uint64_t i = 0; i < m; ++i) {
        uint64_t u, v;
        if (!(in >> u >> v)) {
            set_error("error: expecting " + std::to_string(m) + " edges, getting " + std::to_string(i) + " edges.");
            return SGE_EIO;
        }
        if (u >= n || v >= n) {
            set_error("edge endpoint out of range");
            return SGE_EIO;
        }
        edges.push_back((u << 32) | v);
    }
    return SGE_OK;
}

void normalize_edges(EdgeVec &edges) {
    size_t w = 0;
    for (size_t i = 0; i < edges.size(); ++i) {
        uint32_t u = (uint32_t)(edges[i] >> 32), v = (uint32_t)edges[i];
        if (u == v) continue;
        if (u > v) std::swap(u, v);
        edges[w++] = ((uint64_t)u << 32) | v;
    }
    edges.resize(w);
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
}

static void csr_from_unique(uint32_t n, const EdgeVec &edges, HostGraph &g) {
    g.n = n;
    g.nnz = 2 * (uint64_t)edges.size();
    g.off.assign((size_t)n + 1, 0);
    for (uint64_t e : edges) {
        ++g.off[(e >> 32) + 1];
        ++g.off[(uint32_t)e + 1];
    }
    for (uint32_t v = 0; v < n; ++v) g.off[v + 1] += g.off[v];
    g.nbr.assign(g.nnz, 0);
    std::vector<uint32_t> fill(g.off.begin(), g.off.end() - 1);
    // edges are sorted by (u, v): first pass writes the smaller-id halves in
    // order, second pass the larger-id halves; rows end up ascending.
    for (uint64_t e : edges) {
        uint32_t u = e >> 32, v = (uint32_t)e;
        g.nbr[fill[v]++] = u;  // u < v: all smaller neighbours of v arrive in ascending u
    }
    for (uint64_t e : edges) {
        uint32_t u = e >> 32, v = (uint32_t)e;
        g.nbr[fill[u]++] = v;
    }
    g.max_degree = 0;
    for (uint32_t v = 0; v < n; ++v) g.max_degree = std::max(g.max_degree, g.off[v + 1] - g.off[v]);
}

int build_graph(uint32_t n, EdgeVec &edges, uint32_t flags, HostGraph &g) {
    normalize_edges(edges);
    if (2 * (uint64_t)edges.size() >= (1ULL << 32)) {
        set_error("graph too large for u32 CSR offsets");
        return SGE_ERANGE;
    }
    g = HostGraph();
    if (flags & SGE_GRAPH_DEGREE_ORDER) {
        std::vector<uint32_t> deg(n, 0);
        for (uint64_t e : edges) {
            ++deg[e >> 32];
            ++deg[(uint32_t)e];
        }
        std::vector<uint32_t> perm(n);
        std::iota(perm.begin(), perm.end(), 0u);
        std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return deg[a] < deg[b]; });
        std::vector<uint32_t> inv(n);
        for (uint32_t i = 0; i < n; ++i) inv[perm[i]] = i;
        parallel_for(edges.size(), [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; ++i) {
                uint32_t u = inv[edges[i] >> 32], v = inv[(uint32_t)edges[i]];
                if (u > v) std::swap(u, v);
                edges[i] = ((uint64_t)u << 32) | v;
            }
        });
        std::sort(edges.begin(), edges.end());
        csr_from_unique(n, edges, g);
        g.perm = std::move(perm);
        g.degree_ordered = true;
    } else {
        csr_from_unique(n, edges, g);
        g.perm.resize(n);
        std::iota(g.perm.begin(), g.perm.end(), 0u);
    }
    return SGE_OK;
}

// Chung-Lu: expected degree d_i = min(C (i+1)^(-1/(gamma-1)), cap) with C set so
// that sum d_i = 2m; endpoints are drawn independently with probability
// proportional to d_i; self-loops and repeats are rejected until m distinct
// undirected edges exist.
void chung_lu_edges(uint32_t n, uint64_t m, double gamma, double cap, uint64_t seed, EdgeVec &edges) {
    std::vector<double> w(n);
    double ex = -1.0 / (gamma - 1.0);
    for (uint32_t i = 0; i < n; ++i) w[i] = std::pow((double)i + 1.0, ex);
    auto total = [&](double C) {
        double s = 0;
        for (uint32_t i = 0; i < n; ++i) s += std::min(C * w[i], cap);
        return s;
    };
    double target = 2.0 * (double)m, lo = 0, hi = 1;
    while (total(hi) < target && hi < 1e18) hi *= 2;
    for (int it = 0; it < 200; ++it) {
        double mid = 0.5 * (lo + hi);
        if (total(mid) < target) lo = mid; else hi = mid;
    }
    std::vector<double> cdf(n);
    double s = 0;
    for (uint32_t i = 0; i < n; ++i) {
        s += std::min(hi * w[i], cap);
        cdf[i] = s;
    }
    Rng rng(seed);
    auto draw = [&]() -> uint32_t {
        double r = rng.uniform() * s;
        uint32_t i = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), r) - cdf.begin());
        return std::min(i, n - 1);
    };
    std::unordered_set<uint64_t> seen;
    seen.reserve(m * 2);
    edges.clear();
    uint64_t limit = m * 1000 + 1000000;
    for (uint64_t tries = 0; edges.size() < m && tries < limit; ++tries) {
        uint32_t u = draw(), v = draw();
        if (u == v) continue;
        if (u > v) std::swap(u, v);
        uint64_t key = ((uint64_t)u << 32) | v;
        if (seen.insert(key).second) edges.push_back(key);
    }
    std::sort(edges.begin(), edges.end());
}

// R-MAT: 2^scale vertices, edge_factor * 2^scale draws; each draw descends
// `scale` levels choosing quadrant (0,0),(0,1),(1,0),(1,1) with probabilities
// a, b, c, 1-a-b-c.  Draws are grouped in fixed blocks of 2^20 with their own
// seeded stream, so the edge multiset does not depend on the thread count.
// Undirected, deduplicated, self-loops dropped (by normalize_edges).
void rmat_edges(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, EdgeVec &edges) {
    uint64_t draws = (uint64_t)ef << scale;
    const uint64_t block = 1ULL << 20;
    uint64_t nblocks = (draws + block - 1) / block;
    edges.assign(draws, 0);
    const uint64_t ta = (uint64_t)(a * 4294967296.0), tb = (uint64_t)((a + b) * 4294967296.0),
                   tc = (uint64_t)((a + b + c) * 4294967296.0);
    parallel_for(nblocks, [&](uint64_t b0, uint64_t b1) {
        for (uint64_t blk = b0; blk < b1; ++blk) {
            uint64_t x = seed ^ (blk * 0xd1342543de82ef95ULL);
            Rng rng(Rng::splitmix(x));
            uint64_t beg = blk * block, end = std::min(draws, beg + block);
            for (uint64_t i = beg; i < end; ++i) {
                uint32_t u = 0, v = 0;
                for (uint32_t l = 0; l < scale; ++l) {
                    uint64_t r = rng.next() >> 32;
                    uint32_t bu, bv;
                    if (r < ta) { bu = 0; bv = 0; }
                    else if (r < tb) { bu = 0; bv = 1; }
                    else if (r < tc) { bu = 1; bv = 0; }
                    else { bu = 1; bv = 1; }
                    u = (u << 1) | bu;
                    v = (v << 1) | bv;
                }
                edges[i] = ((uint64_t)u << 32) | v;
            }
        }
    });
}

}  // namespace sge
