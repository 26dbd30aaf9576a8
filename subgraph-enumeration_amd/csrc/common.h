// Internal host-side structures shared by the C++ translation units of libsge.
#pragma once
#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>
#include <utility>

#include "sge.h"

struct sge_graph;

namespace sge {

void set_error(const std::string &msg);

// Every int-returning C entry point runs its body inside this pair, so that no
// C++ exception unwinds into a C, ctypes or cgo caller (which would abort the
// process): host allocation failures become SGE_ENOMEM, anything else
// SGE_EINVAL, each with its message in sge_last_error().
#define SGE_API_BEGIN try {
#define SGE_API_END                                                               \
    }                                                                             \
    catch (const std::bad_alloc &) {                                              \
        sge::set_error("host memory allocation failed");                          \
        return SGE_ENOMEM;                                                        \
    }                                                                             \
    catch (const std::length_error &e) {                                          \
        sge::set_error(std::string("host allocation too large: ") + e.what());    \
        return SGE_ENOMEM;                                                        \
    }                                                                             \
    catch (const std::exception &e) {                                             \
        sge::set_error(std::string("internal error: ") + e.what());               \
        return SGE_EINVAL;                                                        \
    }

// Symmetric CSR, ascending neighbour lists, u32 offsets (the reference's EdgeID
// is u32 too, config/config.h:43-45).  perm[new] = original id.
struct HostGraph {
    uint32_t n = 0;
    uint64_t nnz = 0;
    uint32_t max_degree = 0;
    bool degree_ordered = false;
    std::vector<uint32_t> off;   // n + 1
    std::vector<uint32_t> nbr;   // nnz
    std::vector<uint32_t> perm;  // n (new -> original)
    uint32_t deg(uint32_t v) const { return off[v + 1] - off[v]; }
};

using EdgeVec = std::vector<uint64_t>;  // (u << 32) | v, u < v

// graph.cpp
int read_edge_file(const char *path, uint32_t &n, EdgeVec &edges);
void normalize_edges(EdgeVec &edges);  // drop self loops, orient u<v, sort, unique
int build_graph(uint32_t n, EdgeVec &edges, uint32_t flags, HostGraph &g);
void chung_lu_edges(uint32_t n, uint64_t m, double gamma, double cap, uint64_t seed, EdgeVec &edges);
void rmat_edges(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, EdgeVec &edges);

// Deterministic RNG (splitmix64 seeding + xoshiro256**).
struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto &w : s) w = splitmix(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

// pattern.cpp
struct Pattern {
    int k = 0;
    uint32_t adj[SGE_MAX_K] = {0};
    bool edge(int a, int b) const { return (adj[a] >> b) & 1u; }
    int num_edges() const;
    bool connected() const;
};
Pattern from_c(const sge_pattern &p);
// A caller's pattern for entry point `fn`: non-null with 1..8 vertices (SGE_EINVAL otherwise).
int pattern_arg(const sge_pattern *cp, const char *fn, Pattern &out);
// A caller's vertex order for a k-vertex pattern: non-null and a permutation of 0..k-1.
int order_arg(const uint32_t *order, int k, const char *fn);
sge_pattern to_c(const Pattern &p);
std::vector<std::vector<int>> automorphisms(const Pattern &p);
uint64_t canonical_code(const Pattern &p);
struct Decomposition {
    double fhw = 0;
    int core_size = 0;
    int num_bags = 0;
    bool single_bag = false;
    std::vector<uint32_t> bags;  // bitmasks of the chosen tree (incl. peripheral edge bags)
};
Decomposition decompose(const Pattern &p, const std::vector<std::vector<int>> &aut);

// trie.hip: level-order trie of a bag relation (device arrays, see sge_trie in sge.h)
struct BagTrie {
    uint32_t *val = nullptr, *child = nullptr;
    uint32_t vb[SGE_MAX_K + 1] = {0};  // first value index of each depth
    uint32_t depth = 0;
    uint64_t tuples = 0;
    uint64_t nodes[SGE_MAX_K] = {0};
    double build_ms = 0;
    void *desc = nullptr;  // executor's device descriptor
};
// rows: n x p images (row-major, device); ids < n_vertices; stream: hipStream_t
int trie_from_rows(const uint32_t *d_rows, uint64_t n, uint32_t p, uint32_t n_vertices, void *stream, BagTrie &t);
void trie_free(BagTrie &t);

// graph_build.hip: the SGE_GRAPH_BUILD_DEVICE path (fills g->h and g's device mirror)
int build_graph_device(uint32_t n, const EdgeVec &edges, uint32_t flags, sge_graph *g);

// planner.cpp
struct GraphSummary {
    double n = 1, m = 0, avg_deg = 0, d2 = 0;  // d2 = E[deg^2]/E[deg]
    double tri_per_wedge = 0;                 // sampled clustering
    bool valid = false;
};
GraphSummary summarize(const HostGraph *g);
// estimator.cpp: the reference's sampled cost model of one plan's order
double sampled_order_cost(const Pattern &p, const HostGraph &g, const sge_plan &pl, uint64_t seed);
double order_cost(const Pattern &p, const GraphSummary &s, const int *order, const std::vector<std::vector<int>> &aut,
                  uint32_t mode);
int build_plan(const Pattern &p, const HostGraph *g, uint32_t mode, const uint32_t *order, sge_plan &out);
std::vector<std::vector<int>> connected_orders(const Pattern &p, const std::vector<std::vector<int>> &aut);
int bag_check(const sge_plan &pl, uint32_t level_mask);  // SGE_OK, or SGE_EINVAL with the reason
int bags_check(const sge_plan &pl);                        // every bag of pl.bag (bits 0-7, 8-15)
void bag_subplan(const sge_plan &pl, uint32_t level_mask, sge_plan &sub);  // plan of one bag's relation

}  // namespace sge

struct sge_graph {
    sge::HostGraph h;
    double build_ms = 0;  // wall time of the CSR build
    // device mirror (executor.hip)
    void *d_off = nullptr, *d_nbr = nullptr, *d_src = nullptr;
    void *d_skip = nullptr;  // skip index of the rows (executor.hip), built at the first launch
    int device = -1;
    void *split_cache = nullptr;  // executor-owned: per-shape subtask prefix arrays on the device
    sge::GraphSummary summary;
    bool summary_ready = false;
    // batches borrow the graph's device arrays: sge_graph_destroy defers the
    // release while any batch of it is alive, and the last one releases it
    uint32_t batches = 0;
    bool destroyed = false;
};

// Free a graph's device mirror and the graph itself (api.cpp).
void sge_graph_release(sge_graph *g);
