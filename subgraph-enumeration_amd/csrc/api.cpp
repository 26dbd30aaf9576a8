// C-ABI entry points for graphs and errors (include/sge.h).
#include <chrono>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>

#include "common.h"

namespace sge {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
}  // namespace sge

using namespace sge;

extern "C" void sge_graph_device_free(sge_graph *g);

extern "C" const char *sge_last_error(void) { return g_err.c_str(); }
extern "C" int sge_version(void) { return 10000; }

static int finish_graph(uint32_t n, EdgeVec &edges, uint32_t flags, sge_graph **out) {
    if (!out) {
        set_error("null output handle");
        return SGE_EINVAL;
    }
    std::unique_ptr<sge_graph> g(new sge_graph());  // freed if the build fails or throws
    const auto t0 = std::chrono::steady_clock::now();
    int rc = (flags & SGE_GRAPH_BUILD_DEVICE) ? build_graph_device(n, edges, flags, g.get())
                                              : build_graph(n, edges, flags, g->h);
    if (rc) return rc;
    g->build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = g.release();
    return SGE_OK;
}

extern "C" int sge_graph_load(const char *path, uint32_t flags, sge_graph **out) {
    SGE_API_BEGIN
    if (!path) {
        set_error("null path");
        return SGE_EINVAL;
    }
    uint32_t n = 0;
    EdgeVec edges;
    int rc = read_edge_file(path, n, edges);
    if (rc) return rc;
    return finish_graph(n, edges, flags, out);
    SGE_API_END
}

extern "C" int sge_graph_from_edges(uint32_t n, const uint32_t *uv, uint64_t m, uint32_t flags, sge_graph **out) {
    SGE_API_BEGIN
    if (m && !uv) {
        set_error("null edge list");
        return SGE_EINVAL;
    }
    EdgeVec edges(m);
    for (uint64_t i = 0; i < m; ++i) {
        if (uv[2 * i] >= n || uv[2 * i + 1] >= n) {
            set_error("edge endpoint out of range");
            return SGE_EINVAL;
        }
        edges[i] = ((uint64_t)uv[2 * i] << 32) | uv[2 * i + 1];
    }
    return finish_graph(n, edges, flags, out);
    SGE_API_END
}

extern "C" int sge_graph_from_csr(const uint32_t *off, const uint32_t *nbr, uint32_t n, uint64_t nnz, uint32_t flags,
                                  sge_graph **out) {
    SGE_API_BEGIN
    if (!off || (nnz && !nbr)) {
        set_error("null CSR array");
        return SGE_EINVAL;
    }
    if (off[n] != nnz) {
        set_error("offsets[n] != nnz");
        return SGE_EINVAL;
    }
    EdgeVec edges;
    edges.reserve(nnz / 2);
    for (uint32_t v = 0; v < n; ++v) {
        if (off[v + 1] < off[v]) {
            set_error("offsets not monotone");
            return SGE_EINVAL;
        }
        for (uint32_t e = off[v]; e < off[v + 1]; ++e) {
            uint32_t u = nbr[e];
            if (u >= n) {
                set_error("neighbour id out of range");
                return SGE_EINVAL;
            }
            if (v < u) edges.push_back(((uint64_t)v << 32) | u);
        }
    }
    return finish_graph(n, edges, flags, out);
    SGE_API_END
}

extern "C" int sge_graph_chung_lu(uint32_t n, uint64_t m, double gamma, double cap, uint64_t seed, uint32_t flags,
                                  sge_graph **out) {
    SGE_API_BEGIN
    if (n < 2 || gamma <= 1.0 || m > (uint64_t)n * (n - 1) / 2) {
        set_error("chung_lu: need n >= 2, gamma > 1, m <= n(n-1)/2");
        return SGE_EINVAL;
    }
    EdgeVec edges;
    chung_lu_edges(n, m, gamma, cap > 0 ? cap : 1e300, seed, edges);
    if (edges.size() != m) {  // the draw limit ran out (m near n(n-1)/2 under a steep degree law)
        set_error("chung_lu: drew only " + std::to_string(edges.size()) + " of " + std::to_string(m) +
                  " distinct edges; lower m or gamma");
        return SGE_ERANGE;
    }
    return finish_graph(n, edges, flags, out);
    SGE_API_END
}

extern "C" int sge_graph_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, uint32_t flags,
                              sge_graph **out) {
    SGE_API_BEGIN
    if (scale < 1 || scale > 30 || ef == 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) {
        set_error("rmat: bad parameters");
        return SGE_EINVAL;
    }
    EdgeVec edges;
    rmat_edges(scale, ef, a, b, c, seed, edges);
    return finish_graph(1u << scale, edges, flags, out);
    SGE_API_END
}

extern "C" int sge_graph_stats_get(const sge_graph *g, sge_graph_stats *s) {
    SGE_API_BEGIN
    if (!g || !s) {
        set_error("null argument");
        return SGE_EINVAL;
    }
    std::memset(s, 0, sizeof(*s));
    s->n = g->h.n;
    s->m = g->h.nnz / 2;
    s->max_degree = g->h.max_degree;
    s->degree_ordered = g->h.degree_ordered;
    s->csr_bytes = 4ull * ((uint64_t)g->h.n + 1) + 4ull * g->h.nnz;
    s->build_ms = g->build_ms;
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_graph_csr(const sge_graph *g, uint32_t *off, uint32_t *nbr) {
    SGE_API_BEGIN
    if (!g || !off || (g->h.nnz && !nbr)) {
        set_error("null argument");
        return SGE_EINVAL;
    }
    std::memcpy(off, g->h.off.data(), 4ull * ((uint64_t)g->h.n + 1));
    if (g->h.nnz) std::memcpy(nbr, g->h.nbr.data(), 4ull * g->h.nnz);
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_graph_perm(const sge_graph *g, uint32_t *perm) {
    SGE_API_BEGIN
    if (!g || !perm) {
        set_error("null argument");
        return SGE_EINVAL;
    }
    std::memcpy(perm, g->h.perm.data(), 4ull * g->h.n);
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_graph_write(const sge_graph *g, const char *path, int binary) {
    SGE_API_BEGIN
    if (!g || !path) {
        set_error("null argument");
        return SGE_EINVAL;
    }
    const HostGraph &h = g->h;
    std::vector<uint64_t> es;
    es.reserve(h.nnz / 2);
    for (uint32_t v = 0; v < h.n; ++v)
        for (uint32_t e = h.off[v]; e < h.off[v + 1]; ++e) {
            uint32_t a = h.perm[v], b = h.perm[h.nbr[e]];
            if (a < b) es.push_back(((uint64_t)a << 32) | b);
        }
    std::sort(es.begin(), es.end());
    if (binary && es.size() > 0xffffffffull) {  // BINGRPH stores m as u32
        set_error("BINGRPH holds at most 2^32 - 1 edges");
        return SGE_ERANGE;
    }
    FILE *f = std::fopen(path, binary ? "wb" : "w");
    if (!f) {
        set_error(std::string("Can not create file ") + path + "!");
        return SGE_EIO;
    }
    bool ok = true;
    if (binary) {
        uint32_t hdr[2] = {h.n, (uint32_t)es.size()};
        ok = std::fwrite("BINGRPH", 1, 7, f) == 7 && std::fwrite(hdr, 4, 2, f) == 2;
        for (size_t i = 0; ok && i < es.size(); ++i) {
            uint32_t uv[2] = {(uint32_t)(es[i] >> 32), (uint32_t)es[i]};
            ok = std::fwrite(uv, 4, 2, f) == 2;
        }
    } else {
        ok = std::fprintf(f, "%u %llu\n", h.n, (unsigned long long)es.size()) > 0;
        for (size_t i = 0; ok && i < es.size(); ++i)
            ok = std::fprintf(f, "%u %u\n", (uint32_t)(es[i] >> 32), (uint32_t)es[i]) > 0;
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        set_error(std::string("write failed: ") + path);
        return SGE_EIO;
    }
    return SGE_OK;
    SGE_API_END
}

void sge_graph_release(sge_graph *g) {
    sge_graph_device_free(g);
    delete g;
}

extern "C" void sge_graph_destroy(sge_graph *g) {
    if (!g) return;
    if (g->batches) {  // a live batch still reads the device arrays: the last batch releases the graph
        g->destroyed = true;
        return;
    }
    sge_graph_release(g);
}
