// Degree-ordered CSR build on the GPU (the SGE_GRAPH_BUILD_DEVICE path of the
// graph constructors).  Same result as the host builder in graph.cpp, array for
// array: undirected edges deduplicated without self-loops, vertices relabelled
// by ascending degree with ties by original id, ascending rows.  The reference
// builds its CSR on the host (relation/graph.cpp:143-173, 255-314, 444-461).
//
//   1. normalise: (min, max) keys, self-loops to a sentinel that sorts last
//   2. radix sort + unique  -> m' undirected edges
//   3. degrees (atomics, one per run of equal smaller endpoints per wavefront),
//      rank = sort of (degree << 32 | id) -> perm / inverse
//   4. both directions in new ids, radix sort -> src (high word), nbr (low word)
//   5. offsets = exclusive scan of the relabelled degrees
// The device arrays become the graph's device mirror; off, nbr and perm are
// copied back for the host planner.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <vector>

#include "common.h"

namespace {

constexpr uint64_t SENTINEL = ~0ull;

#define BUILD_TRY(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) {                                                          \
            sge::set_error(std::string("device graph build: ") + hipGetErrorString(e_)); \
            return SGE_EDEVICE;                                                          \
        }                                                                                \
    } while (0)

__global__ void normalize_kernel(const uint64_t *__restrict__ in, uint64_t m, uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t e = in[i];
    const uint32_t u = (uint32_t)(e >> 32), v = (uint32_t)e;
    out[i] = (u == v) ? SENTINEL : (((uint64_t)min(u, v) << 32) | max(u, v));
}

// Degrees of the sorted unique edges.  The smaller endpoints arrive in runs (the
// keys are sorted), so each wavefront adds a run's length once from its first
// lane; hub vertices would otherwise serialise thousands of atomics.
__global__ void degree_kernel(const uint64_t *__restrict__ e, uint64_t m, uint32_t *__restrict__ deg) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = __lane_id();
    const bool valid = i < m;
    const uint32_t u = valid ? (uint32_t)(e[i] >> 32) : 0xffffffffu;
    const uint32_t prev = __shfl_up(u, 1);
    const bool head = valid && (lane == 0 || prev != u);
    const uint64_t heads = __ballot(head);
    const uint64_t valid_lanes = __ballot(valid);  // outside the branch: every lane votes
    if (head) {
        const uint64_t after = heads & ~((lane == 63) ? ~0ull : ((2ull << lane) - 1ull));
        const uint32_t end = after ? (uint32_t)__builtin_ctzll(after) : (uint32_t)(64 - __builtin_clzll(valid_lanes));
        atomicAdd(&deg[u], end - (uint32_t)lane);
    }
    if (valid) atomicAdd(&deg[(uint32_t)e[i]], 1u);
}

__global__ void rank_key_kernel(const uint32_t *__restrict__ deg, uint32_t n, uint64_t *__restrict__ key) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) key[v] = ((uint64_t)deg[v] << 32) | v;
}

// perm[new] = old, inv[old] = new, ndeg[new] = degree
__global__ void perm_kernel(const uint64_t *__restrict__ ranked, uint32_t n, uint32_t *__restrict__ perm,
                            uint32_t *__restrict__ inv, uint32_t *__restrict__ ndeg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = ranked[i];
    const uint32_t v = (uint32_t)k;
    perm[i] = v;
    inv[v] = i;
    ndeg[i] = (uint32_t)(k >> 32);
}

__global__ void identity_kernel(const uint32_t *__restrict__ deg, uint32_t n, uint32_t *__restrict__ perm,
                                uint32_t *__restrict__ inv, uint32_t *__restrict__ ndeg) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[i] = i;
    inv[i] = i;
    ndeg[i] = deg[i];
}

__global__ void directed_kernel(const uint64_t *__restrict__ e, uint64_t m, const uint32_t *__restrict__ inv,
                                uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t a = inv[(uint32_t)(e[i] >> 32)], b = inv[(uint32_t)e[i]];
    out[2 * i] = (a << 32) | b;
    out[2 * i + 1] = (b << 32) | a;
}

__global__ void split_kernel(const uint64_t *__restrict__ keys, uint64_t nnz, uint32_t *__restrict__ src,
                             uint32_t *__restrict__ nbr) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    src[i] = (uint32_t)(keys[i] >> 32);
    nbr[i] = (uint32_t)keys[i];
}

__global__ void set_last_kernel(uint32_t *__restrict__ off, uint32_t n, uint32_t nnz) {
    if (blockIdx.x == 0 && threadIdx.x == 0) off[n] = nnz;
}

unsigned blocks(uint64_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

// Device scratch of one build, released on every exit path.
struct Scratch {
    std::vector<void *> p;
    ~Scratch() {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename T>
    hipError_t alloc(T **out, size_t count) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) p.push_back(q);
        *out = static_cast<T *>(q);
        return e;
    }
    void release(void *q) {  // hand ownership to the caller
        p.erase(std::remove(p.begin(), p.end(), q), p.end());
    }
};

template <typename F>
hipError_t with_temp(Scratch &s, F f) {
    size_t bytes = 0;
    hipError_t e = f(nullptr, bytes);
    if (e != hipSuccess) return e;
    void *t = nullptr;
    e = s.alloc(reinterpret_cast<unsigned char **>(&t), bytes);
    if (e != hipSuccess) return e;
    return f(t, bytes);
}

}  // namespace

namespace sge {

int build_graph_device(uint32_t n, const EdgeVec &edges, uint32_t flags, sge_graph *g) {
    const uint64_t m_in = edges.size();
    if (2 * m_in >= (1ULL << 32)) {
        set_error("graph too large for u32 CSR offsets");
        return SGE_ERANGE;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device available for the device graph build");
        return SGE_EDEVICE;
    }
    for (uint64_t e : edges)
        if ((e >> 32) >= n || (uint32_t)e >= n) {
            set_error("edge endpoint out of range");
            return SGE_EINVAL;
        }
    hipStream_t st = nullptr;
    BUILD_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    Scratch s;
    uint64_t *d_in, *d_a, *d_b;
    BUILD_TRY(s.alloc(&d_in, m_in));
    BUILD_TRY(s.alloc(&d_a, 2 * m_in));
    BUILD_TRY(s.alloc(&d_b, 2 * m_in));
    BUILD_TRY(hipMemcpyAsync(d_in, edges.data(), sizeof(uint64_t) * m_in, hipMemcpyHostToDevice, st));
    // 1-2: normalise, sort, unique
    if (m_in) normalize_kernel<<<blocks(m_in), 256, 0, st>>>(d_in, m_in, d_a);
    BUILD_TRY(hipGetLastError());
    BUILD_TRY(with_temp(s, [&](void *t, size_t &b) {
        return hipcub::DeviceRadixSort::SortKeys(t, b, d_a, d_b, (int)m_in, 0, 64, st);
    }));
    uint64_t *d_count;
    BUILD_TRY(s.alloc(&d_count, 1));
    BUILD_TRY(with_temp(s, [&](void *t, size_t &b) {
        return hipcub::DeviceSelect::Unique(t, b, d_b, d_a, d_count, (int)m_in, st);
    }));
    uint64_t mu = 0, last = 0;
    BUILD_TRY(hipMemcpyAsync(&mu, d_count, sizeof(mu), hipMemcpyDeviceToHost, st));
    BUILD_TRY(hipStreamSynchronize(st));
    if (mu) {
        BUILD_TRY(hipMemcpyAsync(&last, d_a + (mu - 1), sizeof(last), hipMemcpyDeviceToHost, st));
        BUILD_TRY(hipStreamSynchronize(st));
        if (last == SENTINEL) --mu;  // self-loops
    }
    const uint64_t nnz = 2 * mu;
    // 3: degrees and ranking
    uint32_t *d_deg, *d_perm, *d_inv, *d_ndeg;
    BUILD_TRY(s.alloc(&d_deg, n));
    BUILD_TRY(s.alloc(&d_perm, n));
    BUILD_TRY(s.alloc(&d_inv, n));
    BUILD_TRY(s.alloc(&d_ndeg, n));
    BUILD_TRY(hipMemsetAsync(d_deg, 0, sizeof(uint32_t) * std::max<uint32_t>(n, 1), st));
    if (mu) degree_kernel<<<blocks(mu), 256, 0, st>>>(d_a, mu, d_deg);
    const bool ordered = flags & SGE_GRAPH_DEGREE_ORDER;
    if (ordered) {
        uint64_t *d_rk, *d_rk2;
        BUILD_TRY(s.alloc(&d_rk, n));
        BUILD_TRY(s.alloc(&d_rk2, n));
        if (n) rank_key_kernel<<<blocks(n), 256, 0, st>>>(d_deg, n, d_rk);
        BUILD_TRY(with_temp(s, [&](void *t, size_t &b) {
            return hipcub::DeviceRadixSort::SortKeys(t, b, d_rk, d_rk2, (int)n, 0, 64, st);
        }));
        if (n) perm_kernel<<<blocks(n), 256, 0, st>>>(d_rk2, n, d_perm, d_inv, d_ndeg);
    } else if (n) {
        identity_kernel<<<blocks(n), 256, 0, st>>>(d_deg, n, d_perm, d_inv, d_ndeg);
    }
    BUILD_TRY(hipGetLastError());
    // 4: both directions in new ids, sorted
    if (mu) directed_kernel<<<blocks(mu), 256, 0, st>>>(d_a, mu, d_inv, d_b);
    BUILD_TRY(with_temp(s, [&](void *t, size_t &b) {
        return hipcub::DeviceRadixSort::SortKeys(t, b, d_b, d_a, (int)nnz, 0, 64, st);
    }));
    uint32_t *d_src, *d_nbr, *d_off;
    BUILD_TRY(s.alloc(&d_src, nnz));
    BUILD_TRY(s.alloc(&d_nbr, nnz));
    BUILD_TRY(s.alloc(&d_off, (size_t)n + 1));
    if (nnz) split_kernel<<<blocks(nnz), 256, 0, st>>>(d_a, nnz, d_src, d_nbr);
    // 5: offsets
    BUILD_TRY(with_temp(s, [&](void *t, size_t &b) {
        return hipcub::DeviceScan::ExclusiveSum(t, b, d_ndeg, d_off, (int)n, st);
    }));
    set_last_kernel<<<1, 64, 0, st>>>(d_off, n, (uint32_t)nnz);
    BUILD_TRY(hipGetLastError());
    // host copy for the planner
    HostGraph &h = g->h;
    h = HostGraph();
    h.n = n;
    h.nnz = nnz;
    h.degree_ordered = ordered;
    h.off.resize((size_t)n + 1);
    h.nbr.resize(nnz);
    h.perm.resize(n);
    BUILD_TRY(hipMemcpyAsync(h.off.data(), d_off, sizeof(uint32_t) * ((size_t)n + 1), hipMemcpyDeviceToHost, st));
    if (nnz) BUILD_TRY(hipMemcpyAsync(h.nbr.data(), d_nbr, sizeof(uint32_t) * nnz, hipMemcpyDeviceToHost, st));
    if (n) BUILD_TRY(hipMemcpyAsync(h.perm.data(), d_perm, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
    BUILD_TRY(hipStreamSynchronize(st));
    h.max_degree = 0;
    for (uint32_t v = 0; v < n; ++v) h.max_degree = std::max(h.max_degree, h.off[v + 1] - h.off[v]);
    // the device arrays become the graph's device mirror (nothing can fail after this)
    int dev = -1;
    BUILD_TRY(hipGetDevice(&dev));
    s.release(d_off);
    s.release(d_nbr);
    s.release(d_src);
    g->d_off = d_off;
    g->d_nbr = d_nbr;
    g->d_src = d_src;
    g->device = dev;
    return SGE_OK;
}

}  // namespace sge
