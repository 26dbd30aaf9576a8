// Level-order trie of a bag relation on the GPU: the reference's buildTrie
// (matching/join.cpp:1713-1717) -> TrieLevel::buildTrieFromSortedMatchesBatch
// (relation/relation.cpp:340-379), which std::sort()s the bag's tuples and
// groups them recursively by depth into nested TrieLevel objects.  Here the
// tuples (N rows of p images, written by the executor's enumeration mode) are
//
//   1. sorted lexicographically by an LSD radix sort of a row permutation,
//      one stable pass per column from the last to the first (hipCUB onesweep),
//      then gathered column-major;
//   2. each row gets the first column in which it differs from the row before
//      (row 0: column 0): row i starts a new trie node at every depth j >= it;
//   3. per depth j a scan of those "starts a node" flags numbers the nodes;
//      node values are scattered to val[vb[j] + id] and each node's first
//      child, the node number of its first row at depth j + 1, to the offsets.
//
// The result is the reference's nested trie flattened depth by depth: the same
// values, with every node's children as one contiguous sorted range.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <string>
#include <vector>

#include "common.h"

namespace {

#define TRIE_TRY(expr)                                                                  \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            sge::set_error(std::string("bag trie build: ") + hipGetErrorString(e_));    \
            return SGE_EDEVICE;                                                         \
        }                                                                               \
    } while (0)

unsigned blocks(uint64_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

__global__ void iota_kernel(uint32_t *__restrict__ a, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (uint32_t)i;
}

// out[i] = rows[perm[i]][c]
__global__ void column_kernel(const uint32_t *__restrict__ rows, const uint32_t *__restrict__ perm, uint64_t n,
                              uint32_t p, uint32_t c, uint32_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = rows[(uint64_t)perm[i] * p + c];
}

// dcol[i] = first column where sorted row i differs from row i - 1 (0 for row 0)
__global__ void first_diff_kernel(const uint32_t *__restrict__ S, uint64_t n, uint32_t p, uint32_t *__restrict__ dcol) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c = 0;
    if (i > 0)
        while (c < p && S[(uint64_t)c * n + i] == S[(uint64_t)c * n + i - 1]) ++c;
    dcol[i] = c;
}

// hist[c] += rows whose first differing column is c (c <= p)
__global__ void histogram_kernel(const uint32_t *__restrict__ dcol, uint64_t n, uint32_t p,
                                 unsigned long long *__restrict__ hist) {
    __shared__ unsigned int h[SGE_MAX_K + 1];
    if (threadIdx.x <= p) h[threadIdx.x] = 0;
    __syncthreads();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[dcol[i]], 1u);
    __syncthreads();
    if (threadIdx.x <= p && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

__global__ void flag_kernel(const uint32_t *__restrict__ dcol, uint64_t n, uint32_t j, uint32_t *__restrict__ f) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = dcol[i] <= j ? 1u : 0u;
}

// Depth j: node values, and (j < p - 1) each node's first child.
__global__ void scatter_kernel(const uint32_t *__restrict__ Sj, const uint32_t *__restrict__ f,
                               const uint32_t *__restrict__ pos, const uint32_t *__restrict__ pos_next, uint64_t n,
                               uint32_t vb_j, uint32_t vb_next, uint32_t j, uint32_t *__restrict__ val,
                               uint32_t *__restrict__ child) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !f[i]) return;
    val[vb_j + pos[i]] = Sj[i];
    if (pos_next) child[vb_j + j + pos[i]] = vb_next + pos_next[i];
}

__global__ void put_kernel(uint32_t *__restrict__ a, uint64_t at, uint32_t v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) a[at] = v;
}

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
    void release(void *q) { p.erase(std::remove(p.begin(), p.end(), q), p.end()); }
};

}  // namespace

namespace sge {

int trie_from_rows(const uint32_t *d_rows, uint64_t n, uint32_t p, uint32_t n_vertices, void *stream, BagTrie &t) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (p < 2 || p > SGE_MAX_K) {
        set_error("bag trie build: a bag holds 2..8 levels");
        return SGE_EINVAL;
    }
    if (n >= (1ull << 31) || (uint64_t)n * p >= 0xffffffffull - p) {
        set_error("bag trie build: relation too large for 32-bit trie indexes (" + std::to_string(n) + " tuples)");
        return SGE_ERANGE;
    }
    const int key_bits = n_vertices > 1 ? 32 - __builtin_clz(n_vertices - 1) : 1;
    Scratch s;
    uint32_t *perm, *perm2, *keys, *keys2, *S;
    TRIE_TRY(s.alloc(&perm, n));
    TRIE_TRY(s.alloc(&perm2, n));
    TRIE_TRY(s.alloc(&keys, n));
    TRIE_TRY(s.alloc(&keys2, n));
    TRIE_TRY(s.alloc(&S, n * p));
    // one temporary buffer serves every radix pass and every scan below
    size_t sort_bytes = 0, scan_bytes = 0;
    TRIE_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, keys2, perm, perm2, (int)n, 0, key_bits, st));
    TRIE_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, keys2, perm, (int)n, st));
    size_t tmp_bytes = std::max(sort_bytes, scan_bytes);
    unsigned char *tmp;
    TRIE_TRY(s.alloc(&tmp, tmp_bytes));
    // 1. lexicographic order: stable radix passes from the last column to the first
    if (n) iota_kernel<<<blocks(n), 256, 0, st>>>(perm, n);
    for (int c = (int)p - 1; c >= 0 && n; --c) {
        column_kernel<<<blocks(n), 256, 0, st>>>(d_rows, perm, n, p, (uint32_t)c, keys);
        TRIE_TRY(hipGetLastError());
        size_t b = tmp_bytes;
        TRIE_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, b, keys, keys2, perm, perm2, (int)n, 0, key_bits, st));
        std::swap(perm, perm2);
    }
    for (uint32_t c = 0; c < p && n; ++c) column_kernel<<<blocks(n), 256, 0, st>>>(d_rows, perm, n, p, c, S + (uint64_t)c * n);
    TRIE_TRY(hipGetLastError());
    // 2. node starts per depth
    uint32_t *dcol = keys, *f = keys2, *pos = perm, *pos_next = perm2;  // sort buffers reused
    unsigned long long *hist;
    TRIE_TRY(s.alloc(&hist, SGE_MAX_K + 1));
    TRIE_TRY(hipMemsetAsync(hist, 0, sizeof(unsigned long long) * (SGE_MAX_K + 1), st));
    if (n) {
        first_diff_kernel<<<blocks(n), 256, 0, st>>>(S, n, p, dcol);
        histogram_kernel<<<std::min<unsigned>(blocks(n), 1024), 256, 0, st>>>(dcol, n, p, hist);
    }
    TRIE_TRY(hipGetLastError());
    unsigned long long h[SGE_MAX_K + 1];
    TRIE_TRY(hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, st));
    TRIE_TRY(hipStreamSynchronize(st));
    if (h[p] != 0) {
        set_error("bag trie build: duplicate tuples in the bag relation");
        return SGE_EINVAL;
    }
    t.depth = p;
    t.tuples = n;
    uint64_t acc = 0;
    for (uint32_t j = 0; j < p; ++j) {
        acc += h[j];  // rows that start a node at depth j
        t.nodes[j] = acc;
    }
    t.vb[0] = 0;
    for (uint32_t j = 0; j < p; ++j) t.vb[j + 1] = t.vb[j] + (uint32_t)t.nodes[j];
    for (uint32_t j = p + 1; j <= SGE_MAX_K; ++j) t.vb[j] = t.vb[p];
    const uint64_t nval = t.vb[p], nchild = (uint64_t)t.vb[p - 1] + (p - 1);
    uint32_t *val, *child;
    TRIE_TRY(s.alloc(&val, nval));
    TRIE_TRY(s.alloc(&child, nchild));
    // 3. depths from the deepest up: node ids, values, first children, sentinels
    for (int j = (int)p - 1; j >= 0; --j) {
        if (n) {
            flag_kernel<<<blocks(n), 256, 0, st>>>(dcol, n, (uint32_t)j, f);
            TRIE_TRY(hipGetLastError());
            size_t b = tmp_bytes;
            TRIE_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, b, f, pos, (int)n, st));
            const bool inner = j < (int)p - 1;
            scatter_kernel<<<blocks(n), 256, 0, st>>>(S + (uint64_t)j * n, f, pos, inner ? pos_next : nullptr, n,
                                                      t.vb[j], inner ? t.vb[j + 1] : 0u, (uint32_t)j, val, child);
            TRIE_TRY(hipGetLastError());
        }
        if (j < (int)p - 1) put_kernel<<<1, 64, 0, st>>>(child, (uint64_t)t.vb[j] + j + t.nodes[j], t.vb[j + 1] + (uint32_t)t.nodes[j + 1]);
        TRIE_TRY(hipGetLastError());
        std::swap(pos, pos_next);
    }
    TRIE_TRY(hipStreamSynchronize(st));
    s.release(val);
    s.release(child);
    t.val = val;
    t.child = child;
    return SGE_OK;
}

void trie_free(BagTrie &t) {
    if (t.val) (void)hipFree(t.val);
    if (t.child) (void)hipFree(t.child);
    if (t.desc) (void)hipFree(t.desc);
    t.val = t.child = nullptr;
    t.desc = nullptr;
}

}  // namespace sge
