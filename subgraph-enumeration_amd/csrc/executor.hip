// MI355X (gfx950) WCOJ enumeration executor.
//
// Replaces the reference's join executors -- sharedJoin (matching/join.cpp:812),
// parSharedJoin WCOJ / exploration (matching/parallel.cpp:1154, 1488) and their
// set-intersection kernels (utility/compute_set_intersection.cpp) -- with one
// persistent kernel:
//
//  * one depth-first search per wavefront (64 lanes); the wave's per-level
//    candidate buffers live in a per-wave global scratch slab, its DFS cursor
//    (images, lengths, positions, list descriptors) in LDS;
//  * root tasks are directed data edges (CSR positions), handed out by a global
//    atomic queue (the OpenMP schedule(dynamic) root split of
//    parallel.cpp:1254-1299 / 1414-1458) from the high-degree end of the
//    degree-ordered CSR down: the heaviest subtasks one per item, the rest in
//    chunks; a root edge whose level-2 set can be large is split into slices;
//  * every level's candidate set is an intersection of sorted neighbour lists
//    clamped by symmetry-breaking bounds (setBeginPos/setEndPos, utils.cpp:581-612)
//    computed by lane-parallel binary search (windows staged in LDS when short),
//    the shortest list driving; survivors are compacted with a wave64 ballot +
//    popcount prefix;
//  * the last level is never enumerated: like TrieLevel::numResults
//    (relation.cpp:301-338) it is counted in closed form, and all siblings of the
//    second-to-last level are counted together with a flattened, load-balanced
//    lane mapping (per-sibling work prefix-summed in LDS, F sets staged in LDS as
//    bitmaps or sorted copies);
//  * IEP mode counts a trailing symmetric group with choosec (utils.cpp:80-94,
//    join.cpp:1532-1537), wrapping uint64 arithmetic included, and an
//    independent 2- or 3-vertex tail by inclusion-exclusion;
//  * exploration mode forms each level's parent intersection as soon as its last
//    parent is matched (attributesAfter, join.cpp:423-563) and only filters it later;
//  * a wavefront that has spent its work budget inside one item hands its open
//    candidate ranges to the next pass as continuation items;
//  * batched mode runs many plans in the same launch: queue item i is item i/P of
//    plan i%P, so all patterns drain the shared queue together.
// Counts are exact uint64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "common.h"

#define HIP_TRY(x)                                                                         \
    do {                                                                                   \
        hipError_t _e = (x);                                                               \
        if (_e != hipSuccess) {                                                            \
            sge::set_error(std::string("HIP error: ") + hipGetErrorString(_e) + " at " #x); \
            return SGE_EDEVICE;                                                            \
        }                                                                                  \
    } while (0)

namespace {

constexpr int WAVE = 64;
constexpr int MAXK = SGE_MAX_K;
#ifndef SGE_SIB_BLOCK
#define SGE_SIB_BLOCK 64
#endif
constexpr int SIB_BLOCK = SGE_SIB_BLOCK;  // (sibling, term) entries per flattened last-level batch
static_assert(SIB_BLOCK <= 64, "the probe-window entry mapping covers at most 64 entries per batch");
constexpr int MAX_TERMS = 7;    // inclusion-exclusion terms of a 3-level independent tail
constexpr int NBUF_MAX = 2 * MAXK + 1 + MAX_TERMS;  // scratch buffers per wave at most (see BufLayout)
constexpr uint32_t NODE_STEP = 1024;  // work-cap accounting granularity per wavefront (wave steps)
#ifdef SGE_PHASE_PROFILE
// per-item record of the phase-profile build: the phase cycles and events of
// each wavefront's longest item
constexpr int PH_ITEM_PHASES = 12;  // PH_OTHER .. PH_DONATE
constexpr int PH_ITEM_EVENTS = 8;   // PHE_*
constexpr int PH_ITEM_WORDS = PH_ITEM_PHASES + PH_ITEM_EVENTS + 4;
constexpr int PH_ITEM_WAVES = 8192;  // wavefronts recorded (root passes of 8,192-wave launches)
#endif
#ifndef SGE_FSTAGE
#define SGE_FSTAGE 395
#endif
constexpr int FSTAGE = SGE_FSTAGE;  // LDS words per wave for staging the terms' F sets (bitmaps or copies)
#ifndef SGE_SEARCH_STAGE
#define SGE_SEARCH_STAGE 256
#endif
constexpr uint32_t SEARCH_STAGE = SGE_SEARCH_STAGE;  // source elements staged in LDS per intersection chunk (in fst)
static_assert(SEARCH_STAGE <= (uint32_t)FSTAGE, "the intersection stage reuses the F staging words");
constexpr uint32_t NONE = 0xffffffffu;

// Bag relation as a level-order trie (built by trie.hip): depth j's values are
// val[vb[j] .. vb[j+1]), sorted within each parent's child list; the children
// of the depth-j value at absolute index a (j < depth-1) are
// val[child[a + j] .. child[a + j + 1]).  The flat form of the reference's
// nested TrieLevel (relation/relation.h:49-69).
struct TrieDesc {
    const uint32_t *val;
    const uint32_t *child;
    uint32_t vb[MAXK + 1];
    uint32_t depth;
};

// Device copy of a plan (levels masks) plus derived fields.
struct DPlan {
    uint32_t k, tail, mode, last, sib;  // last: level counted in closed form; sib = last - 1
    uint32_t parent[MAXK], lower[MAXK], upper[MAXK], check[MAXK];
    // bytes: rarely read, and every byte of the per-wave LDS copy is one
    // fewer for the F stage
    int8_t reuse[MAXK];
    uint8_t ready[MAXK];
    // last-level decomposition (relative to the sibling level s)
    uint32_t f_parent;     // parents of `last` other than s
    uint32_t f_ready;      // exploration: level after which F's intersection is formed
    uint32_t order[MAXK];  // level -> pattern vertex (enumeration output)
    // closed-form terms counted per sibling: kind 0/1 use one term (the first
    // closed-form level); kind 2 uses one term per non-empty subset B of the
    // independent tail (index = bitmask(B) - 1).  Masks are over levels below
    // the sibling level s; sflags bit0: s is a parent, bit1: lower bound from s,
    // bit2: upper bound from s, bit3: must differ from s's image.
    uint32_t tail_kind, nterms;
    uint32_t t_par[MAX_TERMS], t_lower[MAX_TERMS], t_upper[MAX_TERMS], t_check[MAX_TERMS], t_sflags[MAX_TERMS];
    // term j's F set starts from the candidate set of level (t_reuse >> 4j) & 15
    // (15: from the parents' rows alone), a level whose constraints it implies
    uint32_t t_reuse;
    // materialised bags: levels whose candidates come from a bag trie (the
    // trie list replaces the rows of the bag's parents); bag = the levels of
    // any bag (0 = none), bagm[b] = bag b's levels
    uint32_t bag;
    uint32_t bagm[SGE_MAX_BAGS];
    const TrieDesc *trie[SGE_MAX_BAGS];
    // root-task split: split[i] = first subtask of sampled root edge i (prefix
    // over ntask + 1 entries); nsub = total subtasks.  Subtask (i, ci) of
    // split[i+1] - split[i] processes slice ci of the level-2 candidates.
    const unsigned long long *split;
    unsigned long long nsub;
    // this plan's root tasks: CSR positions task_begin + i * stride, i < ntask
    unsigned long long task_begin, ntask;
};

struct PlanOut {
    unsigned long long count, inter, bytes, nodes, tasks;
    unsigned long long cycles;  // wavefront clock cycles spent on the plan's work items
    unsigned long long model;   // list bytes of the per-pair model: 4(|A'|+|B'|) + 4|out| + 8 per row
    unsigned long long pad;
};

struct KParams {
    const uint32_t *off, *nbr, *src;
    // skip index of the CSR: skip[i] = nbr[SKIP_G * i] (null: searches read the rows only)
    const uint32_t *skip;
    uint32_t n;
    uint64_t nnz;
    const DPlan *plans;
    uint32_t nplans;
    uint64_t stride;  // root tasks of plan p are CSR positions p.task_begin + i * stride, i < p.ntask
    uint32_t chunk;
    uint32_t head;    // the heaviest `head` subtasks of each plan are one queue item each
    uint64_t nitems;
    unsigned long long *queue;
    uint32_t *scratch;
    uint64_t wave_stride;  // u32 per wave
    uint32_t buf_stride;   // u32 per buffer
    // buffer indices in a wave's slab: level d's set at d; exploration's early
    // parent intersections at ex_base + level and F's at pre_f; the last-level
    // terms' F sets at term_base + term
    uint32_t ex_base, pre_f, term_base;
    // per-wave bitmap region in global memory for last-level F sets too large
    // for the LDS stage (all zero between count_last calls; null: not used)
    uint32_t *gbm;
    uint32_t gbm_words;
    uint32_t gbm_min;  // F sets staged as global bitmaps have at least this many elements
    PlanOut *out;
    uint32_t *emb;
    uint64_t emb_cap;
    unsigned long long *emb_n;
    // work donation between passes: a wave that spends more than `budget`
    // wave steps (one per node plus one per 64 candidates formed, and the
    // 64-lane iterations of each last-level count) inside one item writes the
    // untouched candidate ranges of its open levels as continuation items for
    // the next pass
    const uint32_t *in_items;          // this pass's input items (null: root pass)
    const unsigned long long *in_count;
    uint32_t *out_items;
    unsigned long long *out_count;
    uint64_t item_cap;                 // capacity of each item buffer
    uint32_t budget;                   // 0: never donate
    uint32_t split;                    // items per donated range (at most; >= 1)
    uint32_t probe_ratio;              // drive with the sibling's row while |row| <= ratio * |F| (F staged in LDS)
    uint32_t bitmap_ratio;             // the same when F is staged as a bitmap
    uint32_t own_items;                // 1: idle waves take this pass's own items (see live_claim)
    unsigned long long node_cap;       // stop after about this many wave steps of search work (0 = no cap)
    unsigned long long *node_ctr;      // wave steps spent so far, reported in blocks of >= NODE_STEP
};

// continuation item (64 bytes): plan, level l, absolute candidate range [b, e)
// of level l, images of levels 0..l-1, then the pass epoch that publishes it
// (written after the body, so a wave that takes the item within the same pass
// reads it only once it is whole)
constexpr int ITEM_BODY = 4 + MAXK;
constexpr int ITEM_EPOCH = ITEM_BODY;
constexpr int ITEM_WORDS = 16;
static_assert(ITEM_BODY < ITEM_WORDS, "the epoch word follows the body");

// control words of a pass, after the root / input queue counter at [0]:
// tickets for this pass's own output items, waves owning an item, the pass
// epoch (incremented by reset_pass_kernel before every pass), failed waits
// (an error the host reports), own-output items finished and waves exited
constexpr int Q_TAKEN = 4, Q_ACTIVE = 5, Q_EPOCH = 6, Q_ERR = 7, Q_DONE = 8, Q_EXITED = 9, Q_WORDS = 16;
// polls of a wave waiting for its ticket's item before it reports an error and
// exits (about a minute at the longest back-off), and iterations of the wait
// for a claimed item's publication (about a second)
constexpr uint32_t LIVE_POLLS = 1u << 20;
// WaveLds::qstate: the wave owns an item (counts in queue[Q_ACTIVE]); its input
// queue is exhausted; its last item was one of its pass's own
constexpr uint32_t QS_HOLD = 1, QS_INPUT_DONE = 2, QS_OWN = 4;
constexpr uint32_t ITEM_WAIT_SPINS = 1u << 24;

// list descriptor: [ptr + b, ptr + e)
typedef const __attribute__((address_space(1))) uint32_t guint;
__device__ __forceinline__ const guint *gl(const uint32_t *p) { return (const guint *)p; }
struct Src {
    const guint *ptr;
    uint32_t b, e;
};

// Per-wave LDS state.  Uniform values are written by every lane (same value),
// so no cross-lane ordering is needed for them.  Candidate sets are
// descriptors: a clamped CSR row range when a level has one source list (no
// copy; injectivity checks are then applied when the set is read), otherwise a
// range of the wave's scratch buffer for that level.
struct WaveLds {
    DPlan pl;
    uint32_t phi[MAXK];
    uint32_t pos[MAXK];  // next candidate position (absolute index into cand[d].ptr)
    uint32_t iend[MAXK]; // end of this wave's iteration range at each level
    uint32_t tpos[SGE_MAX_BAGS][MAXK];  // bag levels: absolute value index of phi[level] in each bag's trie
    unsigned long long accv[7];  // statistics counters (Acc), lane 0 only
    Src cand[MAXK];      // candidate set per level (whole set: reuse reads it)
    Src pre[MAXK + 1];   // exploration: parent intersections formed early ([MAXK] = F)
    Src srcs[MAXK + 2];  // sources of the intersection being formed
    uint32_t cur[MAXK + 2];  // per-source cursor of the running merge
    uint32_t chk[MAXK];  // check images of the intersection being formed
    // flattened sibling batch
    uint32_t pre_sum[SIB_BLOCK + 1];
    uint4 ent[SIB_BLOCK];      // per entry: first probe slot, driver start, target range [z, w)
    uint32_t cnt[SIB_BLOCK];   // per-sibling counts
    uint32_t flag[SIB_BLOCK];  // 1: driver is the sibling's CSR row (target F), 0: driver F (target row)
    uint32_t mark[WAVE];       // phase 2: entry starting at each lane's probe slot (index + 1)
    uint32_t qstate;           // work-queue state bits (QS_*), kept here rather than in registers
    uint32_t fst[FSTAGE];      // staged copies of small F sets (probed through LDS)
#ifdef SGE_PHASE_PROFILE
    unsigned long long ph[16];  // wavefront clock cycles per phase (PH_*), lane 0 only; [14], [15]: the
                                // longest F-set formation and last-level count call
    unsigned long long ph_t;    // clock at the last phase switch
    uint32_t ph_cur;            // current phase
    unsigned long long ph0[PH_ITEM_PHASES];  // ph[] at the current item's start
    unsigned long long ev[PH_ITEM_EVENTS];   // the current item's events (PHE_*)
#endif
};

// Phase profile (builds with -DSGE_PHASE_PROFILE -DSGE_FSTAGE=308, so that the
// counters fit the same LDS): every wavefront clock cycle is charged to the
// phase the wave is in (phases nest: a function enters its phase and restores
// the caller's on return), summed over all waves and read back with
// sge_phase_profile_read.
enum {
    PH_OTHER, PH_MAT, PH_EXPL, PH_TERMS, PH_STAGE, PH_P1, PH_P2, PH_P3, PH_QUEUE, PH_ROOT, PH_DFS, PH_DONATE,
    PH_ALL, PH_NUM
};
#ifdef SGE_PHASE_PROFILE
// [PH_NUM]: queue fetches, [+1] waves, [+2] longest wave; root passes only: [+3]
// summed wave cycles, [+4] longest wave, [+5] waves, [+6] the longest item
// (cycles >> 8 in the high word, its queue index in the low word), [+7] the
// same item's start (cycles >> 8 since the wave started) in the low word
__device__ unsigned long long g_phase[PH_NUM + 10];  // [+8], [+9]: longest F formation / count call
__device__ __forceinline__ uint32_t ph_switch(WaveLds &w, uint32_t i) {
    uint32_t old = 0;
    if (__lane_id() == 0) {
        const unsigned long long now = __builtin_amdgcn_s_memtime();
        old = w.ph_cur;
        w.ph[old] += now - w.ph_t;
        w.ph_t = now;
        w.ph_cur = i;
    }
    return old;
}
#define PH_ENTER(w, i) const uint32_t _ph_prev = ph_switch(w, i)
#define PH_SET(w, i) ph_switch(w, i)
#define PH_LEAVE(w) ph_switch(w, _ph_prev)
// per-item events: count calls, sibling hand-overs, DFS donations that wrote
// items, materialised sets, the longest count call and F formation (cycles),
// siblings counted, wave steps of the counts
enum { PHE_COUNTS, PHE_HANDOVERS, PHE_DONATIONS, PHE_MATS, PHE_COUNT_MAX, PHE_F_MAX, PHE_SIBS, PHE_STEPS };
#define PH_EV(w, k, v) do { if (__lane_id() == 0) (w).ev[k] += (v); } while (0)
#define PH_EVMAX(w, k, v) do { if (__lane_id() == 0) (w).ev[k] = max((w).ev[k], (unsigned long long)(v)); } while (0)
// wave index -> its longest item's record (PH_ITEM_WORDS: phase cycles,
// events, cycles, queue index, plan, start); g_item_key = the longest item of
// the pass as (cycles >> 8) << 32 | wave index
__device__ unsigned long long g_item[PH_ITEM_WAVES * PH_ITEM_WORDS];
__device__ unsigned long long g_item_key;
#else
#define PH_ENTER(w, i)
#define PH_SET(w, i)
#define PH_LEAVE(w)
#define PH_EV(w, k, v)
#define PH_EVMAX(w, k, v)
#endif

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// first index in [lo, hi) with a[i] >= key
template <typename PT>
__device__ __forceinline__ uint32_t lower_bound(PT a, uint32_t lo, uint32_t hi, uint32_t key) {
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}
// first index in [lo, hi) with a[i] > key
template <typename PT>
__device__ __forceinline__ uint32_t upper_bound(PT a, uint32_t lo, uint32_t hi, uint32_t key) {
    while (lo < hi) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] <= key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

template <typename PT>
__device__ __forceinline__ bool contains(PT a, uint32_t lo, uint32_t hi, uint32_t key) {
    uint32_t p = lower_bound(a, lo, hi, key);
    return p < hi && a[p] == key;
}

#ifndef SGE_PROBE_ARITY
#define SGE_PROBE_ARITY 1
#endif
// Membership by a search of fan-out 4 while the range is long: three
// independent pivot loads per step (issued together, one round trip) cut the
// range to a quarter, so a search of a row in L2 takes about half the
// dependent round trips of a binary search; the last few elements are searched
// as before.
template <typename PT>
__device__ __forceinline__ uint32_t lower_bound_wide(PT a, uint32_t lo, uint32_t hi, uint32_t key) {
    if (SGE_PROBE_ARITY == 4) {
        // invariant: a[lo - 1] < key (or lo is the start) and a[hi] >= key (or hi is the end)
        while (hi - lo > 16) {
            const uint32_t q = (hi - lo) >> 2, m1 = lo + q, m2 = m1 + q, m3 = m2 + q;
            const uint32_t v1 = a[m1], v2 = a[m2], v3 = a[m3];
            if (key <= v1) {
                hi = m1;
            } else if (key <= v2) {
                lo = m1 + 1;
                hi = m2;
            } else if (key <= v3) {
                lo = m2 + 1;
                hi = m3;
            } else {
                lo = m3 + 1;
            }
        }
    }
    return lower_bound(a, lo, hi, key);
}

template <typename PT>
__device__ __forceinline__ bool contains_wide(PT a, uint32_t lo, uint32_t hi, uint32_t key) {
    if (SGE_PROBE_ARITY == 4) {
        while (hi - lo > 16) {
            const uint32_t q = (hi - lo) >> 2, m1 = lo + q, m2 = m1 + q, m3 = m2 + q;
            const uint32_t v1 = a[m1], v2 = a[m2], v3 = a[m3];
            if (key < v1) {
                hi = m1;
            } else if (key < v2) {
                lo = m1;
                hi = m2;
            } else if (key < v3) {
                lo = m2;
                hi = m3;
            } else {
                lo = m3;
            }
        }
    }
    return contains(a, lo, hi, key);
}

// Clamp [b,e) of a sorted list to the open interval (lo, hi) (setBeginPos /
// setEndPos).  Every lane computes the same result.
template <typename PT>
__device__ __forceinline__ void clamp_range(PT a, uint32_t &b, uint32_t &e, bool has_lo, uint32_t lo,
                                            bool has_hi, uint32_t hi) {
    if (has_lo) b = upper_bound(a, b, e, lo);
    if (has_hi) e = lower_bound(a, b, e, hi);
    if (e < b) e = b;
}

// Searches of long CSR rows go through the skip index first: skip[i] =
// nbr[SKIP_G * i] for every SKIP_G-th CSR position, so a search of a row range
// [b, e) is a binary search over the range's block starts in the skip array
// (32x smaller, so the hub rows' blocks stay in L2) and then one within one
// block of SKIP_G elements (one or two cache lines), instead of a binary
// search whose middle steps each miss L2 on a 64K-element hub row.  Rows
// shorter than SKIP_MIN are searched directly.
constexpr uint32_t SKIP_G = 32;
constexpr uint32_t SKIP_MIN = 1024;

// first index in [b, e) of the CSR with nbr[i] >= key
__device__ __forceinline__ uint32_t row_lower_bound(const KParams &P, uint32_t b, uint32_t e, uint32_t key) {
    if (P.skip == nullptr || e - b < SKIP_MIN) return lower_bound(gl(P.nbr), b, e, key);
    const uint32_t ib = (b + SKIP_G - 1) / SKIP_G, ie = (e - 1) / SKIP_G;  // blocks starting in [b, e)
    const uint32_t j = lower_bound(gl(P.skip), ib, ie + 1, key);          // first block start >= key
    const uint32_t lo = (j == ib) ? b : SKIP_G * (j - 1) + 1;             // nbr[SKIP_G (j-1)] < key
    const uint32_t hi = (j <= ie) ? SKIP_G * j : e;                        // nbr[SKIP_G j] >= key
    return lower_bound(gl(P.nbr), lo, hi, key);
}

// first index in [b, e) of the CSR with nbr[i] > key
__device__ __forceinline__ uint32_t row_upper_bound(const KParams &P, uint32_t b, uint32_t e, uint32_t key) {
    return key == 0xffffffffu ? e : row_lower_bound(P, b, e, key + 1);
}

__device__ __forceinline__ bool row_contains(const KParams &P, uint32_t b, uint32_t e, uint32_t key) {
    const uint32_t p = row_lower_bound(P, b, e, key);
    return p < e && P.nbr[p] == key;
}

// clamp_range for a CSR row range
__device__ __forceinline__ void row_clamp(const KParams &P, uint32_t &b, uint32_t &e, bool has_lo, uint32_t lo,
                                          bool has_hi, uint32_t hi) {
    if (has_lo) b = row_upper_bound(P, b, e, lo);
    if (has_hi) e = row_lower_bound(P, b, e, hi);
    if (e < b) e = b;
}

// x / d for d < 2^16 in three 32-bit divisions (a 64-bit division expands to a
// long sequence whose temporaries raise the register pressure of the loops
// around it): the high word, then the remainder with each 16-bit half of the
// low word.
__device__ __forceinline__ uint64_t div_small(uint64_t x, uint32_t d) {
    const uint32_t hi = (uint32_t)(x >> 32), lo = (uint32_t)x;
    const uint32_t qh = hi / d, rh = hi - qh * d;
    const uint32_t t1 = (rh << 16) | (lo >> 16);
    const uint32_t q1 = t1 / d, r1 = t1 - q1 * d;
    const uint32_t t2 = (r1 << 16) | (lo & 0xffffu);
    const uint32_t q2 = t2 / d;
    return ((uint64_t)qh << 32) | ((uint64_t)q1 << 16) | q2;
}

__device__ __forceinline__ uint64_t choosec(uint64_t n, uint32_t k) {
    if ((uint64_t)k > n) return 0;
    if (k == 0) return 1;
    uint64_t nom = n;
    uint32_t den = 1;  // k! <= 8! < 2^16
    for (uint32_t i = 1; i < k; ++i) {
        nom *= (n - i);
        den *= (i + 1);
    }
    return div_small(nom, den);
}

// Start of piece q of [0, len) cut into np near-equal pieces (the first
// len % np pieces one longer); piece_start(len, np, np) == len.
__device__ __forceinline__ uint32_t piece_start(uint32_t len, uint32_t q, uint32_t np) {
    const uint32_t per = len / np, rem = len - per * np;
    return q * per + min(q, rem);
}

// Per-wave statistics counters (matches, pairwise intersections, algorithmic
// bytes of the executed strategy, search-tree nodes, root tasks, clock cycles,
// list bytes of the per-pair model), kept in the wave's LDS and updated by lane
// 0 only, so they occupy no registers.  The two byte counts differ only at the
// last level: the strategy charges 8 bytes per probe it runs, the model
// 4(|A'|+|B'|) per (sibling, term) pair whatever the probe direction.
enum { A_COUNT, A_INTER, A_BYTES, A_NODES, A_TASKS, A_CYCLES, A_MODEL, A_NUM };
static_assert(A_NUM == sizeof(((WaveLds *)0)->accv) / sizeof(unsigned long long), "one LDS slot per counter");
static_assert(A_NUM <= 8, "PlanOut holds the counters in Acc order");
struct Acc {
    unsigned long long *c;  // WaveLds::accv
    __device__ __forceinline__ void add(int k, unsigned long long v) const {
        if (__lane_id() == 0) c[k] += v;
    }
};

// Bounds from the images of the levels in the masks.
__device__ __forceinline__ void level_bounds(const WaveLds &w, uint32_t lmask, uint32_t umask, bool &has_lo,
                                             uint32_t &lo, bool &has_hi, uint32_t &hi) {
    has_lo = lmask != 0;
    has_hi = umask != 0;
    lo = 0;
    hi = 0xffffffffu;
    for (uint32_t m = lmask; m; m &= m - 1) lo = max(lo, w.phi[__builtin_ctz(m)]);
    for (uint32_t m = umask; m; m &= m - 1) hi = min(hi, w.phi[__builtin_ctz(m)]);
}

// Load the check images of `mask` into w.chk; returns their number.
__device__ __forceinline__ int load_checks(WaveLds &w, uint32_t mask) {
    int n = 0;
    for (uint32_t m = mask; m; m &= m - 1) w.chk[n++] = w.phi[__builtin_ctz(m)];
    return n;
}

__device__ __forceinline__ bool is_checked(const WaveLds &w, uint32_t mask, uint32_t x) {
    bool hit = false;
    for (uint32_t m = mask; m; m &= m - 1) hit |= (w.phi[__builtin_ctz(m)] == x);
    return hit;
}

// Intersect the ns sources in w.srcs (already clamped), drop the nchk check
// images in w.chk, and write survivors to out.  The shortest source drives; for
// every other source each lane binary-searches its element within a window that
// starts at the previous chunk's last position.  When that window is expected to
// be short (source and driver of similar density), its first SEARCH_STAGE
// elements are loaded into LDS with coalesced reads and searched there; elements
// beyond them are searched in global memory.  Survivors are compacted with a wave
// ballot + popcount prefix.  Returns the survivor count (uniform).
__device__ __forceinline__ uint32_t wave_intersect(const KParams &P, WaveLds &w, int ns, int nchk, uint32_t *out,
                                                   Acc &acc) {
    if (ns <= 0) return 0;
    int drv = 0;
    uint64_t bytes = 0;
    for (int i = 0; i < ns; ++i) {
        uint32_t len = w.srcs[i].e - w.srcs[i].b;
        bytes += 4ull * len;
        if (len < w.srcs[drv].e - w.srcs[drv].b) drv = i;
        w.cur[i] = w.srcs[i].b;
    }
    acc.add(A_BYTES, bytes);
    acc.add(A_MODEL, bytes);
    acc.add(A_INTER, (ns > 1) ? (ns - 1) : 0);
    const guint *D = w.srcs[drv].ptr;
    const uint32_t db = w.srcs[drv].b, de = w.srcs[drv].e;
    const int lane = lane_id();
    uint32_t total = 0;
    for (uint32_t base = db; base < de; base += WAVE) {
        uint32_t idx = base + lane;
        bool ok = idx < de;
        uint32_t x = ok ? D[idx] : 0xffffffffu;
        const uint32_t clen = min((uint32_t)WAVE, de - base);
        for (int i = 0; i < ns; ++i) {
            if (i == drv) continue;
            const guint *a = w.srcs[i].ptr;
            const bool row = a == gl(P.nbr);  // a CSR row: searched through the skip index
            const uint32_t e = w.srcs[i].e, c0 = w.cur[i];
            const uint32_t rem = e - c0;
            uint32_t p;
            bool in;
            if (rem > 0 && (uint64_t)rem * clen <= (uint64_t)SEARCH_STAGE * (de - base)) {
                const uint32_t n = min((uint32_t)SEARCH_STAGE, rem);
                for (uint32_t q = lane; q < n; q += WAVE) w.fst[q] = a[c0 + q];
                wave_sync();
                if (x <= w.fst[n - 1]) {
                    const uint32_t pos = lower_bound(w.fst, 0u, n, x);
                    p = c0 + pos;
                    in = w.fst[pos] == x;  // pos < n: fst[n - 1] >= x
                } else {
                    p = row ? row_lower_bound(P, c0 + n, e, x) : lower_bound_wide(a, c0 + n, e, x);
                    in = p < e && a[p] == x;
                }
                wave_sync();  // the next source is staged over these words
            } else {
                p = row ? row_lower_bound(P, c0, e, x) : lower_bound_wide(a, c0, e, x);
                in = p < e && a[p] == x;
            }
            ok = ok && in;
            w.cur[i] = __shfl(p, WAVE - 1);
        }
        for (int c = 0; c < nchk; ++c) ok = ok && (x != w.chk[c]);
        uint64_t mask = __ballot(ok);
        uint32_t pre = __popcll(mask & ((1ull << lane) - 1ull));
        if (ok) out[total + pre] = x;
        total += __popcll(mask);
    }
    acc.add(A_BYTES, 4ull * total);
    acc.add(A_MODEL, 4ull * total);
    wave_sync();  // other lanes read the survivors next
    return total;
}

__device__ __forceinline__ uint32_t *buf_ptr(const KParams &P, uint32_t *wave_base, int i) {
    return wave_base + (uint64_t)i * P.buf_stride;
}

// Trie list of level d in bag b (unclamped): the bag's first level reads the
// root values, a later one the children of the bag's previous level's node.
__device__ __forceinline__ Src trie_list(const WaveLds &w, int b, int d) {
    const TrieDesc *T = w.pl.trie[b];
    const uint32_t below = w.pl.bagm[b] & ((1u << d) - 1u);
    Src s;
    s.ptr = gl(T->val);
    if (below == 0) {
        s.b = T->vb[0];
        s.e = T->vb[1];
    } else {
        const uint32_t j = __popc(below);                 // depth of level d
        const uint32_t a = w.tpos[b][31 - __clz(below)];  // parent node (depth j - 1)
        s.b = T->child[a + j - 1];
        s.e = T->child[a + j];
    }
    return s;
}

// Descend bag b's trie at level d after phi[d] = x: record x's value index
// (`idx` is its index in cand[d], NONE when unknown; when cand[d] is this trie
// list itself that index is the value index).  Returns false when x is not in
// the trie list, which only happens for the root levels (every other bag
// level's candidates are a subset of its trie lists).
__device__ __forceinline__ bool trie_enter(WaveLds &w, int b, int d, uint32_t x, uint32_t idx) {
    const guint *val = gl(w.pl.trie[b]->val);
    if (idx != NONE && w.cand[d].ptr == val) {
        w.tpos[b][d] = idx;
        return true;
    }
    const Src s = trie_list(w, b, d);
    const uint32_t p = lower_bound(s.ptr, s.b, s.e, x);
    w.tpos[b][d] = p;
    return p < s.e && val[p] == x;
}

// Descend every bag trie that holds level d and a later level (all = every bag
// holding d, for root levels); false when x is missing from one of them.  Kept
// out of line: only bag plans call it, and inlined at its five call sites it
// cost bagless plans 3% (C2 six-vertex batch 432 -> 446 ms) through register
// allocation in the DFS loop.
__device__ __noinline__ bool trie_enter_all(WaveLds &w, int d, uint32_t x, uint32_t idx, bool all) {
    bool ok = true;
    for (int b = 0; b < SGE_MAX_BAGS; ++b) {
        const uint32_t bm = w.pl.bagm[b];
        if (((bm >> d) & 1u) && (all || (bm >> (d + 1))) && !trie_enter(w, b, d, x, idx)) ok = false;
    }
    return ok;
}

// Sources of the parent intersection of level d into w.srcs, bounds applied.
__device__ int gather_sources(const KParams &P, WaveLds &w, int d, uint32_t pmask, int reuse, bool use_pre,
                              bool has_lo, uint32_t lo, bool has_hi, uint32_t hi, Acc &acc) {
    int ns = 0;
    if (use_pre) {
        Src s = w.pre[d];
        clamp_range(s.ptr, s.b, s.e, has_lo, lo, has_hi, hi);
        w.srcs[ns++] = s;
        return ns;
    }
    if (reuse >= 0) {
        Src s = w.cand[reuse];
        pmask &= ~w.pl.parent[reuse];
        clamp_range(s.ptr, s.b, s.e, has_lo, lo, has_hi, hi);
        w.srcs[ns++] = s;
    }
    for (int b = 0; ((w.pl.bag >> d) & 1u) && b < SGE_MAX_BAGS; ++b) {
        const uint32_t bm = w.pl.bagm[b];
        if (!((bm >> d) & 1u)) continue;
        // the trie list already satisfies every edge, bound and check among
        // the bag's levels: the rows of the bag's parents are not read
        Src s = trie_list(w, b, d);
        if (bm & ((1u << d) - 1u)) pmask &= ~bm;
        clamp_range(s.ptr, s.b, s.e, has_lo, lo, has_hi, hi);
        w.srcs[ns++] = s;
        acc.add(A_BYTES, 8);
        acc.add(A_MODEL, 8);
    }
    for (uint32_t m = pmask; m; m &= m - 1) {
        uint32_t v = w.phi[__builtin_ctz(m)];
        Src s;
        s.ptr = gl(P.nbr);
        s.b = P.off[v];
        s.e = P.off[v + 1];
        row_clamp(P, s.b, s.e, has_lo, lo, has_hi, hi);
        w.srcs[ns++] = s;
        acc.add(A_BYTES, 8);
        acc.add(A_MODEL, 8);
    }
    return ns;
}

// Form a set from w.srcs: alias the single source when allowed, else
// intersect into the scratch buffer `dst`.
__device__ Src form_set(const KParams &P, WaveLds &w, int ns, int nchk, bool may_alias, uint32_t *dst, Acc &acc) {
    if (ns <= 0) {
        Src z;
        z.ptr = gl(dst);
        z.b = z.e = 0;
        return z;
    }
    if (ns == 1 && may_alias) {
        acc.add(A_BYTES, 4ull * (w.srcs[0].e - w.srcs[0].b));
        acc.add(A_MODEL, 4ull * (w.srcs[0].e - w.srcs[0].b));
        return w.srcs[0];
    }
    Src r;
    r.ptr = gl(dst);
    r.b = 0;
    r.e = wave_intersect(P, w, ns, nchk, dst, acc);
    return r;
}

// Wave steps of intersecting w.srcs[0..ns): one per 64-element round of the
// shortest source for every other source (the searches wave_intersect makes).
__device__ __forceinline__ uint32_t intersect_steps(const WaveLds &w, int ns) {
    if (ns < 2) return 0;
    uint32_t m = 0xffffffffu;
    for (int i = 0; i < ns; ++i) m = min(m, w.srcs[i].e - w.srcs[i].b);
    return ((m + WAVE - 1) / WAVE) * (uint32_t)(ns - 1);
}

// Candidate set of level d (checks of level d applied now when the set is
// materialised, or when it is read if it aliases a CSR row).  Returns the wave
// steps of its intersection (intersect_steps).
__device__ __forceinline__ uint32_t materialize(const KParams &P, WaveLds &w, uint32_t *wb, int d, Acc &acc) {
    PH_ENTER(w, PH_MAT);
    PH_EV(w, PHE_MATS, 1);
    const DPlan &pl = w.pl;
    bool has_lo, has_hi;
    uint32_t lo, hi;
    level_bounds(w, pl.lower[d], pl.upper[d], has_lo, lo, has_hi, hi);
    // exploration: the parent intersection formed early (bag levels: formed now
    // from the trie list and the non-bag parents' rows, like the other modes)
    bool use_pre = pl.mode == SGE_MODE_EXPLORE && !((pl.bag >> d) & 1u);
    int ns = gather_sources(P, w, d, pl.parent[d], use_pre ? -1 : pl.reuse[d], use_pre, has_lo, lo, has_hi, hi, acc);
    int nchk = load_checks(w, pl.check[d]);
    const uint32_t st = intersect_steps(w, ns);
    const Src c = form_set(P, w, ns, nchk, true, buf_ptr(P, wb, d), acc);
    w.cand[d] = c;
    w.pos[d] = c.b;
    w.iend[d] = c.e;
    acc.add(A_NODES, 1);
    PH_LEAVE(w);
    return st;
}

// Exploration: form the parent intersection of every level whose last parent
// is level d (and of the last level's sibling-independent part F), unclamped.
__device__ __forceinline__ void explore_after(const KParams &P, WaveLds &w, uint32_t *wb, int d, Acc &acc) {
    PH_ENTER(w, PH_EXPL);
    const DPlan &pl = w.pl;
    for (uint32_t j = d + 1; j <= pl.sib; ++j) {
        // a bag level's candidates come from its trie list, which exists only
        // once the previous bag level is matched: it is formed when it is reached
        if (pl.ready[j] != (uint32_t)d || j < 2 || ((pl.bag >> j) & 1u)) continue;
        int ns = gather_sources(P, w, j, pl.parent[j], -1, false, false, 0, false, 0, acc);
        w.pre[j] = form_set(P, w, ns, 0, true, buf_ptr(P, wb, P.ex_base + j), acc);
    }
    if (pl.f_parent && pl.f_ready == (uint32_t)d) {
        int ns = gather_sources(P, w, pl.last, pl.f_parent, -1, false, false, 0, false, 0, acc);
        w.pre[MAXK] = form_set(P, w, ns, 0, true, buf_ptr(P, wb, P.pre_f), acc);
    }
    PH_LEAVE(w);
}

// Per-term setup kept in LDS for a sibling batch.
struct TermLds {
    Src F;            // sibling-independent part (bounds/checks below s applied)
    uint32_t lo, hi;  // fixed bounds (has_lo / has_hi in flags)
    uint32_t flags;   // bit0 has_lo, bit1 has_hi, bit2 universe, bit3 staged copy, bit4 staged bitmap
    uint32_t sflags;  // as DPlan::t_sflags
    uint32_t chk;     // check mask below s (needed in the universe case)
    uint32_t st_off;  // word offset of the staged copy / bitmap in WaveLds::fst
    uint32_t bm_lo;   // bitmap: id of bit 0 (= smallest element)
    uint32_t bm_hi;   // bitmap: largest element
};

// Membership of x in term t's F between F positions [tlo, thi) (a sibling's row
// drove x, so its bounds already hold): one bit test for a bitmap, else a
// binary search in the LDS copy or in F itself.
__device__ __forceinline__ bool term_has(const KParams &P, const WaveLds &w, const TermLds &t, const uint32_t *gb,
                                         uint32_t x, uint32_t tlo, uint32_t thi) {
    if (t.flags & 48u) {
        if (x < t.bm_lo || x > t.bm_hi) return false;
        uint32_t r = x - t.bm_lo;
        // a global bitmap is read at L2 (agent scope), where its atomics landed
        const uint32_t word = (t.flags & 16u) ? w.fst[t.st_off + (r >> 5)]
                                              : __hip_atomic_load(gb + t.st_off + (r >> 5), __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT);
        return (word >> (r & 31u)) & 1u;
    }
    if (t.flags & 8u) return contains(&w.fst[t.st_off], tlo - t.F.b, thi - t.F.b, x);
    if (t.F.ptr == gl(P.nbr)) return row_contains(P, tlo, thi, x);  // F is a CSR row
    return contains_wide(t.F.ptr, tlo, thi, x);
}

// Inclusion-exclusion over set partitions of the tail (phase 3 of count_last):
// injective assignments = sum over partitions pi of prod_{B in pi} (-1)^{|B|-1}
// (|B|-1)! |cap_{t in B} C_t|; r = 2: ab - |ab|; r = 3: abc - |ab|c - |ac|b -
// |bc|a + 2|abc|.  u64 arithmetic is exact modulo 2^64 and the true count is
// below 2^64.

// Work cap: report `since` wave steps once at least NODE_STEP have gathered;
// true when the call's total has reached the cap (every wave then stops at its
// next check, so a capped call ends with a normal kernel exit).
template <bool CAP>
__device__ __forceinline__ bool cap_reached(const KParams &P, uint32_t &since) {
    if (!CAP || since < NODE_STEP) return false;
    unsigned long long tot = 0;
    if (lane_id() == 0) tot = atomicAdd(P.node_ctr, (unsigned long long)since);
    tot = __shfl(tot, 0);
    since = 0;
    return tot >= P.node_cap;
}

// Probe direction for an F staged as a global bitmap: the sibling's row drives
// while |row| <= SGE_GBM_RATIO |F| (a probe is then one L2 read, against a
// binary search of the row per F element the other way).
#ifndef SGE_GBM_RATIO
#define SGE_GBM_RATIO 2
#endif

// Whether a global bitmap of a large F (len elements) pays for its build and
// clear (about 2 len / 64 wave steps): the siblings' rows short enough to
// probe F (|row| <= 2 len) save about log2(len) loads per element.  Estimated
// from the first 64 siblings' degrees, scaled to all nsib.
__device__ __forceinline__ bool gbm_pays(const KParams &P, const WaveLds &w, int s, uint32_t len) {
    const int lane = lane_id();
    const uint32_t b = (s == 1) ? 0u : w.pos[s];
    const uint32_t nsib = (s == 1) ? 1u : w.iend[s] - b;
    const uint32_t nsamp = min(nsib, (uint32_t)WAVE);
    uint64_t probes = 0;
    if ((uint32_t)lane < nsamp) {
        const uint32_t c = (s == 1) ? w.phi[1] : w.cand[s].ptr[b + lane];
        const uint32_t d = P.off[c + 1] - P.off[c];
        if ((uint64_t)d <= (uint64_t)SGE_GBM_RATIO * len) probes = d;
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) probes += __shfl_xor(probes, off);
    // probes * nsib / nsamp: nsamp is 64 when nsib >= 64, else nsib
    const uint64_t est = nsib >= (uint32_t)WAVE ? (probes * nsib) >> 6 : probes;
    return est * (uint64_t)(32 - __clz(len)) >= 4ull * len;
}

// Count the closed-form level(s) for every sibling in cand[sib] (or the single
// sibling phi[1] when sib == 1).  Each term's F = the candidates that do not
// depend on the sibling, with every bound/check from levels below the sibling
// applied; per (sibling, term) the count is |F ∩ N(c) ∩ range(c)| (or a range
// size), computed with the flattened probe loop.
// The wave steps it spends (64-lane iterations of the entry setup and the
// flattened probes), the unit of the donation budget and the work cap, are
// added to `spent` and `since`; the cap is checked after every sibling batch.
// A long count hands siblings to other wavefronts (when donation is on): after
// the entry set-up of a batch, the time left in the call is estimated -- at the
// first batch from its probes (rounds x SIB_CYCLES_PER_STEP clock cycles), at
// later ones from the clock cycles the batches so far took per sibling, times
// the siblings left; at SIB_DONATE_AFTER donation budgets of steps' worth of
// cycles or more, the call is cut into pieces of about half that (at most
// `split`), this wave keeps the first piece's share of siblings (at least one)
// and writes the rest as continuation items of level sib, once per call (each
// receiver re-forms the F sets once and may cut its piece again: handing over
// half of a hub's siblings at every budget was 4.7x slower on the R-MAT 20
// ranges).  The clock matters on hub rows: a 64-lane probe round through a
// 64K-id row costs about 8,000 cycles of dependent loads, eight times the
// rounds of an L2-resident graph, and an estimate in rounds let one count of
// 8,784 siblings hold a wavefront for 62 M cycles (29 ms) while the rest of the
// pass had finished.  Returns true when the cap stopped it (its count is then
// partial).
constexpr uint32_t SIB_DONATE_AFTER = 16;
constexpr float SIB_CYCLES_PER_STEP = 1024.0f;

// Item words are written and read with device-coherent (agent-scope relaxed
// atomic) stores and loads: a wave on another XCD may take an item while the
// pass runs, and per-XCD L2s are not coherent for ordinary accesses.  Fences
// at agent scope would write back or invalidate a whole L2 instead.
__device__ __forceinline__ void item_store(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t item_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ctl_load(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ctl_add(unsigned long long *p, unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// every memory operation this wave issued has completed (is device-visible)
__device__ __forceinline__ void mem_drain() { __builtin_amdgcn_s_waitcnt(0); }

// Publish output slots [base, base + n) whose bodies this wave has written:
// once the body stores have completed, write the pass epoch into each.
__device__ __forceinline__ void publish_items(const KParams &P, uint64_t base, uint64_t n) {
    mem_drain();
    const uint32_t epoch = (uint32_t)ctl_load(P.queue + Q_EPOCH);
    for (uint64_t q = lane_id(); q < n; q += WAVE) item_store(P.out_items + (base + q) * ITEM_WORDS + ITEM_EPOCH, epoch);
}

// Write [b, e) of level l's candidates as np continuation items of near-equal
// length (plan, level, range, images of levels < l); false (nothing written)
// when the output buffer is full.
__device__ __forceinline__ bool give_range(const KParams &P, const WaveLds &w, uint32_t plan, int l, uint32_t b,
                                           uint32_t e, uint32_t np) {
    const int lane = lane_id();
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(P.out_count, (unsigned long long)np);
    base = __shfl(base, 0);
    if (base + np > P.item_cap) {
        for (uint64_t slot = base; slot < P.item_cap && slot < base + np; ++slot)
            if (lane == 0) item_store(P.out_items + slot * ITEM_WORDS, NONE);
        if (base < P.item_cap) publish_items(P, base, min<uint64_t>(np, P.item_cap - base));
        return false;
    }
    const uint32_t len = e - b;
    for (uint32_t q = 0; q < np; ++q) {
        uint32_t *it = P.out_items + (base + q) * ITEM_WORDS;
        if (lane < ITEM_BODY) {
            uint32_t v;
            if (lane == 0) v = plan;
            else if (lane == 1) v = (uint32_t)l;
            else if (lane == 2) v = b + piece_start(len, q, np);
            else if (lane == 3) v = b + piece_start(len, q + 1, np);
            else v = (lane - 4 < l) ? w.phi[lane - 4] : 0u;
            item_store(it + lane, v);
        }
    }
    publish_items(P, base, np);
    return true;
}

template <bool CAP>
__device__ __forceinline__ bool count_last(const KParams &P, WaveLds &w, TermLds *terms, uint32_t *wb, uint32_t *gb,
                                           Acc &acc, uint32_t &spent, uint32_t &since, uint32_t plan) {
    PH_ENTER(w, PH_TERMS);
#ifdef SGE_PHASE_PROFILE
    const unsigned long long ph_t0 = __builtin_amdgcn_s_memtime();
    PH_EV(w, PHE_COUNTS, 1);
#endif
    const DPlan &pl = w.pl;
    const int s = pl.sib;
    const int lane = lane_id();
    const uint32_t K = pl.nterms;
    // wave steps of the F-set intersections: on hub rows a search whose every
    // sibling is pruned can spend most of its time here, and an item that
    // left them uncounted never reached its donation budget
    uint32_t fsteps = 0;
    // ---- per-term fixed parts
    for (uint32_t j = 0; j < K; ++j) {
        bool has_lo, has_hi;
        uint32_t lo, hi;
        level_bounds(w, pl.t_lower[j], pl.t_upper[j], has_lo, lo, has_hi, hi);
        TermLds t;
        t.lo = lo;
        t.hi = hi;
        t.sflags = pl.t_sflags[j];
        t.chk = pl.t_check[j];
        t.flags = (has_lo ? 1u : 0u) | (has_hi ? 2u : 0u);
        t.F.ptr = nullptr;
        t.F.b = t.F.e = 0;
        // an inclusion-exclusion term over several tail members has exactly
        // the union of their constraints, so its F is the intersection of
        // two smaller terms' F sets (built earlier: their masks are smaller)
        const uint32_t B = j + 1, b1 = B & (~B + 1u);
        const bool from_terms = pl.tail_kind == 2 && B != b1 && pl.t_par[j] != 0 &&
                                !((terms[b1 - 1].flags | terms[(B ^ b1) - 1].flags) & 4u);
        if (pl.t_par[j] == 0) {
            t.flags |= 4u;  // universe: only the sibling's row constrains it
        } else if (from_terms) {
            w.srcs[0] = terms[b1 - 1].F;
            w.srcs[1] = terms[(B ^ b1) - 1].F;
            fsteps += intersect_steps(w, 2);
            t.F = form_set(P, w, 2, 0, false, buf_ptr(P, wb, P.term_base + j), acc);
            acc.add(A_NODES, 1);
        } else {
            int ns;
            if (pl.mode == SGE_MODE_EXPLORE && K == 1) {
                Src sp = w.pre[MAXK];
                clamp_range(sp.ptr, sp.b, sp.e, has_lo, lo, has_hi, hi);
                w.srcs[0] = sp;
                ns = 1;
            } else {
                const uint32_t tr = (pl.t_reuse >> (4u * j)) & 15u;
                ns = gather_sources(P, w, pl.last, pl.t_par[j], tr == 15u ? -1 : (int)tr, false, has_lo, lo, has_hi,
                                    hi, acc);
            }
            int nchk = load_checks(w, pl.t_check[j]);
            fsteps += intersect_steps(w, ns);
            t.F = form_set(P, w, ns, nchk, nchk == 0, buf_ptr(P, wb, P.term_base + j), acc);
            acc.add(A_NODES, 1);
        }
        terms[j] = t;
    }
    wave_sync();
#ifdef SGE_PHASE_PROFILE
    if (lane_id() == 0) w.ph[14] = max(w.ph[14], __builtin_amdgcn_s_memtime() - ph_t0);
    PH_EVMAX(w, PHE_F_MAX, __builtin_amdgcn_s_memtime() - ph_t0);
#endif
    PH_SET(w, PH_STAGE);
    // ---- stage F sets in LDS so per-sibling probes hit LDS: a bitmap over
    // [min F, max F] when it fits, else a sorted copy when that fits, else (a
    // large F that the siblings' rows will probe) a bitmap in the wave's
    // global region
    uint32_t used_g = 0;
    {
        uint32_t used = 0;
        for (uint32_t j = 0; j < K; ++j) {
            TermLds t = terms[j];
            const uint32_t len = t.F.e - t.F.b;
            if ((t.flags & 4u) || len == 0) continue;
            const uint32_t fmin = t.F.ptr[t.F.b], fmax = t.F.ptr[t.F.e - 1];
            const uint64_t words = ((uint64_t)(fmax - fmin) >> 5) + 1;
            if (used + words <= (uint64_t)FSTAGE) {
                for (uint32_t q = lane; q < (uint32_t)words; q += WAVE) w.fst[used + q] = 0u;
                wave_sync();
                for (uint32_t q = lane; q < len; q += WAVE) {
                    uint32_t r = t.F.ptr[t.F.b + q] - fmin;
                    atomicOr(&w.fst[used + (r >> 5)], 1u << (r & 31u));
                }
                t.flags |= 16u;
                t.st_off = used;
                t.bm_lo = fmin;
                t.bm_hi = fmax;
                used += (uint32_t)words;
            } else if (used + len <= (uint32_t)FSTAGE) {
                for (uint32_t q = lane; q < len; q += WAVE) w.fst[used + q] = t.F.ptr[t.F.b + q];
                t.flags |= 8u;
                t.st_off = used;
                used += len;
            } else if (gb && (t.sflags & 1u) && len >= P.gbm_min && used_g + words <= (uint64_t)P.gbm_words &&
                       gbm_pays(P, w, s, len)) {
                // too large for LDS: a bitmap in the wave's global region, so a
                // probe is one L2 read instead of a binary search of F
                for (uint32_t q = lane; q < len; q += WAVE) {
                    const uint32_t r = t.F.ptr[t.F.b + q] - fmin;
                    atomicOr(gb + used_g + (r >> 5), 1u << (r & 31u));
                }
                t.flags |= 32u;
                t.st_off = used_g;
                t.bm_lo = fmin;
                t.bm_hi = fmax;
                used_g += (uint32_t)words;
            } else {
                continue;
            }
            terms[j] = t;
        }
        // the bits are set before any probe: the region is this wave's own, so
        // workgroup scope orders it (an agent-scope fence would write back and
        // invalidate the XCD's whole L2)
        if (used_g) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        wave_sync();
    }
    PH_SET(w, PH_P1);
    // ---- siblings
    Src sibs;
    uint32_t nsib;
    const uint32_t one = w.phi[s];
    if (s == 1) {
        sibs.ptr = nullptr;
        sibs.b = sibs.e = 0;
        nsib = 1;
    } else {
        sibs = w.cand[s];
        sibs.b = w.pos[s];
        sibs.e = w.iend[s];
        nsib = sibs.e - sibs.b;
    }
    const uint32_t schk_mask = pl.check[s];
    const uint32_t kind = pl.tail_kind;
    const bool per_entry = (kind != 0) || K > 1;  // counts needed per (sibling, term)
    const uint32_t per_block = SIB_BLOCK / K;     // siblings per batch
    unsigned long long local = 0;                 // per-lane partial
    uint32_t steps = 1 + fsteps;
    bool stopped = false, gave = false;
    const unsigned long long t_batches = __builtin_amdgcn_s_memtime();  // start of the sibling batches
    for (uint32_t sb = 0; sb < nsib; sb += per_block) {
        if (CAP) {
            spent += steps;
            since += steps;
            steps = 0;
            if (cap_reached<CAP>(P, since)) {
                stopped = true;
                break;
            }
        }
        uint32_t nb = min(per_block, nsib - sb);
        uint32_t ne = nb * K;  // entries: sibling i, term j -> i*K + j
        // phase 1: per-entry ranges and work (lane per entry)
        uint32_t carry = 0;
        uint32_t my_mw = 0;
        for (uint32_t r0 = 0; r0 < ne; r0 += WAVE) {
            uint32_t ei = r0 + lane;
            uint32_t work = 0, dstart = 0, tl = 0, th = 0;
            if (ei < ne) {
                const uint32_t i = ei / K, j = ei - i * K;
                const TermLds &t = terms[j];
                uint32_t c = sibs.ptr ? sibs.ptr[sibs.b + sb + i] : one;
                uint32_t cntv = 0;
                uint32_t flag = 0;
                // a sibling aliasing a CSR row may still equal a checked image
                bool valid = (s == 1) || !is_checked(w, schk_mask, c);
                if (valid) {
                    bool clo = t.flags & 1u, chi = t.flags & 2u;
                    uint32_t clo_v = t.lo, chi_v = t.hi;
                    const bool s_lower = t.sflags & 2u, s_upper = t.sflags & 4u;
                    if (s_lower) { clo_v = clo ? max(clo_v, c) : c; clo = true; }
                    if (s_upper) { chi_v = chi ? min(chi_v, c) : c; chi = true; }
                    if (t.sflags & 1u) {  // the sibling is a parent: intersect with its row
                        uint32_t rb = P.off[c], re = P.off[c + 1];
                        row_clamp(P, rb, re, clo, clo_v, chi, chi_v);
                        if (t.flags & 4u) {
                            cntv = re - rb;
                            for (uint32_t m = t.chk; m; m &= m - 1)
                                if (row_contains(P, rb, re, w.phi[__builtin_ctz(m)])) --cntv;
                        } else {
                            uint32_t fb = t.F.b, fe = t.F.e;
                            clamp_range(t.F.ptr, fb, fe, s_lower, c, s_upper, c);
                            // probes into a staged F cost LDS reads (one for a bitmap),
                            // probes into a row cost a global binary search
                            const uint64_t ratio =
                                (t.flags & 32u) ? (uint32_t)SGE_GBM_RATIO
                                                : ((t.flags & 16u) ? P.bitmap_ratio : ((t.flags & 8u) ? P.probe_ratio : 1u));
                            if ((uint64_t)(re - rb) <= ratio * (uint64_t)(fe - fb)) {
                                flag = 1;
                                dstart = rb;
                                tl = fb;
                                th = fe;
                                work = re - rb;
                            } else {
                                flag = 0;
                                dstart = fb;
                                tl = rb;
                                th = re;
                                work = fe - fb;
                            }
                        }
                    } else {
                        uint32_t fb = t.F.b, fe = t.F.e;
                        clamp_range(t.F.ptr, fb, fe, s_lower, c, s_upper, c);
                        cntv = fe - fb;
                        if ((t.sflags & 8u) && contains(t.F.ptr, fb, fe, c)) --cntv;
                    }
                }
                w.cnt[ei] = cntv;
                w.flag[ei] = flag | (j << 1);
            }
            // inclusive scan of work within the wave
            uint32_t incl = work;
            for (int off = 1; off < WAVE; off <<= 1) {
                uint32_t tt = __shfl_up(incl, off);
                if (lane >= off) incl += tt;
            }
            if (ei < ne) {
                w.pre_sum[ei + 1] = carry + incl;
                w.ent[ei] = make_uint4(carry + incl - work, dstart, tl, th);
            }
            carry += __shfl(incl, WAVE - 1);
            // model bytes: both clamped lists of every probed (sibling, term)
            // pair (entry ei = lane: ne <= SIB_BLOCK <= 64, one iteration)
            my_mw = work + (th - tl);
        }
        if (lane == 0) w.pre_sum[0] = 0;
        wave_sync();
        if (!gave && s >= 2 && P.budget && P.out_items && nsib - sb >= 2) {
            // the cycles this call has left: from this batch's probes at the
            // first batch, from the clock at later ones
            const uint32_t rem = nsib - sb;
            float est;
            if (sb == 0) {
                const uint32_t per_sib = max(1u, w.pre_sum[ne] / nb);
                est = (float)per_sib * (float)rem * (SIB_CYCLES_PER_STEP / (float)WAVE);
            } else {
                const float took = (float)(__builtin_amdgcn_s_memtime() - t_batches);
                est = took * (float)rem / (float)sb;
            }
            const float unit = (float)SIB_DONATE_AFTER * (float)P.budget * SIB_CYCLES_PER_STEP;
            if (est >= unit) {
                gave = true;
                // pieces of about unit / 2 cycles; this wave keeps as much
                const float want = est / (0.5f * unit);
                const uint32_t cap = min(P.split, rem - 1);
                const uint32_t np = want >= (float)cap ? cap : max(1u, (uint32_t)want);
                const uint32_t keep = min(nb, max(1u, rem / (np + 1)));
                if (give_range(P, w, plan, s, sibs.b + sb + keep, sibs.e, np)) {
                    PH_EV(w, PHE_HANDOVERS, 1);
                    // drop the given siblings' entries from this batch
                    nb = keep;
                    ne = keep * K;
                    nsib = sb + keep;
                }
            }
        }
        {
            uint32_t mw = (uint32_t)lane < ne ? my_mw : 0u;
            for (int off = WAVE / 2; off > 0; off >>= 1) mw += __shfl_xor(mw, off);
            acc.add(A_MODEL, 4ull * mw);
        }
        PH_SET(w, PH_P2);
        const uint32_t W = w.pre_sum[ne];
        steps += (ne + W + WAVE - 1) / WAVE;
        acc.add(A_INTER, ne);
        acc.add(A_BYTES, 8ull * nb);
        acc.add(A_MODEL, 8ull * nb);
        // phase 2: flattened probes, a window of 64 per round.  Lane x's entry is
        // the last entry starting at or before u0 + x: entries e0+1.. (at most 64,
        // as ne <= SIB_BLOCK <= 64) mark their start slot in the window, and each
        // lane takes the highest marked slot at or below it.
        uint32_t e0 = 0;  // an entry starting at or before u0 (uniform)
        for (uint32_t u0 = 0; u0 < W; u0 += WAVE) {
            const uint32_t u = u0 + lane;
            w.mark[lane] = 0;
            wave_sync();
            {
                const uint32_t ej = e0 + 1 + lane;
                if (ej < ne) {
                    const uint32_t st = w.pre_sum[ej];
                    if (st < u0 + WAVE) atomicMax(&w.mark[st - u0], lane + 1);
                }
            }
            wave_sync();
            const uint64_t marked = __ballot(w.mark[lane] != 0);
            const uint64_t upto = marked & ((lane == WAVE - 1) ? ~0ull : ((2ull << lane) - 1ull));
            uint32_t a = e0;
            if (upto) a = e0 + w.mark[63 - __builtin_clzll(upto)];
            e0 = __shfl(a, WAVE - 1);
            wave_sync();  // marks are rewritten next round
            if (u < W) {
                const uint4 en = w.ent[a];
                const uint32_t off = u - en.x;
                const uint32_t fl = w.flag[a];
                const TermLds &t = terms[fl >> 1];
                bool hit;
                if (fl & 1u) {
                    uint32_t x = P.nbr[en.y + off];
                    hit = term_has(P, w, t, gb, x, en.z, en.w);
                } else {
                    uint32_t x = t.F.ptr[en.y + off];
                    hit = row_contains(P, en.z, en.w, x);
                }
                if (hit) {
                    if (per_entry) atomicAdd(&w.cnt[a], 1u);
                    else local += 1;
                }
            }
        }
        acc.add(A_BYTES, 4ull * W * 2);
        wave_sync();
        PH_SET(w, PH_P3);
        // phase 3: closed-form contributions (lane per sibling)
        for (uint32_t r0 = 0; r0 < nb; r0 += WAVE) {
            uint32_t i = r0 + lane;
            if (i < nb) {
                if (kind == 2) {
                    const uint32_t *cn = &w.cnt[i * K];
                    if (K == 3) {
                        uint64_t a = cn[0], b = cn[1], ab = cn[2];
                        local += a * b - ab;
                    } else {
                        // a=1 b=2 ab=3 c=4 ac=5 bc=6 abc=7 (index = mask - 1)
                        uint64_t a = cn[0], b = cn[1], ab = cn[2], c = cn[3], ac = cn[4], bc = cn[5], abc = cn[6];
                        local += a * b * c - ab * c - ac * b - bc * a + 2ull * abc;
                    }
                } else if (kind == 1) {
                    local += choosec(w.cnt[i], pl.tail);
                } else {
                    local += w.cnt[i];
                }
            }
        }
        wave_sync();
        PH_SET(w, PH_P1);
    }
    PH_SET(w, PH_P3);
#ifdef SGE_PHASE_PROFILE
    if (lane_id() == 0) w.ph[15] = max(w.ph[15], __builtin_amdgcn_s_memtime() - ph_t0);
    PH_EVMAX(w, PHE_COUNT_MAX, __builtin_amdgcn_s_memtime() - ph_t0);
    PH_EV(w, PHE_STEPS, steps);
    PH_EV(w, PHE_SIBS, nsib);
#endif
    for (int off = WAVE / 2; off > 0; off >>= 1) local += __shfl_xor(local, off);
    acc.add(A_COUNT, local);
    if (used_g) {
        // leave the global region all zero for the next call
        for (uint32_t k = 0; k < K; ++k) {
            const TermLds t = terms[k];
            if (!(t.flags & 32u)) continue;
            for (uint32_t q = lane; q < t.F.e - t.F.b; q += WAVE)
                gb[t.st_off + ((t.F.ptr[t.F.b + q] - t.bm_lo) >> 5)] = 0u;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  // this wave's next call sets bits again
    }
    PH_LEAVE(w);
    spent += steps;
    since += steps;
    return stopped;
}

__device__ void flush(const KParams &P, uint32_t plan, Acc &acc) {
    if (lane_id() == 0) {
        unsigned long long *o = &P.out[plan].count;  // PlanOut: the counters in Acc order
        for (int k = 0; k < A_NUM; ++k) {
            if (acc.c[k]) atomicAdd(o + k, acc.c[k]);
            acc.c[k] = 0;
        }
    }
}

// Enumeration: form the last level and write one row per embedding, images
// stored at their pattern vertex (row[order[j]] = phi[j]).
__device__ void emit_last(const KParams &P, WaveLds &w, uint32_t *wb, Acc &acc) {
    const DPlan &pl = w.pl;
    const int L = pl.k - 1;
    materialize(P, w, wb, L, acc);
    const Src c = w.cand[L];
    const uint32_t cmask = pl.check[L];
    const int lane = lane_id();
    for (uint32_t base = c.b; base < c.e; base += WAVE) {
        uint32_t idx = base + lane;
        bool ok = idx < c.e;
        uint32_t x = ok ? c.ptr[idx] : 0;
        ok = ok && !is_checked(w, cmask, x);
        uint64_t mask = __ballot(ok);
        uint32_t n = __popcll(mask);
        if (n == 0) continue;
        unsigned long long slot0 = 0;
        if (lane == 0) slot0 = atomicAdd(P.emb_n, (unsigned long long)n);
        slot0 = __shfl(slot0, 0);
        if (ok) {
            unsigned long long slot = slot0 + __popcll(mask & ((1ull << lane) - 1ull));
            if (slot < P.emb_cap) {
                uint32_t *row = P.emb + slot * pl.k;
                for (int j = 0; j < L; ++j) row[pl.order[j]] = w.phi[j];
                row[pl.order[L]] = x;
            }
        }
        acc.add(A_COUNT, n);
    }
}

__device__ void load_plan(const KParams &P, WaveLds &w, uint32_t plan) {
    const uint32_t *srcw = reinterpret_cast<const uint32_t *>(P.plans + plan);
    uint32_t *dstw = reinterpret_cast<uint32_t *>(&w.pl);
    for (int q = lane_id(); q < (int)(sizeof(DPlan) / 4); q += WAVE) dstw[q] = srcw[q];
    wave_sync();
}

// Write the untouched ranges of levels l0..d (below `stop`) as continuation
// items; the deepest level gives away only the upper half of its range.
// Returns false (and keeps all work) when the output buffer is full.
__device__ __forceinline__ uint32_t give_from(const WaveLds &w, int j, int d) {
    uint32_t b = w.pos[j], e = w.iend[j];
    return (j == d) ? b + (e - b + 1) / 2 : b;
}

// Each given range goes as min(length, P.split) items of near-equal length,
// so a wide range left at a shallow level reaches many waves of the next pass
// at once instead of one candidate per pass.
__device__ __forceinline__ uint32_t pieces_of(const KParams &P, uint32_t b, uint32_t e) {
    return b < e ? min(e - b, P.split) : 0u;
}

__device__ __forceinline__ void donate(const KParams &P, WaveLds &w, uint32_t plan, int l0, int d, int stop) {
    PH_ENTER(w, PH_DONATE);
    const int lane = lane_id();
    uint32_t n = 0;
    for (int j = l0; j <= d && j < stop; ++j) n += pieces_of(P, give_from(w, j, d), w.iend[j]);
    if (n == 0) {
        PH_LEAVE(w);
        return;
    }
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(P.out_count, (unsigned long long)n);
    base = __shfl(base, 0);
    if (base + n > P.item_cap) {
        // no room: mark the reserved slots that exist as empty, keep the work
        for (uint64_t slot = base; slot < P.item_cap && slot < base + n; ++slot)
            if (lane == 0) item_store(P.out_items + slot * ITEM_WORDS, NONE);
        if (base < P.item_cap) publish_items(P, base, min<uint64_t>(n, P.item_cap - base));
        PH_LEAVE(w);
        return;
    }
    PH_EV(w, PHE_DONATIONS, 1);
    uint64_t slot = base;
    for (int j = l0; j <= d && j < stop; ++j) {
        const uint32_t b = give_from(w, j, d), e = w.iend[j];
        const uint32_t np = pieces_of(P, b, e);
        const uint32_t len = e - b;
        for (uint32_t q = 0; q < np; ++q, ++slot) {
            uint32_t *it = P.out_items + slot * ITEM_WORDS;
            if (lane < ITEM_BODY) {
                uint32_t v;
                if (lane == 0) v = plan;
                else if (lane == 1) v = (uint32_t)j;
                else if (lane == 2) v = b + piece_start(len, q, np);
                else if (lane == 3) v = b + piece_start(len, q + 1, np);
                else v = (lane - 4 < j) ? w.phi[lane - 4] : 0u;
                item_store(it + lane, v);
            }
        }
        if (np) w.iend[j] = b;  // this wave keeps [pos, b)
    }
    publish_items(P, base, n);
    PH_LEAVE(w);
}

// A wave whose input is exhausted takes items of its own pass's output (the
// continuation items other waves are writing now) instead of leaving them to
// the next pass.  A wave counts in queue[Q_ACTIVE] while it owns an item
// (`holding`; every wave starts as owner).  Here it counts the item it owned as
// finished (queue[Q_DONE], when it was one of these items), gives up
// ownership, and takes a ticket t (fetch-add on queue[Q_TAKEN]): its item is
// output slot t once t is below the output count, and it owns it from then.
// With no owner left and as many items finished as written, no item can
// appear any more: the pass is complete and the waiting waves exit (the last
// wave to exit sets the output count to 0: nothing is left to a next pass).  Every written slot has
// one ticket holder, so no item is lost; tickets at or above the buffer's
// capacity never get one and exit at once.  A wait longer than LIVE_POLLS polls
// is counted as an error (the host fails the call) and ends the wave, as does
// the work cap.  Returns the slot, or ~0 to exit.
template <bool CAP>
__device__ __forceinline__ uint64_t live_claim(const KParams &P, WaveLds &w) {
    unsigned long long got = ~0ull;
    const uint32_t qs = w.qstate;
    if (lane_id() == 0) {
        unsigned long long *q = P.queue;
        // every operation below completes before the next is issued
        // (mem_drain): each load sees a state no older than the previous one's
        if (qs & QS_HOLD) {
            mem_drain();  // the items this wave wrote are published
            if (qs & QS_OWN) ctl_add(q + Q_DONE, 1ull);
            mem_drain();
            ctl_add(q + Q_ACTIVE, ~0ull);
        }
        const unsigned long long t = atomicAdd(q + Q_TAKEN, 1ull);
        uint32_t reps = 1;
        for (uint32_t polls = 0; t < P.item_cap; ++polls) {
            if (CAP && ctl_load(P.node_ctr) >= P.node_cap) break;
            unsigned long long c = ctl_load(P.out_count);
            mem_drain();
            if (t < c) {
                ctl_add(q + Q_ACTIVE, 1ull);
                got = t;
                break;
            }
            if (ctl_load(q + Q_ACTIVE) == 0) {
                mem_drain();
                const unsigned long long d = ctl_load(q + Q_DONE);
                mem_drain();
                c = ctl_load(P.out_count);
                mem_drain();
                if (t < c) continue;  // written meanwhile
                if (d >= (c < P.item_cap ? c : (unsigned long long)P.item_cap)) break;
            }
            if (polls >= LIVE_POLLS) {
                atomicAdd(q + Q_ERR, 1ull);
                break;
            }
            for (uint32_t r = 0; r < reps; ++r) __builtin_amdgcn_s_sleep(127);
            if (reps < 16) reps *= 2;
        }
    }
    got = __shfl(got, 0);
    w.qstate = (qs & QS_INPUT_DONE) | QS_OWN | (got != ~0ull ? QS_HOLD : 0u);
    return got;
}

// Word `lane` of claimed output slot j's body once its writer has published
// it (ok = false, and an error counted, if that takes longer than
// ITEM_WAIT_SPINS polls).
__device__ __forceinline__ uint32_t wait_item(const KParams &P, uint64_t j, bool &ok) {
    const uint32_t *it = P.out_items + j * ITEM_WORDS;
    uint32_t good = 1;
    if (lane_id() == 0) {
        const uint32_t epoch = (uint32_t)ctl_load(P.queue + Q_EPOCH);
        uint32_t spins = 0;
        while (item_load(it + ITEM_EPOCH) != epoch) {
            if (++spins >= ITEM_WAIT_SPINS) {
                atomicAdd(P.queue + Q_ERR, 1ull);
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        mem_drain();
    }
    ok = __shfl(good, 0) != 0;
    return (ok && lane_id() < ITEM_BODY) ? item_load(it + lane_id()) : 0u;
}

// Depth-first search from level l0 whose iteration range [pos, iend) is set,
// images of levels < l0 fixed.  Levels below `stop` are enumerated; `stop`
// is counted for all its siblings at once (or emitted when enumerating).
// Returns true when the work cap stopped the search (the wave then exits).
template <bool ENUM, bool CAP>
__device__ __forceinline__ bool run_dfs(const KParams &P, WaveLds &w, TermLds *terms, uint32_t *wb, uint32_t *gb,
                                        Acc &acc, uint32_t plan, int l0, int stop) {
    PH_ENTER(w, PH_DFS);
    const DPlan &pl = w.pl;
    uint32_t spent = 0, since = 0;  // wave steps since the last donation / node-cap report
    int d = l0;
    while (d >= l0) {
        if (d == stop) {
            if (ENUM) {
                const Src c = w.cand[d];
                const bool descend = (pl.bag >> d) & 1u;
                for (uint32_t i = w.pos[d]; i < w.iend[d]; ++i) {
                    uint32_t x = c.ptr[i];
                    if (is_checked(w, pl.check[d], x)) continue;
                    w.phi[d] = x;
                    if (descend) trie_enter_all(w, d, x, i, false);
                    emit_last(P, w, wb, acc);
                }
            } else if (count_last<CAP>(P, w, terms, wb, gb, acc, spent, since, plan)) {
                PH_LEAVE(w);
                return true;  // work cap reached (reported by count_last)
            }
            --d;
            continue;
        }
        const Src c = w.cand[d];
        uint32_t p = w.pos[d];
        const uint32_t pe = w.iend[d];
        // next candidate, skipping images excluded by injectivity checks
        uint32_t x = 0;
        bool found = false;
        while (p < pe) {
            x = c.ptr[p];
            ++p;
            if (!is_checked(w, pl.check[d], x)) {
                found = true;
                break;
            }
        }
        w.pos[d] = p;
        if (found) {
            w.phi[d] = x;
            // x is in d's trie list (a source of cand[d]); a miss would leave no
            // valid trie node to descend from, so it is skipped, never read
            if (((pl.bag >> d) & 1u) && !trie_enter_all(w, d, x, p - 1, false)) continue;
            if (pl.mode == SGE_MODE_EXPLORE && !ENUM) explore_after(P, w, wb, d, acc);
            const uint32_t ms = materialize(P, w, wb, d + 1, acc);
            ++d;
            // the searches of the intersection and the survivors written
            const uint32_t st = 1 + ms + ((w.cand[d].e - w.cand[d].b) >> 6);
            spent += st;
            since += st;
            if (P.budget && P.out_items && spent >= P.budget) {
                donate(P, w, plan, l0, d, stop);
                spent = 0;
            }
            if (cap_reached<CAP>(P, since)) {
                PH_LEAVE(w);
                return true;  // the host reports the call as stopped
            }
        } else {
            --d;
        }
    }
    if (CAP && since && lane_id() == 0) atomicAdd(P.node_ctr, (unsigned long long)since);
    PH_LEAVE(w);
    return false;
}

// Resident wavefronts per SIMD the register allocation is held to: the
// hardware maximum, 8 (64 VGPRs).  The search is latency bound (dependent loads
// of binary searches and LDS round trips) and every extra resident wavefront
// hides more of it; FSTAGE is sized so that 32 wavefronts' LDS fits a CU's
// 160 KB.
#ifndef SGE_WAVES_PER_SIMD
#define SGE_WAVES_PER_SIMD 8
#endif
static_assert((sizeof(WaveLds) + MAX_TERMS * sizeof(TermLds)) * 4 * SGE_WAVES_PER_SIMD <= 160 * 1024,
              "the resident wavefronts' LDS state must fit one CU");
// CAP: the call has a work cap (node_cap != 0); the uncapped kernel carries no
// cap accounting at all.
template <bool ENUM, bool CAP>
__global__ void __launch_bounds__(256, SGE_WAVES_PER_SIMD) wcoj_kernel(KParams P) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int lane = lane_id();
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);  // uniform: wave bases in SGPRs
    const int nwaves = blockDim.x / WAVE;
    WaveLds &w = reinterpret_cast<WaveLds *>(lds_raw)[wid];
    TermLds *terms = reinterpret_cast<TermLds *>(lds_raw + sizeof(WaveLds) * nwaves) + wid * MAX_TERMS;
    const uint64_t gwave = (uint64_t)blockIdx.x * (blockDim.x / WAVE) + wid;
    uint32_t *wb = P.scratch + gwave * P.wave_stride;
    uint32_t *gb = P.gbm ? P.gbm + gwave * P.gbm_words : nullptr;
    Acc acc{w.accv};
    if (lane == 0)
        for (int k = 0; k < A_NUM; ++k) w.accv[k] = 0;
#ifdef SGE_PHASE_PROFILE
    if (lane == 0) {
        for (int k = 0; k < 16; ++k) w.ph[k] = 0;
        w.ph_cur = PH_OTHER;
        w.ph_t = __builtin_amdgcn_s_memtime();
    }
    const unsigned long long t_kernel = __builtin_amdgcn_s_memtime();
#endif
    uint32_t cur_plan = NONE;
    const bool cont = P.in_items != nullptr;
    const uint64_t nitems = cont ? min((uint64_t)*P.in_count, P.item_cap) : P.nitems;
    unsigned long long t_item = __builtin_amdgcn_s_memtime();  // start of the current item
    uint32_t fetched = 15;  // the work cap is checked before the first item and every 16th after it
    uint32_t since = 0;     // wave steps of root tasks counted at the first level (not yet reported)
#ifdef SGE_PHASE_PROFILE
    unsigned long long ph_prev = 0, ph_long = 0, ph_long_start = 0;
    unsigned long long ph_wlong = 0;  // this wave's longest item (cycles >> 8)
    bool ph_cont = false;             // the current item is a continuation item
    const uint32_t ph_wid = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
#endif
    w.qstate = QS_HOLD;  // see live_claim (reset_pass_kernel counted every wavefront as an owner)
    while (true) {
        {
            // charge the time since the last item started to that item's plan
            const unsigned long long now = __builtin_amdgcn_s_memtime();
            if (cur_plan != NONE) acc.add(A_CYCLES, now - t_item);
#ifdef SGE_PHASE_PROFILE
            if (cur_plan != NONE) {
                const unsigned long long d = (now - t_item) >> 8;
                if ((d << 32 | (ph_prev & 0xffffffffull)) > ph_long) {
                    ph_long = d << 32 | (ph_prev & 0xffffffffull);
                    ph_long_start = d << 32 | (((t_item - t_kernel) >> 8) & 0xffffffffull);
                }
                if (!cont && d > ph_wlong && ph_wid < (uint32_t)PH_ITEM_WAVES) {
                    // this wave's longest item so far: its phase cycles and events
                    ph_wlong = d;
                    ph_switch(w, w.ph_cur);  // charge the open phase
                    if (lane == 0) {
                        unsigned long long *r = g_item + (size_t)ph_wid * PH_ITEM_WORDS;
                        for (int k = 0; k < PH_ITEM_PHASES; ++k) r[k] = w.ph[k] - w.ph0[k];
                        for (int k = 0; k < PH_ITEM_EVENTS; ++k) r[PH_ITEM_PHASES + k] = w.ev[k];
                        r[PH_ITEM_PHASES + PH_ITEM_EVENTS] = now - t_item;
                        r[PH_ITEM_PHASES + PH_ITEM_EVENTS + 1] = ph_prev;
                        r[PH_ITEM_PHASES + PH_ITEM_EVENTS + 2] = cur_plan | (unsigned long long)ph_cont << 32;
                        r[PH_ITEM_PHASES + PH_ITEM_EVENTS + 3] = t_item - t_kernel;
                    }
                }
            }
            ph_switch(w, w.ph_cur);
            if (lane == 0) {
                for (int k = 0; k < PH_ITEM_PHASES; ++k) w.ph0[k] = w.ph[k];
                for (int k = 0; k < PH_ITEM_EVENTS; ++k) w.ev[k] = 0;
            }
#endif
            t_item = now;
        }
        PH_SET(w, PH_QUEUE);
        unsigned long long item = 0;
        bool is_cont = false;  // a continuation item (else a root item)
        uint32_t iw = 0;       // its word `lane` (lanes below ITEM_BODY)
        if (!(w.qstate & QS_INPUT_DONE)) {
            if (lane == 0) item = atomicAdd(P.queue, 1ull);
            item = __shfl(item, 0);
            if (item >= nitems) {
                w.qstate |= QS_INPUT_DONE;
            } else if (cont) {
                is_cont = true;
                if (lane < ITEM_BODY) iw = item_load(P.in_items + item * ITEM_WORDS + lane);
            }
        }
        if (w.qstate & QS_INPUT_DONE) {
            // ---- this pass's own output, taken while it is being written
            if (!P.out_items || !P.own_items) break;
            item = live_claim<CAP>(P, w);
            if (item == ~0ull) break;
            bool ok;
            iw = wait_item(P, item, ok);
            if (!ok) continue;
            is_cont = true;
        }
#ifdef SGE_PHASE_PROFILE
        if (lane == 0) w.ph[PH_NUM] += 1;  // fetches
        ph_prev = item;
        ph_cont = is_cont;
#endif
        if (CAP && (++fetched & 15u) == 0) {  // every 16th item: has the work cap been reached?
            unsigned long long spent = 0;
            if (lane == 0) spent = __hip_atomic_load(P.node_ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__shfl(spent, 0) >= P.node_cap) break;
        }
        PH_SET(w, PH_ROOT);
        if (is_cont) {
            // ---- continuation item: rebuild the path, then search its range
            const uint32_t plan = __shfl(iw, 0);
            if (plan >= P.nplans) continue;  // empty slot
            if (plan != cur_plan) {
                if (cur_plan != NONE) flush(P, cur_plan, acc);
                cur_plan = plan;
                load_plan(P, w, plan);
            }
            const DPlan &pl = w.pl;
            const int l = (int)__shfl(iw, 1);
            const int stop = ENUM ? (int)pl.k - 2 : (int)pl.sib;
            if (l < 2 || l > stop || (ENUM && l == stop)) continue;  // never written by donate() / count_last
            const uint32_t rb = __shfl(iw, 2), re = __shfl(iw, 3);
            if (lane >= 4 && lane < 4 + l) w.phi[lane - 4] = iw;
            wave_sync();
            if (pl.mode == SGE_MODE_EXPLORE && !ENUM) {
                explore_after(P, w, wb, 0, acc);
                explore_after(P, w, wb, 1, acc);
            }
            // trie nodes of the path's bag levels (all present: the path was live)
            bool live = true;
            if ((pl.bag & 1u) && !trie_enter_all(w, 0, w.phi[0], NONE, true)) live = false;
            if (live && (pl.bag & 2u) && !trie_enter_all(w, 1, w.phi[1], NONE, true)) live = false;
            for (int j = 2; j <= l && live; ++j) {
                materialize(P, w, wb, j, acc);
                if (j < l && ((pl.bag >> j) & 1u) && !trie_enter_all(w, j, w.phi[j], NONE, true)) live = false;
                if (j < l && pl.mode == SGE_MODE_EXPLORE && !ENUM) explore_after(P, w, wb, j, acc);
            }
            if (!live) continue;
            w.pos[l] = max(rb, w.cand[l].b);
            w.iend[l] = min(re, w.cand[l].e);
            if (run_dfs<ENUM, CAP>(P, w, terms, wb, gb, acc, plan, l, stop)) break;  // work cap reached
            continue;
        }
        // ---- root pass: chunk of subtasks of one plan
        uint32_t plan;
        uint64_t ch;
        if ((item >> 32) == 0) {  // 32-bit division (the usual case)
            const uint32_t i32 = (uint32_t)item, c32 = i32 / P.nplans;
            ch = c32;
            plan = i32 - c32 * P.nplans;
        } else {
            ch = item / P.nplans;
            plan = (uint32_t)(item - ch * P.nplans);
        }
        if (plan != cur_plan) {
            if (cur_plan != NONE) flush(P, cur_plan, acc);
            cur_plan = plan;
            load_plan(P, w, plan);
        }
        const DPlan &pl = w.pl;
        if (pl.k < 2) continue;  // never launched by the host; no edge tasks exist for it
        const uint64_t ntask = pl.ntask;
        const uint64_t nsub = pl.split ? pl.nsub : ntask;
        // heavy-first: item 0 covers the highest CSR positions; the first
        // `head` items are single subtasks, the rest chunks of P.chunk
        const uint64_t head = min((uint64_t)P.head, nsub);
        uint64_t hi_u, lo_u;
        if (ch < head) {
            hi_u = nsub - ch;
            lo_u = hi_u - 1;
        } else {
            const uint64_t c2 = ch - head;
            if (head + c2 * P.chunk >= nsub) continue;
            hi_u = nsub - head - c2 * P.chunk;
            lo_u = hi_u > P.chunk ? hi_u - P.chunk : 0;
        }
        bool halt = false;  // the work cap stopped a search: this wave exits
        for (uint64_t u = lo_u; u < hi_u && !halt; ++u) {
            uint64_t ti = u, ci = 0, nch = 1;
            if (pl.split) {
                // sampled root edge owning subtask u: last i with split[i] <= u
                uint64_t a = 0, b = ntask;
                while (b - a > 1) {
                    uint64_t m = (a + b) >> 1;
                    if (pl.split[m] <= u) a = m; else b = m;
                }
                ti = a;
                ci = u - pl.split[a];
                nch = pl.split[a + 1] - pl.split[a];
            }
            const uint64_t t = pl.task_begin + ti * P.stride;
            uint32_t v0 = P.src[t], v1 = P.nbr[t];
            if ((pl.lower[1] & 1u) && !(v1 > v0)) continue;
            if ((pl.upper[1] & 1u) && !(v1 < v0)) continue;
            w.phi[0] = v0;
            w.phi[1] = v1;
            // root levels of the bag: the edge must start a bag tuple
            if ((pl.bag & 1u) && !trie_enter_all(w, 0, v0, NONE, true)) continue;
            if ((pl.bag & 2u) && !trie_enter_all(w, 1, v1, NONE, true)) continue;
            if (ci == 0) acc.add(A_TASKS, 1);
            if (pl.k < 3) {  // an edge pattern counts its root tasks; k = 1 is counted on the host
                if (pl.k == 2) acc.add(A_COUNT, 1);
                continue;
            }
            if (pl.mode == SGE_MODE_EXPLORE && !ENUM) {
                explore_after(P, w, wb, 0, acc);
                explore_after(P, w, wb, 1, acc);
            }
            if (!ENUM && pl.sib == 1) {
                uint32_t spent = 0;
                halt = count_last<CAP>(P, w, terms, wb, gb, acc, spent, since, plan);
                continue;
            }
            if (ENUM && pl.k == 3) {
                emit_last(P, w, wb, acc);
                continue;
            }
            const int stop = ENUM ? (int)pl.k - 2 : (int)pl.sib;  // level whose set feeds the leaf step
            materialize(P, w, wb, 2, acc);
            if (nch > 1) {  // this subtask iterates slice ci of the level-2 candidates
                // slice boundaries floor(len * ci / nch): when nch > len the
                // non-empty slices are spread over ci and the highest ci is
                // never empty.  Which slices are empty
                // changes the schedule, not the counts: leading non-empty
                // slices (the first len of ci) made the R-MAT 20 ranges 27%
                // slower and the R-MAT 23 ranges 40% faster; the same R-MAT 23
                // gain comes from split_span 8 with these slices.
                const Src c2 = w.cand[2];
                const uint64_t len = c2.e - c2.b;
                w.pos[2] = c2.b + (uint32_t)(len * ci / nch);
                w.iend[2] = c2.b + (uint32_t)(len * (ci + 1) / nch);
            }
            halt = run_dfs<ENUM, CAP>(P, w, terms, wb, gb, acc, plan, 2, stop);
        }
        if (halt) break;
    }
    if ((w.qstate & QS_HOLD) && lane == 0) atomicAdd(P.queue + Q_ACTIVE, ~0ull);  // the work cap stopped this wave
    if (lane == 0) {
        // every item written in this pass has been run in it (unless the work
        // cap stopped the pass or a wait failed, both reported to the host):
        // the last wave out leaves no items to a next pass
        const unsigned long long nw = (unsigned long long)gridDim.x * (blockDim.x / WAVE);
        if (P.own_items && atomicAdd(P.queue + Q_EXITED, 1ull) + 1 == nw)
            __hip_atomic_store(P.out_count, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (CAP && since && lane == 0) atomicAdd(P.node_ctr, (unsigned long long)since);
    if (cur_plan != NONE) acc.add(A_CYCLES, __builtin_amdgcn_s_memtime() - t_item);
    if (cur_plan != NONE) flush(P, cur_plan, acc);
#ifdef SGE_PHASE_PROFILE
    PH_SET(w, PH_OTHER);
    if (lane == 0) {
        w.ph[PH_ALL] = __builtin_amdgcn_s_memtime() - t_kernel;
        for (int k = 0; k <= PH_NUM; ++k) atomicAdd(&g_phase[k], w.ph[k]);
        atomicAdd(&g_phase[PH_NUM + 1], 1ull);
        atomicMax(&g_phase[PH_NUM + 2], w.ph[PH_ALL]);
        if (!cont) {
            atomicAdd(&g_phase[PH_NUM + 3], w.ph[PH_ALL]);
            atomicMax(&g_phase[PH_NUM + 4], w.ph[PH_ALL]);
            atomicAdd(&g_phase[PH_NUM + 5], 1ull);
            atomicMax(&g_phase[PH_NUM + 6], ph_long);
            atomicMax(&g_phase[PH_NUM + 7], ph_long_start);
            if (ph_wlong && ph_wid < (uint32_t)PH_ITEM_WAVES) atomicMax(&g_item_key, ph_wlong << 32 | ph_wid);
        }
        atomicMax(&g_phase[PH_NUM + 8], w.ph[14]);
        atomicMax(&g_phase[PH_NUM + 9], w.ph[15]);
    }
#endif
}

}  // namespace

#ifdef SGE_PHASE_PROFILE
// Phase profile of a -DSGE_PHASE_PROFILE build: out[PH_NUM] = summed wavefront
// clock cycles per phase since the last reset (other, materialise, exploration,
// last-level terms, F staging, entry setup, flattened probes, closed forms,
// queue, root and continuation set-up, DFS iteration, donation, whole kernel),
// out[PH_NUM] = queue fetches, out[PH_NUM + 1] = wavefronts, out[PH_NUM + 2] =
// the longest wavefront's cycles, out[PH_NUM + 3 .. 5] = the same three for
// root passes only (summed cycles, longest, wavefronts), out[PH_NUM + 6 .. 7] =
// the root passes' longest item (cycles >> 8 : queue index, cycles >> 8 : start
// within its wavefront), out[PH_NUM + 8 .. 9] = the longest F-set formation and
// the longest last-level count call; reset != 0 zeroes the counters.
// The longest item of the last root passes: out[0] = (cycles >> 8) << 32 |
// wave index, out[1 ..] = that wave's record (PH_ITEM_PHASES phase cycles,
// PH_ITEM_EVENTS events, item cycles, queue index, plan | continuation << 32,
// start since the wave started); reset != 0 clears the key.
extern "C" int sge_phase_item_read(unsigned long long *out, int reset) {
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_item_key), sizeof(unsigned long long)) != hipSuccess) return -1;
    const uint32_t wid = (uint32_t)(out[0] & 0xffffffffull);
    if (out[0] && wid < (uint32_t)PH_ITEM_WAVES &&
        hipMemcpyFromSymbol(out + 1, HIP_SYMBOL(g_item), sizeof(unsigned long long) * PH_ITEM_WORDS,
                            sizeof(unsigned long long) * PH_ITEM_WORDS * wid) != hipSuccess)
        return -1;
    if (reset) {
        static const unsigned long long zero = 0;
        if (hipMemcpyToSymbol(HIP_SYMBOL(g_item_key), &zero, sizeof(zero)) != hipSuccess) return -1;
    }
    return 0;
}

extern "C" int sge_phase_profile_read(unsigned long long *out, int reset) {
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_phase), sizeof(unsigned long long) * (PH_NUM + 10)) != hipSuccess)
        return -1;
    if (reset) {
        static const unsigned long long zero[PH_NUM + 10] = {};
        if (hipMemcpyToSymbol(HIP_SYMBOL(g_phase), zero, sizeof(zero)) != hipSuccess) return -1;
    }
    return 0;
}
#endif

// ============================================================================
// Host side
// ============================================================================
namespace {

struct DeviceCtx {
    bool init = false;
    int dev = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    uint32_t *scratch = nullptr;
    size_t scratch_bytes = 0;
    unsigned long long *queue = nullptr;
    PlanOut *out = nullptr;
    size_t out_cap = 0;
    DPlan *plans = nullptr;
    size_t plans_cap = 0;
    uint32_t *items[2] = {nullptr, nullptr};  // continuation items (ping-pong)
    unsigned long long *icount = nullptr;     // [2] item counts, [2] work-cap counter
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t ex_base = 0, pre_f = 0, term_base = 0, nbuf = NBUF_MAX;  // slab layout of the uploaded plans
    uint32_t *gbm = nullptr;  // per-wave global bitmap regions (kept all zero between uses)
    size_t gbm_bytes = 0;
};

constexpr uint64_t ITEM_CAP = 1ull << 21;  // continuation items per pass buffer (96 MB)

DeviceCtx &dctx() {
    static DeviceCtx c;
    return c;
}

std::mutex &dmutex() {
    static std::mutex m;
    return m;
}

int ensure_ctx() {
    DeviceCtx &c = dctx();
    if (c.init) return SGE_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        sge::set_error("no HIP device available: the GPU executor requires an MI355X (gfx950)");
        return SGE_EDEVICE;
    }
    HIP_TRY(hipGetDevice(&c.dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, c.dev));
    c.num_cu = prop.multiProcessorCount;
    HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&c.queue, sizeof(unsigned long long) * Q_WORDS));
    HIP_TRY(hipMemset(c.queue, 0, sizeof(unsigned long long) * Q_WORDS));
    HIP_TRY(hipEventCreate(&c.ev0));
    HIP_TRY(hipEventCreate(&c.ev1));
    HIP_TRY(hipMalloc(&c.items[0], sizeof(uint32_t) * ITEM_WORDS * ITEM_CAP));
    HIP_TRY(hipMalloc(&c.items[1], sizeof(uint32_t) * ITEM_WORDS * ITEM_CAP));
    // epoch words start at 0, below every pass epoch (reset_pass_kernel counts from 1)
    HIP_TRY(hipMemset(c.items[0], 0, sizeof(uint32_t) * ITEM_WORDS * ITEM_CAP));
    HIP_TRY(hipMemset(c.items[1], 0, sizeof(uint32_t) * ITEM_WORDS * ITEM_CAP));
    HIP_TRY(hipMalloc(&c.icount, sizeof(unsigned long long) * 4));
    c.init = true;
    return SGE_OK;
}

int ensure_buffer(void **p, size_t &cap, size_t need) {
    if (need <= cap && *p) return SGE_OK;
    if (*p) HIP_TRY(hipFree(*p));
    *p = nullptr;
    size_t n = std::max(need, cap * 2);
    HIP_TRY(hipMalloc(p, n));
    cap = n;
    return SGE_OK;
}

DPlan to_dplan(const sge_plan &p) {
    DPlan d;
    std::memset(&d, 0, sizeof(d));
    d.k = p.k;
    d.mode = p.mode;
    d.tail = (p.mode == SGE_MODE_IEP && p.tail >= 2) ? p.tail : 0;
    d.last = d.tail ? p.k - d.tail : (p.k >= 1 ? p.k - 1 : 0);
    d.sib = d.last ? d.last - 1 : 0;
    for (int i = 0; i < MAXK; ++i) {
        d.parent[i] = p.parent[i];
        d.lower[i] = p.lower[i];
        d.upper[i] = p.upper[i];
        d.check[i] = p.check[i];
        d.reuse[i] = (int8_t)p.reuse[i];    // validated: -1 or a level below i
        d.ready[i] = (uint8_t)p.ready[i];   // validated: below i
        d.order[i] = p.order[i];
    }
    uint32_t L = d.last, s = d.sib;
    d.tail_kind = d.tail ? p.tail_kind : 0;
    const uint32_t below = (1u << s) - 1u;
    auto term = [&](uint32_t j, uint32_t members) {
        uint32_t par = 0, lo = 0, up = 0, ck = 0;
        for (uint32_t m = members; m; m &= m - 1) {
            uint32_t lev = __builtin_ctz(m);
            par |= d.parent[lev];
            lo |= d.lower[lev];
            up |= d.upper[lev];
            ck |= d.check[lev];
        }
        d.t_par[j] = par & below;
        d.t_lower[j] = lo & below;
        d.t_upper[j] = up & below;
        d.t_check[j] = ck & below;
        d.t_sflags[j] = ((par >> s) & 1u) | (((lo >> s) & 1u) << 1) | (((up >> s) & 1u) << 2) | (((ck >> s) & 1u) << 3);
    };
    if (d.tail_kind == 2) {
        d.nterms = (1u << d.tail) - 1u;
        for (uint32_t B = 1; B <= d.nterms; ++B) {
            uint32_t members = 0;
            for (uint32_t j = 0; j < d.tail; ++j)
                if ((B >> j) & 1u) members |= 1u << (L + j);
            term(B - 1, members);
        }
    } else {
        d.nterms = 1;
        term(0, 1u << L);
    }
    d.f_parent = d.parent[L] & ~(1u << s);
    d.f_ready = d.f_parent ? 31 - __builtin_clz(d.f_parent) : 0;
    // A term's F formed from rows starts from the candidate set of an earlier
    // level r < s (the reference's candidatesBefore, decomposition.h:15-57,
    // applied per closed-form term) when r's parents are among the term's and
    // every bound and check of r is one the term applies too: cand[r] then
    // holds the intersection of those rows under a subset of the term's
    // constraints, so F = cand[r] ∩ (the other rows) with the term's bounds and
    // checks.  The most parents wins (ties: the later level).  Terms formed
    // from two smaller terms' F sets, universes and bag plans take none.
    d.t_reuse = 0xffffffffu;
    if (!p.bag) {
        for (uint32_t j = 0; j < d.nterms; ++j) {
            const uint32_t par = d.t_par[j];
            if (!par) continue;
            const uint32_t B = j + 1, b1 = B & (~B + 1u);
            if (d.tail_kind == 2 && B != b1 && d.t_par[b1 - 1] && d.t_par[(B ^ b1) - 1]) continue;
            const uint32_t excl = d.t_check[j] | d.t_lower[j] | d.t_upper[j] | par;
            int best = -1, bp = 1;
            for (uint32_t r = 2; r < s; ++r) {
                if ((d.parent[r] & par) != d.parent[r]) continue;
                const int np = __builtin_popcount(d.parent[r]);
                if (np < 2) continue;
                if ((d.lower[r] & ~d.t_lower[j]) || (d.upper[r] & ~d.t_upper[j]) || (d.check[r] & ~excl)) continue;
                if (np >= bp) {
                    best = (int)r;
                    bp = np;
                }
            }
            if (best >= 0) d.t_reuse = (d.t_reuse & ~(15u << (4 * j))) | ((uint32_t)best << (4 * j));
        }
    }
    d.bag = 0;
    for (int b = 0; b < SGE_MAX_BAGS; ++b) {
        d.bagm[b] = (p.bag >> (8 * b)) & 0xffu;
        d.bag |= d.bagm[b];
        d.trie[b] = nullptr;
    }
    if (d.bag) {
        // a bag level's set is not a superset of another level's candidates
        // (its trie list constrains that level alone), and a bag level's own
        // sources are its trie list plus the rows of non-bag parents
        for (int i = 0; i < MAXK; ++i)
            if (((d.bag >> i) & 1u) || (d.reuse[i] >= 0 && ((d.bag >> d.reuse[i]) & 1u))) d.reuse[i] = -1;
    }
    return d;
}

int validate_plan(const sge_plan &p) {
    if (p.k < 1 || p.k > MAXK) {
        sge::set_error("plan: k must be 1..8");
        return SGE_EINVAL;
    }
    if (p.mode > SGE_MODE_IEP) {
        sge::set_error("plan: unknown mode");
        return SGE_EINVAL;
    }
    uint32_t seen = 0;
    for (uint32_t i = 0; i < p.k; ++i) {
        if (p.order[i] >= p.k || ((seen >> p.order[i]) & 1u)) {
            sge::set_error("plan: order is not a permutation");
            return SGE_EINVAL;
        }
        seen |= 1u << p.order[i];
        uint32_t earlier = (1u << i) - 1u;
        if ((p.parent[i] | p.lower[i] | p.upper[i] | p.check[i]) & ~earlier) {
            sge::set_error("plan: masks may only name earlier levels");
            return SGE_EINVAL;
        }
        if (i > 0 && p.parent[i] == 0) {
            sge::set_error("plan: every level after the first needs an earlier neighbour");
            return SGE_EINVAL;
        }
        if (p.reuse[i] >= (int32_t)i || p.reuse[i] < -1 || (p.reuse[i] >= 0 && p.reuse[i] < 2)) {
            sge::set_error("plan: bad reuse level");
            return SGE_EINVAL;
        }
        if (p.ready[i] >= (i ? i : 1)) {
            sge::set_error("plan: bad ready level");
            return SGE_EINVAL;
        }
        // exploration forms a level's parent intersection after level ready[i]:
        // every parent must be matched by then
        if (p.mode == SGE_MODE_EXPLORE && i >= 2 && p.ready[i] < 31u - __builtin_clz(p.parent[i])) {
            sge::set_error("plan: ready level precedes a parent (exploration would read an unmatched image)");
            return SGE_EINVAL;
        }
        // a reused set must contain this level's candidates: its parents and
        // bounds are this level's, and each level it must differ from is
        // excluded here too (by a check, a bound or adjacency)
        if (p.reuse[i] >= 0) {
            const uint32_t r = (uint32_t)p.reuse[i];
            const uint32_t excl = p.check[i] | p.lower[i] | p.upper[i] | p.parent[i];
            if ((p.parent[r] & ~p.parent[i]) || (p.lower[r] & ~p.lower[i]) || (p.upper[r] & ~p.upper[i]) ||
                (p.check[r] & ~excl)) {
                sge::set_error("plan: reuse level's set is not a superset of this level's candidates");
                return SGE_EINVAL;
            }
        }
    }
    if (p.mode == SGE_MODE_IEP && p.tail >= 2) {
        if (p.tail >= p.k || p.k - p.tail < 2) {
            sge::set_error("plan: IEP tail must leave at least two matched levels");
            return SGE_EINVAL;
        }
        uint32_t b = p.k - p.tail, tailmask = ((1u << p.k) - 1u) & ~((1u << b) - 1u);
        if (p.tail_kind == 2) {
            if (p.tail > 3) {
                sge::set_error("plan: inclusion-exclusion tails hold at most 3 levels");
                return SGE_EINVAL;
            }
            for (uint32_t i = b; i < p.k; ++i)
                if ((p.parent[i] | p.lower[i] | p.upper[i]) & tailmask) {
                    sge::set_error("plan: an inclusion-exclusion tail must be independent and unbounded internally");
                    return SGE_EINVAL;
                }
        } else if (p.tail_kind == 1) {
            for (uint32_t i = b; i < p.k; ++i) {
                uint32_t below = (1u << b) - 1u;
                if ((p.parent[i] & ~below) || (p.parent[i] & below) != (p.parent[b] & below) ||
                    (p.lower[i] & below) != (p.lower[b] & below) || (p.upper[i] & below) != (p.upper[b] & below) ||
                    (p.check[i] & below) != (p.check[b] & below) || (p.upper[i] & ~below) ||
                    (p.lower[i] & ~below) != (((1u << i) - 1u) & ~below)) {
                    sge::set_error("plan: a choosec tail must be a chained group of twins");
                    return SGE_EINVAL;
                }
            }
        } else {
            sge::set_error("plan: tail_kind must be 1 or 2 when tail >= 2");
            return SGE_EINVAL;
        }
    }
    return sge::bags_check(p);
}

// Defaults of sge_tuning.donate_budget and donate_split, and the split's
// maximum.  Measured over R-MAT scale 20 root-edge subsets (78 seven-vertex
// plans), R-MAT scale 8 and Chung-Lu 7115 whole jobs: splitting each donated
// range 64 ways ran the subsets 2.4x (cost-model orders) to 5.2x (learned
// orders) faster than one item per range, with a budget of 1024 best or
// within 3% of best on every workload.
constexpr uint32_t DONATE_BUDGET_DEFAULT = 1024;
constexpr uint32_t DONATE_SPLIT_DEFAULT = 64;
constexpr uint32_t DONATE_SPLIT_MAX = 256;

struct LaunchCfg {
    int block = 256;
    int grid = 0;
    size_t lds = 0;
    uint32_t chunk = 64;
    uint32_t head = 0;
    uint32_t budget = DONATE_BUDGET_DEFAULT;
    uint32_t probe_ratio = 2;
    uint32_t bitmap_ratio = 2;
    unsigned long long node_cap = 0;
    uint32_t gbm_min = 2048;
    uint32_t split = 1;      // items per donated range (at most)
    uint32_t pass_mode = 0;  // sge_tuning::pass_items (0 by the graph, 1 next pass, 2 own pass)
    bool budget_auto = true; // sge_tuning::donate_budget 0: the budget follows the graph (budget_for)
};

// Whether idle wavefronts take their own pass's continuation items.  By default
// only on graphs with rows of at least OWN_ITEMS_MIN_DEGREE ids: there long
// items bound each pass and the items they hand over are best run at once
// (R-MAT 20 ranges 72.5 -> 47.8 ms, R-MAT 23 ranges 62.3 -> 35.0 ms); on the
// small graphs the per-item claims cost more than they gain (R-MAT scale 8
// whole job 110 -> 161 ms batched, Chung-Lu n = 7,115 plain 488 -> 545 ms).
constexpr uint32_t OWN_ITEMS_MIN_DEGREE = 1u << 14;
bool own_items_for(const LaunchCfg &cfg, const sge_graph *g) {
    return cfg.pass_mode == 2 || (cfg.pass_mode == 0 && g->h.max_degree >= OWN_ITEMS_MIN_DEGREE);
}

// The default donation budget on graphs with rows of more than
// BUDGET_FULL_DEGREE ids: a wave step costs more cycles there (its searches run
// through longer rows), so the budget in steps shrinks with the largest row,
// DONATE_BUDGET_DEFAULT * BUDGET_FULL_DEGREE / max degree, at least
// BUDGET_MIN_AUTO.  R-MAT scale 23 root ranges (largest row 258,497): 17.0 ms
// at 1,024 steps, 10.4-10.9 ms at 256 or 512; the R-MAT scale 20 ranges
// (64,810) keep 1,024, their best (256: 58.5 ms, 512: 41.6, 1,024: 38.7).
constexpr uint32_t BUDGET_FULL_DEGREE = 1u << 16;
constexpr uint32_t BUDGET_MIN_AUTO = 256;
uint32_t budget_for(const LaunchCfg &cfg, const sge_graph *g) {
    if (!cfg.budget_auto || g->h.max_degree <= BUDGET_FULL_DEGREE) return cfg.budget;
    const uint64_t b = (uint64_t)DONATE_BUDGET_DEFAULT * BUDGET_FULL_DEGREE / g->h.max_degree;
    return (uint32_t)std::max<uint64_t>(b, BUDGET_MIN_AUTO);
}


int launch_cfg(const sge_tuning *t, bool enumerate, LaunchCfg &cfg) {
    DeviceCtx &c = dctx();
    sge_tuning def;
    sge_tuning_default(&def);
    if (!t) t = &def;
    cfg.block = (int)t->block_threads;
    if (cfg.block < 64 || cfg.block > 256 || cfg.block % 64) {
        sge::set_error("tuning: block_threads must be 64, 128, 192 or 256");
        return SGE_EINVAL;
    }
    cfg.chunk = t->chunk ? t->chunk : 4;
    cfg.head = t->head_items;
    cfg.budget = t->donate_budget ? (t->donate_budget >= (1u << 30) ? 0u : t->donate_budget) : DONATE_BUDGET_DEFAULT;
    cfg.budget_auto = t->donate_budget == 0;
    cfg.probe_ratio = t->stage_max ? t->stage_max : 2u;
    cfg.bitmap_ratio = t->bitmap_ratio ? t->bitmap_ratio : cfg.probe_ratio;
    cfg.node_cap = (unsigned long long)t->node_cap_k * NODE_STEP;
    cfg.gbm_min = t->global_bitmap_min ? t->global_bitmap_min : 2048u;
    cfg.split = t->donate_split ? t->donate_split : DONATE_SPLIT_DEFAULT;
    if (cfg.split > DONATE_SPLIT_MAX) {
        sge::set_error("tuning: donate_split must be at most " + std::to_string(DONATE_SPLIT_MAX));
        return SGE_EINVAL;
    }
    if (t->pass_items > 2) {
        sge::set_error("tuning: pass_items must be 0, 1 or 2");
        return SGE_EINVAL;
    }
    cfg.pass_mode = t->pass_items;
    cfg.lds = (size_t)(cfg.block / WAVE) * (sizeof(WaveLds) + MAX_TERMS * sizeof(TermLds));
    int per_cu = (int)t->blocks_per_cu;
    if (per_cu <= 0) {
        int occ = 0;
        if (enumerate)
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, wcoj_kernel<true, false>, cfg.block, cfg.lds));
        else
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, wcoj_kernel<false, false>, cfg.block, cfg.lds));
        per_cu = std::max(1, occ);
    }
    cfg.grid = c.num_cu * per_cu;
    if (cfg.head == 0) cfg.head = (uint32_t)cfg.grid * (uint32_t)(cfg.block / WAVE);  // one per resident wave
    else if (cfg.head >= (1u << 30)) cfg.head = 0;
    return SGE_OK;
}

// Root-queue items of one plan with `nsub` subtasks: single-subtask head items,
// then chunks.
uint64_t items_per_plan(uint64_t nsub, const LaunchCfg &cfg) {
    const uint64_t head = std::min<uint64_t>(cfg.head, nsub);
    return head + (nsub - head + cfg.chunk - 1) / cfg.chunk;
}

// Slab layout for a set of plans: only the buffers they use (level sets up to
// the largest k, exploration's early intersections only when a plan explores,
// as many F sets as the most terms).  At max degree D a wavefront's slab is
// nbuf * (D + 64) words, so this sets the scratch footprint.
struct BufLayout {
    uint32_t ex_base, pre_f, term_base, nbuf;
};

BufLayout buf_layout(const DPlan *dp, uint32_t n) {
    uint32_t kmax = 2, terms = 1;
    bool explore = false;
    for (uint32_t i = 0; i < n; ++i) {
        kmax = std::max(kmax, dp[i].k);
        terms = std::max(terms, dp[i].nterms);
        explore = explore || dp[i].mode == SGE_MODE_EXPLORE;
    }
    BufLayout L;
    if (explore) {
        L.ex_base = kmax;         // levels 2..sib at ex_base + level < 2 kmax
        L.pre_f = 2 * kmax;
        L.term_base = 2 * kmax + 1;
    } else {
        L.ex_base = L.pre_f = L.term_base = kmax;  // no early intersections
    }
    L.nbuf = L.term_base + terms;
    return L;
}

void set_layout(KParams &kp, uint32_t ex_base, uint32_t pre_f, uint32_t term_base) {
    kp.ex_base = ex_base;
    kp.pre_f = pre_f;
    kp.term_base = term_base;
}

// Words of a wave's global bitmap region: three terms' bitmaps over the whole
// id range, at most GBM_MAX_WORDS (3 MB per wave; 26 GB for 8,192 wavefronts).
constexpr uint64_t GBM_MAX_WORDS = 3ull << 18;
uint32_t gbm_words_for(const sge_graph *g) {
    const uint64_t w = 3ull * (((uint64_t)g->h.n >> 5) + 1) + 64;
    return (uint32_t)((std::min(w, GBM_MAX_WORDS) + 63) / 64 * 64);
}

int ensure_scratch(const sge_graph *g, const LaunchCfg &cfg, KParams &kp) {
    DeviceCtx &c = dctx();
    uint32_t stride = ((g->h.max_degree + 64 + 63) / 64) * 64;
    kp.buf_stride = stride;
    kp.wave_stride = (uint64_t)stride * c.nbuf;
    set_layout(kp, c.ex_base, c.pre_f, c.term_base);
    size_t waves = (size_t)cfg.grid * (cfg.block / WAVE);
    size_t need = waves * kp.wave_stride * sizeof(uint32_t);
    int rc = ensure_buffer((void **)&c.scratch, c.scratch_bytes, need);
    if (rc) return rc;
    kp.scratch = c.scratch;
    // the bitmap region starts all zero and every use clears what it set
    kp.gbm_words = gbm_words_for(g);
    const size_t gneed = waves * (size_t)kp.gbm_words * sizeof(uint32_t);
    if (gneed > c.gbm_bytes) {
        if (c.gbm) HIP_TRY(hipFree(c.gbm));
        c.gbm = nullptr;
        c.gbm_bytes = 0;
        HIP_TRY(hipMalloc(&c.gbm, gneed));
        HIP_TRY(hipMemset(c.gbm, 0, gneed));
        c.gbm_bytes = gneed;
    }
    kp.gbm = c.gbm;
    return SGE_OK;
}

// Before every pass: the input queue counter, the pass's own-output counters
// and its output count to 0, a new epoch for the items it writes, and every
// wavefront of the launch counted as an owner (each starts with its input
// items and gives that up once, so no wavefront adds itself at its start).
__global__ void reset_pass_kernel(unsigned long long *queue, unsigned long long *count, unsigned long long nwaves) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        queue[0] = 0;
        queue[Q_TAKEN] = 0;
        queue[Q_ACTIVE] = nwaves;
        queue[Q_DONE] = 0;
        queue[Q_EXITED] = 0;
        queue[Q_EPOCH] += 1;
        *count = 0;
    }
}

// A continuation item that a wave claimed and never saw published (its wait
// timed out) means a lost item: the counts of the call are wrong.
int check_item_waits(const unsigned long long *queue) {
    unsigned long long err = 0;
    HIP_TRY(hipMemcpy(&err, queue + Q_ERR, sizeof(err), hipMemcpyDeviceToHost));
    if (err) {
        sge::set_error("continuation items: " + std::to_string(err) + " waits for a claimed item timed out");
        return SGE_EDEVICE;
    }
    return SGE_OK;
}

__global__ void reset_kernel(PlanOut *out, uint32_t n, unsigned long long *queue, unsigned long long *emb_n,
                             unsigned long long *cap_ctr) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = PlanOut{0, 0, 0, 0, 0, 0, 0, 0};
    if (i == 0) {
        queue[0] = 0;
        queue[Q_ERR] = 0;
        if (emb_n) *emb_n = 0;
        if (cap_ctr) *cap_ctr = 0;
    }
}

// The kernel variant for a launch: enumeration or counting, with or without
// the work cap's accounting.
unsigned long long waves_of(const LaunchCfg &cfg) { return (unsigned long long)cfg.grid * (cfg.block / WAVE); }

void launch_wcoj(bool enumerate, const LaunchCfg &cfg, hipStream_t st, const KParams &kp) {
    if (enumerate)
        wcoj_kernel<true, false><<<cfg.grid, cfg.block, cfg.lds, st>>>(kp);
    else if (kp.node_cap)
        wcoj_kernel<false, true><<<cfg.grid, cfg.block, cfg.lds, st>>>(kp);
    else
        wcoj_kernel<false, false><<<cfg.grid, cfg.block, cfg.lds, st>>>(kp);
}

__global__ void skip_kernel(const uint32_t *__restrict__ nbr, uint64_t nblk, uint32_t *__restrict__ skip) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nblk) skip[i] = nbr[i * SKIP_G];
}

// The graph's skip index (skip[i] = nbr[SKIP_G i], i < ceil(nnz / SKIP_G)),
// built on the device at the first launch that needs it and freed with the
// device mirror; graphs without a row of SKIP_MIN elements get none.
int ensure_skip(sge_graph *g, hipStream_t st) {
    if (g->d_skip || g->h.max_degree < SKIP_MIN || !g->d_nbr) return SGE_OK;
    const uint64_t nblk = (g->h.nnz + SKIP_G - 1) / SKIP_G;
    HIP_TRY(hipMalloc(&g->d_skip, sizeof(uint32_t) * nblk));
    skip_kernel<<<(unsigned)((nblk + 255) / 256), 256, 0, st>>>((const uint32_t *)g->d_nbr, nblk, (uint32_t *)g->d_skip);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(st));
    return SGE_OK;
}

// One counting launch over `np` device plans (already uploaded at `dplans`).
int run_launch(sge_graph *g, const DPlan *dplans, uint32_t np, PlanOut *dout, uint64_t stride, uint64_t max_sub,
               const LaunchCfg &cfg, bool enumerate, uint32_t *emb, uint64_t emb_cap, unsigned long long *emb_n,
               float *ms) {
    DeviceCtx &c = dctx();
    if (int rc = ensure_skip(g, c.stream)) return rc;
    KParams kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.off = (const uint32_t *)g->d_off;
    kp.nbr = (const uint32_t *)g->d_nbr;
    kp.src = (const uint32_t *)g->d_src;
    kp.skip = (const uint32_t *)g->d_skip;
    kp.n = g->h.n;
    kp.nnz = g->h.nnz;
    kp.plans = dplans;
    kp.nplans = np;
    kp.stride = stride ? stride : 1;
    kp.chunk = cfg.chunk;
    kp.head = cfg.head;
    kp.nitems = items_per_plan(max_sub, cfg) * np;
    kp.queue = c.queue;
    kp.out = dout;
    kp.emb = emb;
    kp.emb_cap = emb_cap;
    kp.emb_n = emb_n;
    int rc = ensure_scratch(g, cfg, kp);
    if (rc) return rc;
    kp.item_cap = ITEM_CAP;
    kp.budget = budget_for(cfg, g);
    kp.split = cfg.split;
    kp.own_items = own_items_for(cfg, g);
    kp.probe_ratio = cfg.probe_ratio;
    kp.bitmap_ratio = cfg.bitmap_ratio;
    kp.node_cap = cfg.node_cap;
    kp.gbm_min = cfg.gbm_min;
    kp.node_ctr = c.icount + 2;  // read back with the item counts, in one copy per pass
    reset_kernel<<<(np + 255) / 256, 256, 0, c.stream>>>(dout, np, c.queue, emb_n, kp.node_ctr);
    reset_pass_kernel<<<1, 64, 0, c.stream>>>(c.queue, c.icount, waves_of(cfg));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(c.ev0, c.stream));
    int cur = 0;
    for (int pass = 0;; ++pass) {
        if (pass == 0) {
            kp.in_items = nullptr;
            kp.in_count = nullptr;
        } else {
            unsigned long long hc[3] = {0, 0, 0};  // item counts, work spent
            HIP_TRY(hipMemcpyAsync(hc, c.icount, sizeof(hc), hipMemcpyDeviceToHost, c.stream));
            HIP_TRY(hipStreamSynchronize(c.stream));
            if (hc[cur] == 0) break;
            if (kp.node_cap && hc[2] >= kp.node_cap) break;
            kp.in_items = c.items[cur];
            kp.in_count = c.icount + cur;
            cur ^= 1;
            reset_pass_kernel<<<1, 64, 0, c.stream>>>(c.queue, c.icount + cur, waves_of(cfg));
        }
        kp.out_items = c.items[cur];
        kp.out_count = c.icount + cur;
        if (pass == 0 && kp.nitems == 0) break;
        launch_wcoj(enumerate, cfg, c.stream, kp);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(c.ev1, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, c.ev0, c.ev1));
    if (ms) *ms = t;
    if (int rc = check_item_waits(c.queue)) return rc;
    if (kp.node_cap) {
        unsigned long long spent = 0;
        HIP_TRY(hipMemcpyAsync(&spent, kp.node_ctr, sizeof(spent), hipMemcpyDeviceToHost, c.stream));
        HIP_TRY(hipStreamSynchronize(c.stream));
        if (spent >= kp.node_cap) {
            sge::set_error("work cap reached (" + std::to_string(spent) + " wave steps): counts are partial");
            return SGE_ELIMIT;
        }
    }
    return SGE_OK;
}

// ---- root-task split prefixes (cached per graph)
struct SplitKey {
    uint32_t shape;  // parent mask of level 2 (1, 2, 3), or 0: no level-2 iteration
    uint32_t l1;     // level-1 bound bits: 1 lower, 2 upper
    uint32_t l2;     // level-2 bounds from levels 0, 1: lower[2] & 3 | (upper[2] & 3) << 2
    uint64_t tb, te, stride;
    uint32_t span;
    bool operator<(const SplitKey &o) const {
        return std::tie(shape, l1, l2, tb, te, stride, span) <
               std::tie(o.shape, o.l1, o.l2, o.tb, o.te, o.stride, o.span);
    }
};
struct SplitArr {
    unsigned long long *d = nullptr;
    unsigned long long nsub = 0;
};
using SplitCache = std::map<SplitKey, SplitArr>;

void free_split_cache(sge_graph *g) {
    if (!g->split_cache) return;
    auto *c = static_cast<SplitCache *>(g->split_cache);
    for (auto &kv : *c)
        if (kv.second.d) (void)hipFree(kv.second.d);
    delete c;
    g->split_cache = nullptr;
}

// Subtasks per sampled root edge: 0 when the edge fails the level-1 bound,
// else ceil(bound / span) clamped to [1, 4096], where bound is the size bound of
// the level-2 candidate set given by the endpoint degrees.
int get_split(sge_graph *g, const sge_plan &p, uint64_t tb, uint64_t te, uint64_t stride, uint32_t span,
              const unsigned long long **out, unsigned long long *nsub) {
    *out = nullptr;
    *nsub = 0;
    SplitKey key;
    key.shape = (p.k >= 4 && !(p.mode == SGE_MODE_IEP && p.tail >= 2 && p.k - p.tail <= 2)) ? (p.parent[2] & 3u) : 0u;
    key.l1 = ((p.lower[1] & 1u) ? 1u : 0u) | ((p.upper[1] & 1u) ? 2u : 0u);
    key.l2 = (p.lower[2] & 3u) | ((p.upper[2] & 3u) << 2);
    key.tb = tb;
    key.te = te;
    key.stride = stride;
    key.span = span;
    if (key.shape == 0 || span >= (1u << 30)) return SGE_OK;  // no level-2 iteration: no split
    if (!g->split_cache) g->split_cache = new SplitCache();
    auto *cache = static_cast<SplitCache *>(g->split_cache);
    auto it = cache->find(key);
    if (it == cache->end()) {
        const sge::HostGraph &h = g->h;
        uint64_t ntask = (te - tb + stride - 1) / stride;
        std::vector<unsigned long long> pre(ntask + 1, 0);
        uint32_t v = 0;
        for (uint64_t i = 0; i < ntask; ++i) {
            uint64_t t = tb + i * stride;
            while (h.off[v + 1] <= t) ++v;
            uint32_t v0 = v, v1 = h.nbr[t];
            uint64_t nch;
            if (((key.l1 & 1u) && !(v1 > v0)) || ((key.l1 & 2u) && !(v1 < v0))) {
                nch = 0;
            } else {
                // bound on the level-2 candidates: each parent row clamped to
                // level 2's bounds from the root images (ids are degree-ordered,
                // so a bound like "above the root's image" cuts a hub row to
                // the few heavier vertices; the degree alone overstates it and
                // every surplus subtask materialises the level-2 set again)
                uint64_t lo = 0, hi = UINT64_MAX;  // exclusive
                const uint32_t img[2] = {v0, v1};
                for (int j = 0; j < 2; ++j) {
                    if ((key.l2 >> j) & 1u) lo = std::max<uint64_t>(lo, (uint64_t)img[j] + 1);
                    if ((key.l2 >> (2 + j)) & 1u) hi = std::min<uint64_t>(hi, img[j]);
                }
                uint64_t bound = UINT64_MAX;
                for (int j = 0; j < 2; ++j) {
                    if (!((key.shape >> j) & 1u)) continue;
                    const uint32_t *rb = h.nbr.data() + h.off[img[j]], *re = h.nbr.data() + h.off[img[j] + 1];
                    const uint32_t *a = lo ? std::lower_bound(rb, re, (uint32_t)std::min<uint64_t>(lo, UINT32_MAX)) : rb;
                    const uint32_t *z = hi <= UINT32_MAX ? std::lower_bound(a, re, (uint32_t)hi) : re;
                    bound = std::min<uint64_t>(bound, (uint64_t)(z - a));
                }
                nch = std::min<uint64_t>(4096, std::max<uint64_t>(1, (bound + span - 1) / span));
            }
            pre[i + 1] = pre[i] + nch;
        }
        SplitArr a;
        a.nsub = pre[ntask];
        HIP_TRY(hipMalloc(&a.d, sizeof(unsigned long long) * (ntask + 1)));
        HIP_TRY(hipMemcpy(a.d, pre.data(), sizeof(unsigned long long) * (ntask + 1), hipMemcpyHostToDevice));
        it = cache->emplace(key, a).first;
    }
    *out = it->second.d;
    *nsub = it->second.nsub;
    return SGE_OK;
}

// Root tasks of each plan: CSR positions tb[i] + j * stride < te[i] (clamped to
// the graph's nnz), as (task_begin, ntask).
struct TaskRanges {
    std::vector<uint64_t> begin, ntask;
    TaskRanges(const sge_graph *g, uint32_t n, const uint64_t *tb, const uint64_t *te, uint64_t stride)
        : begin(n), ntask(n) {
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t e = std::min<uint64_t>(te[i], g->h.nnz), b = std::min<uint64_t>(tb[i], e);
            begin[i] = b;
            ntask[i] = (e - b + stride - 1) / stride;
        }
    }
};

// Validate and upload plans with their task ranges (and split prefixes); nsub[i]
// = the root subtasks of plan i.
int upload_plans(sge_graph *g, const sge_plan *plans, uint32_t n, bool enumerate, const TaskRanges &tr,
                 uint64_t stride, uint32_t span, std::vector<uint64_t> &nsub,
                 const std::vector<sge::BagTrie> *tries = nullptr) {
    DeviceCtx &c = dctx();
    std::vector<DPlan> dp(n);
    nsub.assign(n, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t tb = tr.begin[i], ntask = tr.ntask[i], te = tb + ntask * stride;
        int rc = validate_plan(plans[i]);
        if (rc) return rc;
        sge_plan p = plans[i];
        if (enumerate) {
            p.mode = SGE_MODE_WCOJ;
            p.tail = 0;
            p.tail_kind = 0;
            p.bag = 0;  // embeddings are enumerated over CSR rows alone
        }
        dp[i] = to_dplan(p);
        for (int b = 0; b < SGE_MAX_BAGS; ++b) {
            if (!dp[i].bagm[b]) continue;
            if (!tries || !(*tries)[i * SGE_MAX_BAGS + b].desc) {
                sge::set_error("plan has a bag but no trie was built for it");
                return SGE_EINVAL;
            }
            dp[i].trie[b] = static_cast<const TrieDesc *>((*tries)[i * SGE_MAX_BAGS + b].desc);
        }
        const unsigned long long *sp = nullptr;
        unsigned long long ns = 0;
        if (p.k >= 2 && ntask) {
            rc = get_split(g, p, tb, te, stride, span, &sp, &ns);
            if (rc) return rc;
        }
        dp[i].split = sp;
        dp[i].nsub = ns;
        dp[i].task_begin = tb;
        dp[i].ntask = ntask;
        nsub[i] = sp ? ns : ntask;
    }
    const BufLayout L = buf_layout(dp.data(), n);
    c.ex_base = L.ex_base;
    c.pre_f = L.pre_f;
    c.term_base = L.term_base;
    c.nbuf = L.nbuf;
    int rc = ensure_buffer((void **)&c.plans, c.plans_cap, sizeof(DPlan) * std::max<uint32_t>(n, 1));
    if (rc) return rc;
    rc = ensure_buffer((void **)&c.out, c.out_cap, sizeof(PlanOut) * std::max<uint32_t>(n, 1));
    if (rc) return rc;
    HIP_TRY(hipMemcpy(c.plans, dp.data(), sizeof(DPlan) * n, hipMemcpyHostToDevice));
    return SGE_OK;
}

uint32_t span_of(const sge_tuning *t) { return (t && t->split_span) ? t->split_span : 32u; }

void fill_stats(const PlanOut &o, double ms, sge_stats *s) {
    if (!s) return;
    s->intersections += o.inter;
    s->bytes += o.bytes;
    s->tasks += o.tasks;
    s->nodes += o.nodes;
    s->wave_cycles += o.cycles;
    s->model_bytes += o.model;
    s->kernel_ms += ms;
}

constexpr uint64_t BAG_TUPLE_CAP = 1ull << 28;  // tuples of one materialised bag relation

// Frees up to two device buffers when the scope ends, on every return path.
struct DevFree {
    void *a = nullptr, *b = nullptr;
    ~DevFree() {
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
    }
};

// Materialise the relation of the plan's bag with level mask B and build its
// trie (nodeJoin over the bag,
// join.cpp:309-421, then buildTrie, join.cpp:1713-1717): count the relation
// with the bag's own plan, enumerate it into exactly that many rows, sort and
// group on the device.  Uses the context's plan/output buffers: call before the
// caller's own plans are uploaded.
int build_bag_trie(sge_graph *g, const sge_plan &plan, uint32_t B, sge::BagTrie &t) {
    auto t0 = std::chrono::steady_clock::now();
    DeviceCtx &c = dctx();
    sge_plan sub;
    sge::bag_subplan(plan, B, sub);
    const uint64_t zero = 0, nnz = g->h.nnz;
    const TaskRanges all(g, 1, &zero, &nnz, 1);
    std::vector<uint64_t> nsub;
    int rc = upload_plans(g, &sub, 1, false, all, 1, span_of(nullptr), nsub);
    if (rc) return rc;
    LaunchCfg cfg;
    rc = launch_cfg(nullptr, false, cfg);
    if (rc) return rc;
    float ms = 0;
    rc = run_launch(g, c.plans, 1, c.out, 1, nsub[0], cfg, false, nullptr, 0, nullptr, &ms);
    if (rc) return rc;
    PlanOut o;
    HIP_TRY(hipMemcpy(&o, c.out, sizeof(o), hipMemcpyDeviceToHost));
    const uint64_t N = o.count;
    if (N > BAG_TUPLE_CAP) {
        sge::set_error("bag relation has " + std::to_string(N) + " tuples, above the cap of " +
                       std::to_string(BAG_TUPLE_CAP) + ": choose another bag");
        return SGE_ERANGE;
    }
    uint32_t *rows = nullptr;
    unsigned long long *d_n = nullptr;
    DevFree fr;
    HIP_TRY(hipMalloc(&rows, sizeof(uint32_t) * std::max<uint64_t>(N * sub.k, 1)));
    fr.a = rows;
    HIP_TRY(hipMalloc(&d_n, sizeof(unsigned long long)));
    fr.b = d_n;
    if (N) {
        rc = upload_plans(g, &sub, 1, true, all, 1, span_of(nullptr), nsub);
        if (rc) return rc;
        LaunchCfg ecfg;
        rc = launch_cfg(nullptr, true, ecfg);
        if (rc) return rc;
        rc = run_launch(g, c.plans, 1, c.out, 1, nsub[0], ecfg, true, rows, N, d_n, &ms);
        if (rc) return rc;
        unsigned long long got = 0;
        HIP_TRY(hipMemcpy(&got, d_n, sizeof(got), hipMemcpyDeviceToHost));
        if (got != N) {
            sge::set_error("bag relation: enumerated " + std::to_string(got) + " tuples, counted " + std::to_string(N));
            return SGE_EDEVICE;
        }
    }
    rc = sge::trie_from_rows(rows, N, sub.k, g->h.n, c.stream, t);
    if (rc) return rc;
    TrieDesc d;
    std::memset(&d, 0, sizeof(d));
    d.val = t.val;
    d.child = t.child;
    for (int j = 0; j <= MAXK; ++j) d.vb[j] = t.vb[j];
    d.depth = t.depth;
    HIP_TRY(hipMalloc(&t.desc, sizeof(TrieDesc)));
    HIP_TRY(hipMemcpy(t.desc, &d, sizeof(d), hipMemcpyHostToDevice));
    t.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return SGE_OK;
}

// Tries of every bag of every plan: entry i * SGE_MAX_BAGS + b holds bag b of
// plan i (entries of absent bags stay empty); ms_each[i] sums plan i's bags.
struct TrieSet {
    std::vector<sge::BagTrie> t;
    ~TrieSet() {
        for (auto &x : t) sge::trie_free(x);
    }
    int build(sge_graph *g, const sge_plan *plans, uint32_t n, double *ms_total, std::vector<double> *ms_each) {
        t.assign((size_t)n * SGE_MAX_BAGS, sge::BagTrie());
        if (ms_each) ms_each->assign(n, 0.0);
        for (uint32_t i = 0; i < n; ++i) {
            if (!plans[i].bag) continue;
            int rc = validate_plan(plans[i]);
            if (rc) return rc;
            for (int b = 0; b < SGE_MAX_BAGS; ++b) {
                const uint32_t B = (plans[i].bag >> (8 * b)) & 0xffu;
                if (!B) continue;
                sge::BagTrie &x = t[(size_t)i * SGE_MAX_BAGS + b];
                rc = build_bag_trie(g, plans[i], B, x);
                if (rc) return rc;
                if (ms_total) *ms_total += x.build_ms;
                if (ms_each) (*ms_each)[i] += x.build_ms;
            }
        }
        return SGE_OK;
    }
};

}  // namespace

extern "C" void sge_tuning_default(sge_tuning *t) {
    if (!t) return;
    std::memset(t, 0, sizeof(*t));
    t->block_threads = 256;
    t->chunk = 4;
    t->stage_max = 0;
    t->blocks_per_cu = 0;
}

extern "C" int sge_graph_to_device(sge_graph *g) {
    SGE_API_BEGIN
    if (!g) {
        sge::set_error("sge_graph_to_device: null graph");
        return SGE_EINVAL;
    }
    std::lock_guard<std::mutex> lk(dmutex());
    int rc = ensure_ctx();
    if (rc) return rc;
    if (g->d_off) return SGE_OK;
    const sge::HostGraph &h = g->h;
    std::vector<uint32_t> src(h.nnz);
    for (uint32_t v = 0; v < h.n; ++v)
        for (uint32_t e = h.off[v]; e < h.off[v + 1]; ++e) src[e] = v;
    HIP_TRY(hipMalloc(&g->d_off, sizeof(uint32_t) * ((size_t)h.n + 1)));
    HIP_TRY(hipMalloc(&g->d_nbr, sizeof(uint32_t) * std::max<uint64_t>(h.nnz, 1)));
    HIP_TRY(hipMalloc(&g->d_src, sizeof(uint32_t) * std::max<uint64_t>(h.nnz, 1)));
    HIP_TRY(hipMemcpy(g->d_off, h.off.data(), sizeof(uint32_t) * ((size_t)h.n + 1), hipMemcpyHostToDevice));
    if (h.nnz) {
        HIP_TRY(hipMemcpy(g->d_nbr, h.nbr.data(), sizeof(uint32_t) * h.nnz, hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(g->d_src, src.data(), sizeof(uint32_t) * h.nnz, hipMemcpyHostToDevice));
    }
    g->device = dctx().dev;
    return SGE_OK;
    SGE_API_END
}

extern "C" void sge_graph_device_free(sge_graph *g) {
    if (!g) return;
    if (g->d_off) (void)hipFree(g->d_off);
    if (g->d_nbr) (void)hipFree(g->d_nbr);
    if (g->d_src) (void)hipFree(g->d_src);
    if (g->d_skip) (void)hipFree(g->d_skip);
    g->d_off = g->d_nbr = g->d_src = g->d_skip = nullptr;
    free_split_cache(g);
}

// Count plans[i] over its root tasks tb[i] + j * stride < te[i] (one launch per
// plan, or all plans in one launch with SGE_COUNT_BATCHED).
static int count_impl(sge_graph *g, const sge_plan *plans, uint32_t nplans, uint32_t flags, const sge_tuning *tune,
                      const uint64_t *tb, const uint64_t *te, uint64_t stride, uint64_t *counts, sge_stats *stats,
                      sge_stats *per_plan) {
    auto t0 = std::chrono::steady_clock::now();
    if (!g || !plans || !counts || nplans == 0 || !tb || !te) {
        sge::set_error("sge_count: null argument or no plans");
        return SGE_EINVAL;
    }
    int rc = sge_graph_to_device(g);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(dmutex());
    DeviceCtx &c = dctx();
    if (stride == 0) stride = 1;
    const TaskRanges tr(g, nplans, tb, te, stride);
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (per_plan) std::memset(per_plan, 0, sizeof(sge_stats) * nplans);
    TrieSet tries;
    double trie_ms = 0;
    std::vector<double> trie_each;
    rc = tries.build(g, plans, nplans, &trie_ms, &trie_each);
    if (rc) return rc;
    if (stats) stats->trie_ms = trie_ms;
    if (per_plan)
        for (uint32_t i = 0; i < nplans; ++i) per_plan[i].trie_ms = trie_each[i];
    std::vector<uint64_t> nsub;
    rc = upload_plans(g, plans, nplans, false, tr, stride, span_of(tune), nsub, &tries.t);
    if (rc) return rc;
    LaunchCfg cfg;
    rc = launch_cfg(tune, false, cfg);
    if (rc) return rc;
    uint64_t max_sub = 0;
    for (uint64_t x : nsub) max_sub = std::max(max_sub, x);
    std::vector<PlanOut> host(nplans);
    std::vector<float> ms(nplans, 0.f);
    bool any_k1 = false, limited = false;
    for (uint32_t i = 0; i < nplans; ++i) any_k1 = any_k1 || plans[i].k < 2;
    if ((flags & SGE_COUNT_BATCHED) && !any_k1) {
        float t = 0;
        rc = run_launch(g, c.plans, nplans, c.out, stride, max_sub, cfg, false, nullptr, 0, nullptr, &t);
        if (rc == SGE_ELIMIT) limited = true;
        else if (rc) return rc;
        HIP_TRY(hipMemcpy(host.data(), c.out, sizeof(PlanOut) * nplans, hipMemcpyDeviceToHost));
        if (stats) stats->kernel_ms = t;
        for (uint32_t i = 0; i < nplans; ++i) {
            fill_stats(host[i], 0.0, stats);
            if (per_plan) fill_stats(host[i], 0.0, per_plan + i);
        }
    } else {
        for (uint32_t i = 0; i < nplans; ++i) {
            if (plans[i].k < 2) {  // single vertex: n embeddings, no device work
                HIP_TRY(hipMemsetAsync(c.out + i, 0, sizeof(PlanOut), c.stream));
                continue;
            }
            rc = run_launch(g, c.plans + i, 1, c.out + i, stride, nsub[i], cfg, false, nullptr, 0, nullptr, &ms[i]);
            if (rc == SGE_ELIMIT) {  // keep going: the other plans get their own cap
                limited = true;
                continue;
            }
            if (rc) return rc;
        }
        HIP_TRY(hipStreamSynchronize(c.stream));
        HIP_TRY(hipMemcpy(host.data(), c.out, sizeof(PlanOut) * nplans, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < nplans; ++i) {
            fill_stats(host[i], ms[i], stats);
            if (per_plan) fill_stats(host[i], ms[i], per_plan + i);
        }
    }
    for (uint32_t i = 0; i < nplans; ++i) {
        uint64_t cnt = host[i].count;
        if (plans[i].k == 1) cnt = g->h.n;
        counts[i] = cnt * (plans[i].multiplier ? plans[i].multiplier : 1);
    }
    if (stats) stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (limited) {
        sge::set_error("node cap reached: counts are partial (stats cover the work done)");
        return SGE_ELIMIT;
    }
    return SGE_OK;
}

extern "C" int sge_count(sge_graph *g, const sge_plan *plans, uint32_t nplans, uint32_t flags, const sge_tuning *tune,
                         uint64_t *counts, sge_stats *stats, sge_stats *per_plan) {
    SGE_API_BEGIN
    if (!g) {
        sge::set_error("sge_count: null graph");
        return SGE_EINVAL;
    }
    std::vector<uint64_t> tb(nplans, 0), te(nplans, g->h.nnz);
    return count_impl(g, plans, nplans, flags, tune, tb.data(), te.data(), 1, counts, stats, per_plan);
    SGE_API_END
}

extern "C" int sge_count_ranges(sge_graph *g, const sge_plan *plans, uint32_t nplans, const uint64_t *task_begin,
                                const uint64_t *task_end, uint32_t flags, const sge_tuning *tune, uint64_t *counts,
                                sge_stats *stats, sge_stats *per_plan) {
    SGE_API_BEGIN
    for (uint32_t i = 0; plans && i < nplans; ++i)
        if (plans[i].k == 1) {
            sge::set_error("sge_count_ranges: single-vertex patterns have no edge tasks");
            return SGE_EINVAL;
        }
    return count_impl(g, plans, nplans, flags, tune, task_begin, task_end, 1, counts, stats, per_plan);
    SGE_API_END
}

extern "C" int sge_count_range(sge_graph *g, const sge_plan *plan, uint64_t task_begin, uint64_t task_end,
                               const sge_tuning *tune, uint64_t *count, sge_stats *stats) {
    SGE_API_BEGIN
    if (plan && plan->k == 1) {
        sge::set_error("sge_count_range: single-vertex patterns have no edge tasks");
        return SGE_EINVAL;
    }
    return count_impl(g, plan, 1, 0, tune, &task_begin, &task_end, 1, count, stats, nullptr);
    SGE_API_END
}

extern "C" int sge_count_sampled(sge_graph *g, const sge_plan *plans, uint32_t nplans, uint64_t stride,
                                 uint64_t offset, const sge_tuning *tune, uint64_t *counts, sge_stats *per_plan) {
    SGE_API_BEGIN
    if (!g || stride == 0) {
        sge::set_error("sge_count_sampled: null graph or zero stride");
        return SGE_EINVAL;
    }
    for (uint32_t i = 0; i < nplans; ++i)
        if (plans && plans[i].k == 1) {
            sge::set_error("sge_count_sampled: single-vertex patterns have no edge tasks");
            return SGE_EINVAL;
        }
    std::vector<uint64_t> tb(nplans, offset), te(nplans, g->h.nnz);
    return count_impl(g, plans, nplans, 0, tune, tb.data(), te.data(), stride, counts, nullptr, per_plan);
    SGE_API_END
}

extern "C" int sge_num_tasks(const sge_graph *g, const sge_plan *plan, uint64_t *n) {
    SGE_API_BEGIN
    if (!g || !n) {
        sge::set_error("sge_num_tasks: null argument");
        return SGE_EINVAL;
    }
    (void)plan;
    *n = g->h.nnz;
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_enumerate(sge_graph *g, const sge_plan *plan, uint32_t *out, uint64_t cap, uint64_t *n_out) {
    SGE_API_BEGIN
    if (!g || !plan || !n_out || (cap && !out)) {
        sge::set_error("sge_enumerate: bad arguments");
        return SGE_EINVAL;
    }
    if (plan->k < 3) {
        sge::set_error("sge_enumerate: patterns need at least 3 vertices");
        return SGE_EINVAL;
    }
    if (cap > (SIZE_MAX / sizeof(uint32_t)) / SGE_MAX_K) {  // cap * k * 4 bytes must not wrap
        sge::set_error("sge_enumerate: capacity too large");
        return SGE_ERANGE;
    }
    int rc = sge_graph_to_device(g);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(dmutex());
    DeviceCtx &c = dctx();
    std::vector<uint64_t> nsub;
    const uint64_t zero = 0, nnz = g->h.nnz;
    rc = upload_plans(g, plan, 1, true, TaskRanges(g, 1, &zero, &nnz, 1), 1, span_of(nullptr), nsub);
    if (rc) return rc;
    LaunchCfg cfg;
    rc = launch_cfg(nullptr, true, cfg);
    if (rc) return rc;
    uint32_t *d_emb = nullptr;
    unsigned long long *d_n = nullptr;
    DevFree fr;
    HIP_TRY(hipMalloc(&d_emb, (size_t)std::max<uint64_t>(cap, 1) * plan->k * sizeof(uint32_t)));
    fr.a = d_emb;
    HIP_TRY(hipMalloc(&d_n, sizeof(unsigned long long)));
    fr.b = d_n;
    float t = 0;
    rc = run_launch(g, c.plans, 1, c.out, 1, nsub[0], cfg, true, d_emb, cap, d_n, &t);
    if (rc) return rc;
    unsigned long long n = 0;
    HIP_TRY(hipMemcpy(&n, d_n, sizeof(n), hipMemcpyDeviceToHost));
    const uint64_t w = std::min<uint64_t>(n, cap);
    if (w) HIP_TRY(hipMemcpy(out, d_emb, w * plan->k * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < w * plan->k; ++i) out[i] = g->h.perm[out[i]];
    *n_out = n;
    if (n > cap) {
        sge::set_error("sge_enumerate: output capacity exceeded");
        return SGE_ERANGE;
    }
    return SGE_OK;
    SGE_API_END
}

// ------------------------------------------------------------- bag tries
struct sge_trie {
    sge::BagTrie t;
    ~sge_trie() { sge::trie_free(t); }
};

extern "C" int sge_trie_build(sge_graph *g, const sge_plan *plan, sge_trie **out) {
    SGE_API_BEGIN
    if (!g || !plan || !out) {
        sge::set_error("sge_trie_build: null argument");
        return SGE_EINVAL;
    }
    int rc = validate_plan(*plan);
    if (rc) return rc;
    if (!plan->bag) {
        sge::set_error("sge_trie_build: the plan has no bag (sge_plan_set_bag)");
        return SGE_EINVAL;
    }
    rc = sge_graph_to_device(g);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(dmutex());
    sge_trie *t = new sge_trie();
    rc = build_bag_trie(g, *plan, plan->bag & 0xffu, t->t);
    if (rc) {
        delete t;
        return rc;
    }
    *out = t;
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_trie_shape(const sge_trie *t, uint32_t *depth, uint64_t *tuples, uint64_t *nodes, double *build_ms) {
    SGE_API_BEGIN
    if (!t) {
        sge::set_error("sge_trie_shape: null trie");
        return SGE_EINVAL;
    }
    if (depth) *depth = t->t.depth;
    if (tuples) *tuples = t->t.tuples;
    if (nodes)
        for (int j = 0; j < MAXK; ++j) nodes[j] = j < (int)t->t.depth ? t->t.nodes[j] : 0;
    if (build_ms) *build_ms = t->t.build_ms;
    return SGE_OK;
    SGE_API_END
}

extern "C" int sge_trie_copy(const sge_trie *t, uint32_t *values, uint32_t *child) {
    SGE_API_BEGIN
    if (!t || !t->t.depth) {
        sge::set_error("sge_trie_copy: null trie");
        return SGE_EINVAL;
    }
    const uint32_t p = t->t.depth;
    const uint64_t nval = t->t.vb[p], nchild = (uint64_t)t->t.vb[p - 1] + (p - 1);
    if (values && nval) HIP_TRY(hipMemcpy(values, t->t.val, sizeof(uint32_t) * nval, hipMemcpyDeviceToHost));
    if (child && nchild) HIP_TRY(hipMemcpy(child, t->t.child, sizeof(uint32_t) * nchild, hipMemcpyDeviceToHost));
    return SGE_OK;
    SGE_API_END
}

extern "C" void sge_trie_destroy(sge_trie *t) { delete t; }

// ---------------------------------------------------------------- batches
struct sge_batch {
    sge_graph *g = nullptr;  // borrowed: the graph is kept alive until the batch is destroyed
    std::vector<sge_plan> plans;
    TrieSet tries;        // bag tries the plans read (built at creation)
    double trie_ms = 0;
    DPlan *dplans = nullptr;
    PlanOut *dout = nullptr;
    unsigned long long *queue = nullptr;
    uint32_t *scratch = nullptr;
    uint32_t *gbm = nullptr;  // global bitmap regions, all zero between uses
    uint32_t *items[2] = {nullptr, nullptr};
    unsigned long long *icount = nullptr;  // [0], [1] item counts; [2] work-cap counter
    KParams kp;           // root-pass parameters (continuation passes derive from it)
    LaunchCfg cfg;
    int last_out = 0;     // item buffer written by the last captured pass
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// continuation passes inside the graph: a pass that takes the items it writes
// itself leaves nothing to a next one (an empty pass would still cost about a
// millisecond of queue atomics from every wavefront), otherwise eight (an empty
// pass exits at once; sge_batch_run drains anything left)
constexpr int CAPTURED_PASSES_OWN = 0, CAPTURED_PASSES_NEXT = 8;

// Enqueue one continuation pass reading items[in] and writing items[in ^ 1].
static void enqueue_cont_pass(sge_batch *b, int in) {
    KParams kp = b->kp;
    kp.in_items = b->items[in];
    kp.in_count = b->icount + in;
    kp.out_items = b->items[in ^ 1];
    kp.out_count = b->icount + (in ^ 1);
    reset_pass_kernel<<<1, 64, 0, b->stream>>>(b->queue, b->icount + (in ^ 1), waves_of(b->cfg));
    launch_wcoj(false, b->cfg, b->stream, kp);
}

// A batch over plans[i]'s root tasks tb[i] + j * stride < te[i].
static int batch_create(sge_graph *g, const sge_plan *plans, uint32_t nplans, const uint64_t *tb, const uint64_t *te,
                        uint64_t stride, const sge_tuning *tune, sge_batch **out) {
    if (stride == 0) stride = 1;
    if (!g || !plans || !nplans || !out || !tb || !te) {
        sge::set_error("sge_batch_create: bad arguments");
        return SGE_EINVAL;
    }
    int rc = sge_graph_to_device(g);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(dmutex());
    const TaskRanges tr(g, nplans, tb, te, stride);
    std::vector<DPlan> dp(nplans);
    for (uint32_t i = 0; i < nplans; ++i) {
        rc = validate_plan(plans[i]);
        if (rc) return rc;
        if (plans[i].k < 2) {
            sge::set_error("sge_batch_create: patterns need at least 2 vertices");
            return SGE_EINVAL;
        }
        dp[i] = to_dplan(plans[i]);
    }
    TrieSet tries;
    double trie_ms = 0;
    rc = tries.build(g, plans, nplans, &trie_ms, nullptr);
    if (rc) return rc;
    for (uint32_t i = 0; i < nplans; ++i)
        for (int b = 0; b < SGE_MAX_BAGS; ++b)
            if (dp[i].bagm[b]) dp[i].trie[b] = static_cast<const TrieDesc *>(tries.t[(size_t)i * SGE_MAX_BAGS + b].desc);
    LaunchCfg cfg;
    rc = launch_cfg(tune, false, cfg);
    if (rc) return rc;
    uint64_t max_sub = 0;
    for (uint32_t i = 0; i < nplans; ++i) {
        const unsigned long long *sp = nullptr;
        unsigned long long ns = 0;
        if (tr.ntask[i]) {
            rc = get_split(g, plans[i], tr.begin[i], tr.begin[i] + tr.ntask[i] * stride, stride, span_of(tune), &sp,
                           &ns);
            if (rc) return rc;
        }
        dp[i].split = sp;
        dp[i].nsub = ns;
        dp[i].task_begin = tr.begin[i];
        dp[i].ntask = tr.ntask[i];
        max_sub = std::max<uint64_t>(max_sub, sp ? ns : tr.ntask[i]);
    }
    sge_batch *b = new sge_batch();
    b->g = g;
    b->plans.assign(plans, plans + nplans);
    b->tries.t.swap(tries.t);
    b->trie_ms = trie_ms;
    ++g->batches;
    KParams kp;
    std::memset(&kp, 0, sizeof(kp));
    uint32_t bstride = ((g->h.max_degree + 64 + 63) / 64) * 64;
    const BufLayout L = buf_layout(dp.data(), nplans);
    kp.buf_stride = bstride;
    kp.wave_stride = (uint64_t)bstride * L.nbuf;
    set_layout(kp, L.ex_base, L.pre_f, L.term_base);
    size_t waves = (size_t)cfg.grid * (cfg.block / WAVE);
    auto fail = [&](hipError_t e, const char *what) {
        sge::set_error(std::string("HIP error: ") + hipGetErrorString(e) + " at " + what);
        sge_batch_destroy(b);
        return SGE_EDEVICE;
    };
    hipError_t e;
    if ((e = hipMalloc(&b->dplans, sizeof(DPlan) * nplans)) != hipSuccess) return fail(e, "plans");
    if ((e = hipMalloc(&b->dout, sizeof(PlanOut) * nplans)) != hipSuccess) return fail(e, "out");
    if ((e = hipMalloc(&b->queue, sizeof(unsigned long long) * Q_WORDS)) != hipSuccess) return fail(e, "queue");
    if ((e = hipMemset(b->queue, 0, sizeof(unsigned long long) * Q_WORDS)) != hipSuccess) return fail(e, "queue");
    if ((e = hipMalloc(&b->scratch, waves * kp.wave_stride * sizeof(uint32_t))) != hipSuccess) return fail(e, "scratch");
    kp.gbm_words = gbm_words_for(g);
    if ((e = hipMalloc(&b->gbm, waves * (size_t)kp.gbm_words * sizeof(uint32_t))) != hipSuccess) return fail(e, "bitmaps");
    if ((e = hipMemset(b->gbm, 0, waves * (size_t)kp.gbm_words * sizeof(uint32_t))) != hipSuccess)
        return fail(e, "bitmaps");
    for (int q = 0; q < 2; ++q)
        if ((e = hipMalloc(&b->items[q], sizeof(uint32_t) * ITEM_WORDS * ITEM_CAP)) != hipSuccess ||
            (e = hipMemset(b->items[q], 0, sizeof(uint32_t) * ITEM_WORDS * ITEM_CAP)) != hipSuccess)
            return fail(e, "items");
    if ((e = hipMalloc(&b->icount, sizeof(unsigned long long) * 4)) != hipSuccess) return fail(e, "item counts");
    if ((e = hipMemcpy(b->dplans, dp.data(), sizeof(DPlan) * nplans, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "plan upload");
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess) return fail(e, "event");
    if ((e = hipEventCreate(&b->ev1)) != hipSuccess) return fail(e, "event");
    if (int rc = ensure_skip(g, b->stream)) {
        sge_batch_destroy(b);
        return rc;
    }
    kp.off = (const uint32_t *)g->d_off;
    kp.nbr = (const uint32_t *)g->d_nbr;
    kp.src = (const uint32_t *)g->d_src;
    kp.skip = (const uint32_t *)g->d_skip;
    kp.n = g->h.n;
    kp.nnz = g->h.nnz;
    kp.plans = b->dplans;
    kp.nplans = nplans;
    kp.stride = stride;
    kp.chunk = cfg.chunk;
    kp.head = cfg.head;
    kp.nitems = items_per_plan(max_sub, cfg) * nplans;
    kp.queue = b->queue;
    kp.scratch = b->scratch;
    kp.gbm = b->gbm;
    kp.out = b->dout;
    kp.item_cap = ITEM_CAP;
    kp.budget = budget_for(cfg, g);
    kp.split = cfg.split;
    kp.own_items = own_items_for(cfg, g);
    kp.probe_ratio = cfg.probe_ratio;
    kp.bitmap_ratio = cfg.bitmap_ratio;
    kp.gbm_min = cfg.gbm_min;
    kp.node_cap = cfg.node_cap;
    kp.node_ctr = b->icount + 2;
    kp.out_items = b->items[0];
    kp.out_count = b->icount;
    b->kp = kp;
    b->cfg = cfg;
    // capture: resets (fill kernels, no memset nodes), the root pass and a fixed
    // number of continuation passes that read their item counts on the device
    if ((e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess) return fail(e, "capture");
    reset_kernel<<<(nplans + 255) / 256, 256, 0, b->stream>>>(b->dout, nplans, b->queue, nullptr, kp.node_ctr);
    reset_pass_kernel<<<1, 64, 0, b->stream>>>(b->queue, b->icount, waves_of(cfg));
    if (kp.nitems) {
        launch_wcoj(false, cfg, b->stream, kp);
        int in = 0;
        const int npass = kp.own_items ? CAPTURED_PASSES_OWN : CAPTURED_PASSES_NEXT;
        for (int pass = 0; pass < npass; ++pass) {
            enqueue_cont_pass(b, in);
            in ^= 1;
        }
        b->last_out = in;
    } else {
        b->last_out = 0;
    }
    if ((e = hipStreamEndCapture(b->stream, &b->graph)) != hipSuccess) return fail(e, "end capture");
    if ((e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess) return fail(e, "instantiate");
    *out = b;
    return SGE_OK;
}

extern "C" int sge_batch_create(sge_graph *g, const sge_plan *plans, uint32_t nplans, const sge_tuning *tune,
                                sge_batch **out) {
    SGE_API_BEGIN
    if (!g) return batch_create(g, plans, nplans, nullptr, nullptr, 1, tune, out);
    std::vector<uint64_t> tb(nplans, 0), te(nplans, g->h.nnz);
    return batch_create(g, plans, nplans, tb.data(), te.data(), 1, tune, out);
    SGE_API_END
}

extern "C" int sge_batch_create_sampled(sge_graph *g, const sge_plan *plans, uint32_t nplans, uint64_t stride,
                                        uint64_t offset, const sge_tuning *tune, sge_batch **out) {
    SGE_API_BEGIN
    if (!g) return batch_create(g, plans, nplans, nullptr, nullptr, stride, tune, out);
    std::vector<uint64_t> tb(nplans, offset), te(nplans, g->h.nnz);
    return batch_create(g, plans, nplans, tb.data(), te.data(), stride, tune, out);
    SGE_API_END
}

extern "C" int sge_batch_create_ranges(sge_graph *g, const sge_plan *plans, uint32_t nplans, const uint64_t *task_begin,
                                       const uint64_t *task_end, const sge_tuning *tune, sge_batch **out) {
    SGE_API_BEGIN
    return batch_create(g, plans, nplans, task_begin, task_end, 1, tune, out);
    SGE_API_END
}

extern "C" int sge_batch_run(sge_batch *b, uint64_t *counts, sge_stats *stats) {
    SGE_API_BEGIN
    if (!b || !counts) {
        sge::set_error("sge_batch_run: bad arguments");
        return SGE_EINVAL;
    }
    auto t0 = std::chrono::steady_clock::now();
    uint32_t n = (uint32_t)b->plans.size();
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    // items left after the captured passes: drain them eagerly, unless the work
    // cap has stopped the call
    int in = b->last_out;
    bool limited = false;
    for (;;) {
        unsigned long long hc[3] = {0, 0, 0};  // item counts, work spent
        HIP_TRY(hipMemcpyAsync(hc, b->icount, sizeof(hc), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        if (b->kp.node_cap && hc[2] >= b->kp.node_cap) {
            limited = true;
            break;
        }
        if (hc[in] == 0) break;
        enqueue_cont_pass(b, in);
        HIP_TRY(hipGetLastError());
        in ^= 1;
    }
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (int rc = check_item_waits(b->queue)) return rc;
    std::vector<PlanOut> host(n);
    HIP_TRY(hipMemcpy(host.data(), b->dout, sizeof(PlanOut) * n, hipMemcpyDeviceToHost));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        for (auto &o : host) fill_stats(o, 0.0, stats);
        stats->kernel_ms = ms;
        stats->trie_ms = b->trie_ms;
    }
    for (uint32_t i = 0; i < n; ++i) counts[i] = host[i].count * (b->plans[i].multiplier ? b->plans[i].multiplier : 1);
    if (stats) stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (limited) {
        sge::set_error("work cap reached: counts are partial (stats cover the work done)");
        return SGE_ELIMIT;
    }
    return SGE_OK;
    SGE_API_END
}

extern "C" void sge_batch_destroy(sge_batch *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->dplans) (void)hipFree(b->dplans);
    if (b->dout) (void)hipFree(b->dout);
    if (b->queue) (void)hipFree(b->queue);
    if (b->scratch) (void)hipFree(b->scratch);
    if (b->gbm) (void)hipFree(b->gbm);
    for (int q = 0; q < 2; ++q)
        if (b->items[q]) (void)hipFree(b->items[q]);
    if (b->icount) (void)hipFree(b->icount);
    sge_graph *g = b->g;
    delete b;
    // the last batch of a graph its owner already destroyed frees it
    if (g && --g->batches == 0 && g->destroyed) sge_graph_release(g);
}

extern "C" int sge_device_info(int *num_cu, int *num_xcd_hint, char *name, size_t cap) {
    SGE_API_BEGIN
    std::lock_guard<std::mutex> lk(dmutex());
    int rc = ensure_ctx();
    if (rc) return rc;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dctx().dev));
    if (num_cu) *num_cu = prop.multiProcessorCount;
    if (num_xcd_hint) *num_xcd_hint = 8;
    if (name && cap) {
        std::snprintf(name, cap, "%s", prop.gcnArchName);
    }
    return SGE_OK;
    SGE_API_END
}

// ------------------------------------------------------------ calibration
// A plain streaming copy of a known number of bytes, used to scale the
// FETCH_SIZE / WRITE_SIZE counters read for the enumeration kernel.
namespace {
__global__ void __launch_bounds__(256) calib_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                         size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}
}  // namespace

extern "C" int sge_calibration_copy(uint64_t bytes, double *ms) {
    SGE_API_BEGIN
    if (bytes < 16 || bytes % 16) {
        sge::set_error("sge_calibration_copy: bytes must be a positive multiple of 16");
        return SGE_EINVAL;
    }
    std::lock_guard<std::mutex> lk(dmutex());
    int rc = ensure_ctx();
    if (rc) return rc;
    DeviceCtx &c = dctx();
    uint4 *src = nullptr, *dst = nullptr;
    HIP_TRY(hipMalloc(&src, bytes));
    HIP_TRY(hipMalloc(&dst, bytes));
    size_t n = bytes / 16;
    int grid = c.num_cu * 8;
    calib_copy_kernel<<<grid, 256, 0, c.stream>>>(src, dst, n);  // warm (page-in)
    HIP_TRY(hipEventRecord(c.ev0, c.stream));
    calib_copy_kernel<<<grid, 256, 0, c.stream>>>(src, dst, n);
    HIP_TRY(hipEventRecord(c.ev1, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, c.ev0, c.ev1));
    if (ms) *ms = t;
    (void)hipFree(src);
    (void)hipFree(dst);
    return SGE_OK;
    SGE_API_END
}

// ------------------------------------------------------ allocation-free count
// sge_count_noalloc: the batched count with every device buffer the kernels
// need beyond the caller's inputs (the graph's device CSR) and outputs (the
// counts) carved out of one file-scope device array, `scratch`, of
// NOALLOC_SCRATCH_BYTES.  Each GPU the library runs on holds its own copy
// until the process ends; its address on a device is taken once, with
// hipGetSymbolAddress, on the first call with that device current (the
// array's name is a host-side stand-in, never passed to a kernel).  Layout:
//   [0, 64)                   queue, item counts [2], work-cap counter
//   PlanOut[NOALLOC_MAX_PLANS], DPlan[NOALLOC_MAX_PLANS]
//   two continuation-item buffers of NOALLOC_ITEMS items
//   the wavefronts' scratch slabs (as many wavefronts as fit, up to the
//   resident maximum)
// No hipMalloc/hipFree, no library allocation, no memset node: control words
// and counters are cleared by fill kernels.  Every call on a GPU shares that
// GPU's array and leaves the last call's contents in it; calls that overlap on
// one GPU race (this library serialises its own calls with a mutex).
constexpr size_t NOALLOC_SCRATCH_BYTES = 64ull << 20;
constexpr uint32_t NOALLOC_MAX_PLANS = 256;
constexpr uint64_t NOALLOC_ITEMS = 1ull << 12;
constexpr int NOALLOC_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

struct NoallocLayout {
    size_t out, plans, items0, items1, slabs;
};

constexpr size_t align256(size_t x) { return (x + 255) / 256 * 256; }

constexpr NoallocLayout noalloc_layout() {
    NoallocLayout L{};
    L.out = 256;  // the queue counter and the control words first
    L.plans = align256(L.out + sizeof(PlanOut) * NOALLOC_MAX_PLANS);
    L.items0 = align256(L.plans + sizeof(DPlan) * NOALLOC_MAX_PLANS);
    L.items1 = align256(L.items0 + sizeof(uint32_t) * ITEM_WORDS * NOALLOC_ITEMS);
    L.slabs = align256(L.items1 + sizeof(uint32_t) * ITEM_WORDS * NOALLOC_ITEMS);
    return L;
}
static_assert(noalloc_layout().slabs < NOALLOC_SCRATCH_BYTES / 2, "the slabs get most of the array");

// The array's device address on each device, taken once per device at its
// first call (under dmutex()).
unsigned char *noalloc_base(int dev) {
    static unsigned char *ptr[NOALLOC_MAX_DEVICES] = {};
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return nullptr;
    if (!ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            sge::set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return nullptr;
        }
        ptr[dev] = static_cast<unsigned char *>(p);
    }
    return ptr[dev];
}

}  // namespace

extern "C" int sge_count_noalloc(sge_graph *g, const sge_plan *plans, uint32_t nplans, const sge_tuning *tune,
                                 uint64_t *counts, sge_stats *stats) {
    SGE_API_BEGIN
    auto t0 = std::chrono::steady_clock::now();
    if (!g || !plans || !counts || nplans == 0) {
        sge::set_error("sge_count_noalloc: null argument or no plans");
        return SGE_EINVAL;
    }
    if (nplans > NOALLOC_MAX_PLANS) {
        sge::set_error("sge_count_noalloc: at most " + std::to_string(NOALLOC_MAX_PLANS) + " plans per call");
        return SGE_ERANGE;
    }
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (!g->d_off || g->device != dev) {
        sge::set_error("sge_count_noalloc: the graph is not resident on the current device (sge_graph_to_device "
                       "or SGE_GRAPH_BUILD_DEVICE first; uploading it would allocate)");
        return SGE_EINVAL;
    }
    std::lock_guard<std::mutex> lk(dmutex());
    unsigned char *base = noalloc_base(dev);
    if (!base) return SGE_EDEVICE;
    constexpr NoallocLayout L = noalloc_layout();
    // launch shape: the uncapped or capped counting kernel, as many resident
    // wavefronts as the slab region holds
    sge_tuning def;
    sge_tuning_default(&def);
    if (!tune) tune = &def;
    LaunchCfg cfg;
    cfg.block = (int)tune->block_threads;
    if (cfg.block < 64 || cfg.block > 256 || cfg.block % 64) {
        sge::set_error("tuning: block_threads must be 64, 128, 192 or 256");
        return SGE_EINVAL;
    }
    cfg.chunk = tune->chunk ? tune->chunk : 4;
    cfg.budget = tune->donate_budget ? (tune->donate_budget >= (1u << 30) ? 0u : tune->donate_budget) : DONATE_BUDGET_DEFAULT;
    cfg.budget_auto = tune->donate_budget == 0;
    cfg.split = tune->donate_split ? tune->donate_split : DONATE_SPLIT_DEFAULT;
    if (cfg.split > DONATE_SPLIT_MAX) {
        sge::set_error("tuning: donate_split must be at most " + std::to_string(DONATE_SPLIT_MAX));
        return SGE_EINVAL;
    }
    if (tune->pass_items > 2) {
        sge::set_error("tuning: pass_items must be 0, 1 or 2");
        return SGE_EINVAL;
    }
    cfg.pass_mode = tune->pass_items;
    cfg.probe_ratio = tune->stage_max ? tune->stage_max : 2u;
    cfg.bitmap_ratio = tune->bitmap_ratio ? tune->bitmap_ratio : cfg.probe_ratio;
    cfg.node_cap = (unsigned long long)tune->node_cap_k * NODE_STEP;
    cfg.lds = (size_t)(cfg.block / WAVE) * (sizeof(WaveLds) + MAX_TERMS * sizeof(TermLds));
    int num_cu = 0, occ = 0;
    HIP_TRY(hipDeviceGetAttribute(&num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, wcoj_kernel<false, false>, cfg.block, cfg.lds));
    const uint64_t resident = (uint64_t)num_cu * std::max(1, occ) * (cfg.block / WAVE);
    // plans: validated, converted and uploaded one at a time (no host buffers)
    DPlan *dplans = reinterpret_cast<DPlan *>(base + L.plans);
    uint32_t kmax = 2, terms = 1;
    bool explore = false;
    for (uint32_t i = 0; i < nplans; ++i) {
        int rc = validate_plan(plans[i]);
        if (rc) return rc;
        if (plans[i].k < 2 || plans[i].bag) {
            sge::set_error("sge_count_noalloc: plans need >= 2 vertices and no bag (a bag trie is built in "
                           "allocated memory)");
            return SGE_EINVAL;
        }
        DPlan d = to_dplan(plans[i]);
        d.split = nullptr;  // no root-edge split: its prefix arrays would be allocated
        d.nsub = 0;
        d.task_begin = 0;
        d.ntask = g->h.nnz;
        kmax = std::max(kmax, d.k);
        terms = std::max(terms, d.nterms);
        explore = explore || d.mode == SGE_MODE_EXPLORE;
        HIP_TRY(hipMemcpy(dplans + i, &d, sizeof(DPlan), hipMemcpyHostToDevice));
    }
    const uint32_t nbuf = explore ? 2 * kmax + 1 + terms : kmax + terms;
    KParams kp;
    std::memset(&kp, 0, sizeof(kp));
    kp.buf_stride = ((g->h.max_degree + 64 + 63) / 64) * 64;
    kp.wave_stride = (uint64_t)kp.buf_stride * nbuf;
    if (explore)
        set_layout(kp, kmax, 2 * kmax, 2 * kmax + 1);
    else
        set_layout(kp, kmax, kmax, kmax);
    const uint64_t fit = (NOALLOC_SCRATCH_BYTES - L.slabs) / (kp.wave_stride * sizeof(uint32_t));
    if (fit == 0) {
        sge::set_error("sge_count_noalloc: one wavefront's scratch slab (" +
                       std::to_string(kp.wave_stride * sizeof(uint32_t)) + " bytes at max degree " +
                       std::to_string(g->h.max_degree) + ") exceeds the " +
                       std::to_string((NOALLOC_SCRATCH_BYTES - L.slabs) >> 20) + " MiB slab region");
        return SGE_ERANGE;
    }
    const uint64_t waves = std::min(resident, fit);
    if (waves < (uint64_t)(cfg.block / WAVE)) cfg.block = WAVE;  // fewer wavefronts than one workgroup holds
    cfg.lds = (size_t)(cfg.block / WAVE) * (sizeof(WaveLds) + MAX_TERMS * sizeof(TermLds));
    cfg.grid = (int)(waves / (uint64_t)(cfg.block / WAVE));
    cfg.head = tune->head_items ? (tune->head_items >= (1u << 30) ? 0u : tune->head_items) : (uint32_t)waves;
    unsigned long long *queue = reinterpret_cast<unsigned long long *>(base);  // [0] queue counter
    unsigned long long *ctl = queue + 1;  // [0..1] continuation item counts, [2] work-cap counter
    PlanOut *dout = reinterpret_cast<PlanOut *>(base + L.out);
    kp.off = (const uint32_t *)g->d_off;
    kp.nbr = (const uint32_t *)g->d_nbr;
    kp.src = (const uint32_t *)g->d_src;
    kp.skip = (const uint32_t *)g->d_skip;  // built by an earlier launch, or none: this call allocates nothing
    kp.n = g->h.n;
    kp.nnz = g->h.nnz;
    kp.plans = dplans;
    kp.nplans = nplans;
    kp.stride = 1;
    kp.chunk = cfg.chunk;
    kp.head = cfg.head;
    kp.nitems = items_per_plan(g->h.nnz, cfg) * nplans;
    kp.queue = queue;
    kp.scratch = reinterpret_cast<uint32_t *>(base + L.slabs);
    kp.gbm = nullptr;  // large F sets are probed by search instead of a global bitmap
    kp.out = dout;
    kp.item_cap = NOALLOC_ITEMS;
    kp.budget = budget_for(cfg, g);
    kp.split = cfg.split;
    kp.own_items = own_items_for(cfg, g);
    kp.probe_ratio = cfg.probe_ratio;
    kp.bitmap_ratio = cfg.bitmap_ratio;
    kp.node_cap = cfg.node_cap;
    kp.node_ctr = ctl + 2;
    uint32_t *items[2] = {reinterpret_cast<uint32_t *>(base + L.items0), reinterpret_cast<uint32_t *>(base + L.items1)};
    hipStream_t st = nullptr;  // the null stream: nothing to create
    reset_kernel<<<(nplans + 255) / 256, 256, 0, st>>>(dout, nplans, queue, nullptr, kp.node_ctr);
    reset_pass_kernel<<<1, 64, 0, st>>>(queue, ctl, waves_of(cfg));
    HIP_TRY(hipGetLastError());
    int cur = 0;
    bool limited = false;
    for (int pass = 0;; ++pass) {
        if (pass > 0) {
            unsigned long long hc[3] = {0, 0, 0};  // item counts, work spent
            HIP_TRY(hipMemcpy(hc, ctl, sizeof(hc), hipMemcpyDeviceToHost));
            if (kp.node_cap && hc[2] >= kp.node_cap) {
                limited = true;
                break;
            }
            if (hc[cur] == 0) break;
            kp.in_items = items[cur];
            kp.in_count = ctl + cur;
            cur ^= 1;
            reset_pass_kernel<<<1, 64, 0, st>>>(queue, ctl + cur, waves_of(cfg));
        }
        kp.out_items = items[cur];
        kp.out_count = ctl + cur;
        if (pass == 0 && kp.nitems == 0) break;
        launch_wcoj(false, cfg, st, kp);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(st));
    if (int rc = check_item_waits(queue)) return rc;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    for (uint32_t i = 0; i < nplans; ++i) {
        PlanOut o;
        HIP_TRY(hipMemcpy(&o, dout + i, sizeof(o), hipMemcpyDeviceToHost));
        counts[i] = o.count * (plans[i].multiplier ? plans[i].multiplier : 1);
        fill_stats(o, 0.0, stats);
    }
    if (stats) {
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        stats->kernel_ms = stats->wall_ms;
    }
    if (limited) {
        sge::set_error("work cap reached: counts are partial (stats cover the work done)");
        return SGE_ELIMIT;
    }
    return SGE_OK;
    SGE_API_END
}
