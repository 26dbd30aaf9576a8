/*
 * sge_oracle.c -- CPU restatement of the reference's WCOJ counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or the timed
 * CPU baseline); the product path never calls it.
 *
 * What each function restates (reference file:line):
 *   oracle_set_begin_pos  utility/utils.cpp:581-597   (setBeginPos: first index with value > max)
 *   oracle_set_end_pos    utility/utils.cpp:599-612   (setEndPos: first index with value >= min)
 *   oracle_intersect      utility/compute_set_intersection.cpp:10-45, 89-147
 *                         (ComputeCandidates: galloping when max/min > 50, merge otherwise)
 *   oracle_choosec        utility/utils.cpp:80-94     (choosec, wrapping uint64 numerator)
 *   count_last            relation/relation.cpp:301-338 (TrieLevel::numResults: |clamped range|
 *                         minus matched pattern vertices whose image lies in the range)
 *                         + matching/join.cpp:1532-1537 (IEP: choosec over the symmetric group)
 *                         + independent tails (no reference counterpart): inclusion-exclusion
 *                         over set partitions from explicitly intersected member sets
 *   dfs                   matching/join.cpp:137-178, 309-421 (generateCandidates + nodeJoin DFS
 *                         with visited[] injectivity), parallel over root edges like
 *                         matching/parallel.cpp:1414-1458 (schedule(dynamic), + reduction)
 *   oracle_count_brute    no reference counterpart: all injective homomorphisms by plain
 *                         backtracking, the independent check for tiny graphs.
 *   bag_trie_build        matching/join.cpp:309-421 (nodeJoin: the bag's tuples), join.cpp:1713-1717
 *                         + relation/relation.cpp:340-379 (buildTrie: std::sort, then recursive
 *                         grouping by depth), flattened depth by depth (oracle_bag_trie)
 *   candidates (bag)      matching/join.cpp:1202-1300 (globalJoin: a level joined with a bag reads
 *                         the bag's current trie level in place of the rows of its bag parents and
 *                         descends into the matched value's children, join.cpp:244-264 setIters);
 *                         a second bag (plan.bag bits 8-15) is joined the same way through its
 *                         own trie, as the reference joins every bag of a multi-bag plan
 *
 * Plans are the flat sge_plan of include/sge.h: level i matches pattern vertex
 * order[i]; parent/lower/upper/check masks are over levels.
 *
 * Parity pins (tests/test_oracle.py): golden counts written by the reference
 * binary built from its own sources (tests/golden/*.json, including every
 * connected class on 4-6 vertices over two graphs; brute force where the binary
 * crashes or its -u mode disagrees with its WCOJ mode), K_n closed forms
 * n!/(n-k)!/|Aut|, brute force on random small graphs, the diamond formula
 * sum over edges of C(common neighbours, 2), mode/order/relabelling invariance,
 * std-library set operations for the primitives, a hand-computed wrap case
 * for choosec, and, for the bag joins, brute-force bag relations and one- and
 * two-bag counts equal to the plain join's and the golden counts.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/sge.h"

#define BIN_THRESHOLD 32 /* BINARY_SEARCH_THRESHOLD, config/config.h:36 */

uint32_t oracle_set_begin_pos(const uint32_t *a, uint32_t count, uint32_t max_compared) {
    if (count <= BIN_THRESHOLD) {
        uint32_t j;
        for (j = 0; j < count; ++j)
            if (a[j] > max_compared) return j;
        return count;
    }
    uint32_t lo = 0, hi = count;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (a[mid] <= max_compared) lo = mid + 1; else hi = mid;
    }
    return lo;
}

uint32_t oracle_set_end_pos(const uint32_t *a, uint32_t count, uint32_t min_compared) {
    if (count <= BIN_THRESHOLD) {
        uint32_t j;
        for (j = 0; j < count; ++j)
            if (a[j] >= min_compared) return j;
        return count;
    }
    uint32_t lo = 0, hi = count;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (a[mid] < min_compared) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* first index in [lo, hi) with a[idx] >= key */
static uint32_t lower_bound_u32(const uint32_t *a, uint32_t lo, uint32_t hi, uint32_t key) {
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (a[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* galloping from position lo: exponential probe then binary search */
static uint32_t gallop(const uint32_t *a, uint32_t lo, uint32_t n, uint32_t key) {
    uint32_t step = 1, hi = lo;
    while (hi < n && a[hi] < key) {
        lo = hi + 1;
        hi += step;
        step <<= 1;
    }
    if (hi > n) hi = n;
    return lower_bound_u32(a, lo, hi, key);
}

uint32_t oracle_intersect(const uint32_t *a, uint32_t na, const uint32_t *b, uint32_t nb, uint32_t *out) {
    uint32_t cnt = 0;
    if (na == 0 || nb == 0) return 0;
    if (na > nb) {
        const uint32_t *t = a; a = b; b = t;
        uint32_t tn = na; na = nb; nb = tn;
    }
    if ((double)nb / (double)na > 50.0) {
        uint32_t pos = 0, i;
        for (i = 0; i < na && pos < nb; ++i) {
            pos = gallop(b, pos, nb, a[i]);
            if (pos < nb && b[pos] == a[i]) {
                if (out) out[cnt] = a[i];
                ++cnt;
                ++pos;
            }
        }
        return cnt;
    }
    uint32_t i = 0, j = 0;
    while (i < na && j < nb) {
        if (a[i] < b[j]) ++i;
        else if (a[i] > b[j]) ++j;
        else {
            if (out) out[cnt] = a[i];
            ++cnt;
            ++i;
            ++j;
        }
    }
    return cnt;
}

uint64_t oracle_choosec(uint64_t n, int k) {
    if ((uint64_t)k > n) return 0;
    if (k == 0) return 1;
    uint64_t nom = n, denom = 1;
    int i;
    for (i = 1; i < k; ++i) {
        nom *= (n - (uint64_t)i);
        denom *= (uint64_t)(i + 1);
    }
    return nom / denom;
}

/* Bag relation trie, depth-major: depth j's values are val[vb[j] .. vb[j+1]);
 * the children of the depth-j value at index a are val[child[a + j] .. child[a + j + 1]). */
typedef struct {
    uint32_t p;
    uint32_t vb[SGE_MAX_K + 1];
    uint32_t *val, *child;
} otrie_t;

/* A plan holds up to two materialised bags: level masks in bits 0-7 (bag 0) and
 * 8-15 (bag 1) of pl->bag. */
#define NBAGS 2
static uint32_t bag_levels(const sge_plan *pl, int b) { return (pl->bag >> (8 * b)) & 0xffu; }
static uint32_t bag_all(const sge_plan *pl) { return (pl->bag | (pl->bag >> 8)) & 0xffu; }

typedef struct {
    const sge_plan *pl;
    const otrie_t *trie[NBAGS];       /* each bag's trie (NULL: no such bag) */
    uint32_t tpos[NBAGS][SGE_MAX_K];  /* bag levels: value index of phi[level] in the bag's trie */
    const uint32_t *off, *nbr;
    uint32_t n, maxdeg;
    uint32_t phi[SGE_MAX_K];
    uint32_t *buf[SGE_MAX_K];  /* candidate set per level */
    uint32_t len[SGE_MAX_K];
    uint32_t *tmp, *tmp2;
    uint32_t tmpcap;           /* elements tmp / tmp2 hold, less one */
    uint32_t *emb;             /* enumeration output (may be NULL) */
    uint64_t emb_cap, *emb_n;
    uint64_t intersections;
    int no_reuse;
    uint32_t *ie[3];           /* independent-tail member sets */
} ctx_t;

/* clamp bounds of level d from the images of its lower/upper levels */
static void bounds(const ctx_t *c, int d, uint32_t *lo_excl, int *has_lo, uint32_t *hi_excl, int *has_hi) {
    const sge_plan *pl = c->pl;
    uint32_t mx = 0, mn = 0xffffffffu;
    int j;
    *has_lo = 0;
    *has_hi = 0;
    for (j = 0; j < d; ++j) {
        if ((pl->lower[d] >> j) & 1u) {
            if (!*has_lo || c->phi[j] > mx) mx = c->phi[j];
            *has_lo = 1;
        }
        if ((pl->upper[d] >> j) & 1u) {
            if (!*has_hi || c->phi[j] < mn) mn = c->phi[j];
            *has_hi = 1;
        }
    }
    *lo_excl = mx;
    *hi_excl = mn;
}

/* Trie list of level d in bag bi: the root values for the bag's first level,
 * else the children of the bag's previous level's matched value. */
static const uint32_t *bag_list(const ctx_t *c, int bi, int d, uint32_t *len, uint32_t *base) {
    const otrie_t *t = c->trie[bi];
    uint32_t below = bag_levels(c->pl, bi) & ((1u << d) - 1u), b, e;
    if (below == 0) {
        b = t->vb[0];
        e = t->vb[1];
    } else {
        uint32_t j = (uint32_t)__builtin_popcount(below), a = c->tpos[bi][31 - __builtin_clz(below)];
        b = t->child[a + j - 1];
        e = t->child[a + j];
    }
    *len = e - b;
    *base = b;
    return t->val + b;
}

/* Record the trie position of phi[d] for level d of bag bi; 0 when it is absent. */
static int bag_enter(ctx_t *c, int bi, int d) {
    uint32_t len, base;
    const uint32_t *a = bag_list(c, bi, d, &len, &base);
    uint32_t p = lower_bound_u32(a, 0, len, c->phi[d]);
    c->tpos[bi][d] = base + p;
    return p < len && a[p] == c->phi[d];
}

/* reuse is off for bag levels and for levels that would reuse a bag level's set */
static int reuse_of(const ctx_t *c, int d) {
    const sge_plan *pl = c->pl;
    const uint32_t all = bag_all(pl);
    int r = pl->reuse[d];
    if (all && r >= 0 && (((all >> d) & 1u) || ((all >> r) & 1u))) r = -1;
    return r;
}

/* Candidate set of level d (ignoring the IEP group's internal chain when
 * `group_base` >= 0: members of the group beyond group_base are not matched yet).
 * Returns the size; elements in c->buf[d] when materialise != 0. */
static uint32_t candidates(ctx_t *c, int d, int materialise, uint32_t lower_mask, uint32_t upper_mask,
                           uint32_t check_mask) {
    const sge_plan *pl = c->pl;
    uint32_t lists = pl->parent[d];
    const uint32_t *src = NULL;
    uint32_t srclen = 0;
    int j;
    int r = c->no_reuse ? -1 : reuse_of(c, d);
    uint32_t mx = 0, mn = 0xffffffffu;
    int has_lo = 0, has_hi = 0;
    for (j = 0; j < d; ++j) {
        if ((lower_mask >> j) & 1u) { if (!has_lo || c->phi[j] > mx) mx = c->phi[j]; has_lo = 1; }
        if ((upper_mask >> j) & 1u) { if (!has_hi || c->phi[j] < mn) mn = c->phi[j]; has_hi = 1; }
    }
    if (pl->mode != SGE_MODE_EXPLORE && r >= 0) {
        src = c->buf[r];
        srclen = c->len[r];
        lists &= ~pl->parent[r];
    }
    /* a bag's trie list satisfies every edge, bound and check among the bag's
     * levels; rows of bag parents are not read (reuse is off here); a level in
     * two bags intersects both lists */
    const uint32_t *tl2 = NULL;
    uint32_t tl2len = 0;
    for (int bi = 0; bi < NBAGS; ++bi) {
        const uint32_t bm = bag_levels(pl, bi);
        if (!c->trie[bi] || !((bm >> d) & 1u)) continue;
        uint32_t base, len;
        const uint32_t *l = bag_list(c, bi, d, &len, &base);
        if (!src) {
            src = l;
            srclen = len;
        } else {
            tl2 = l;
            tl2len = len;
        }
        if (bm & ((1u << d) - 1u)) lists &= ~bm;
    }
    /* generateCandidates: clamp every list, then chain pairwise intersections */
    uint32_t *cur = c->tmp, *nxt = c->tmp2;
    uint32_t curlen = 0;
    int first = 1;
    if (src) {
        uint32_t b = has_lo ? oracle_set_begin_pos(src, srclen, mx) : 0;
        uint32_t e = has_hi ? oracle_set_end_pos(src, srclen, mn) : srclen;
        if (e < b) e = b;
        memcpy(cur, src + b, (size_t)(e - b) * 4);
        curlen = e - b;
        first = 0;
    }
    if (tl2) {
        uint32_t b = has_lo ? oracle_set_begin_pos(tl2, tl2len, mx) : 0;
        uint32_t e = has_hi ? oracle_set_end_pos(tl2, tl2len, mn) : tl2len;
        if (e < b) e = b;
        curlen = oracle_intersect(cur, curlen, tl2 + b, e - b, nxt);
        uint32_t *t = cur; cur = nxt; nxt = t;
        c->intersections++;
    }
    for (j = 0; j < d; ++j) {
        if (!((lists >> j) & 1u)) continue;
        uint32_t v = c->phi[j];
        const uint32_t *row = c->nbr + c->off[v];
        uint32_t deg = c->off[v + 1] - c->off[v];
        uint32_t b = has_lo ? oracle_set_begin_pos(row, deg, mx) : 0;
        uint32_t e = has_hi ? oracle_set_end_pos(row, deg, mn) : deg;
        if (e < b) e = b;
        if (first) {
            memcpy(cur, row + b, (size_t)(e - b) * 4);
            curlen = e - b;
            first = 0;
        } else {
            curlen = oracle_intersect(cur, curlen, row + b, e - b, nxt);
            uint32_t *t = cur; cur = nxt; nxt = t;
            c->intersections++;
        }
        if (curlen == 0) break;
    }
    /* injectivity: drop images of already matched, non-adjacent, unclamped levels */
    uint32_t out = 0, i;
    uint32_t *dst = materialise ? c->buf[d] : NULL;
    for (i = 0; i < curlen; ++i) {
        uint32_t x = cur[i];
        int bad = 0;
        for (j = 0; j < d && !bad; ++j)
            if (((check_mask >> j) & 1u) && c->phi[j] == x) bad = 1;
        if (bad) continue;
        if (dst) dst[out] = x;
        ++out;
    }
    if (materialise) c->len[d] = out;
    return out;
}

static void emit(ctx_t *c, int upto) {
    /* record phi (by pattern vertex) for levels < upto plus every candidate of level upto */
    const sge_plan *pl = c->pl;
    uint32_t i;
    for (i = 0; i < c->len[upto]; ++i) {
        uint64_t slot;
#pragma omp atomic capture
        slot = (*c->emb_n)++;
        if (slot >= c->emb_cap) continue;
        uint32_t *row = c->emb + slot * pl->k;
        int j;
        for (j = 0; j < upto; ++j) row[pl->order[j]] = c->phi[j];
        row[pl->order[upto]] = c->buf[upto][i];
    }
}

static uint64_t dfs(ctx_t *c, int d) {
    const sge_plan *pl = c->pl;
    int k = (int)pl->k;
    int tail = pl->tail >= 2 ? (int)pl->tail : 0;
    int last = tail ? k - tail : k - 1;
    if (d == last) {
        if (c->emb) {
            candidates(c, d, 1, pl->lower[d], pl->upper[d], pl->check[d]);
            emit(c, d);
            return c->len[d];
        }
        if (tail && pl->tail_kind == 2) {
            /* independent tail: materialise each member's set (bounds/checks from
             * the levels below the tail), then inclusion-exclusion over set
             * partitions of the tail, from explicit intersections */
            uint32_t below = (1u << d) - 1, nm[3], t;
            for (t = 0; t < (uint32_t)tail; ++t) {
                int lev = d + (int)t;
                c->no_reuse = 1;
                uint32_t n = candidates(c, lev, 1, pl->lower[lev] & below, pl->upper[lev] & below,
                                        pl->check[lev] & below);
                c->no_reuse = 0;
                memcpy(c->ie[t], c->buf[lev], (size_t)n * 4);
                nm[t] = n;
            }
            if (tail == 2) {
                uint64_t n01 = oracle_intersect(c->ie[0], nm[0], c->ie[1], nm[1], NULL);
                return (uint64_t)nm[0] * nm[1] - n01;
            }
            uint64_t n01 = oracle_intersect(c->ie[0], nm[0], c->ie[1], nm[1], c->tmp);
            uint64_t n012 = oracle_intersect(c->tmp, (uint32_t)n01, c->ie[2], nm[2], NULL);
            uint64_t n02 = oracle_intersect(c->ie[0], nm[0], c->ie[2], nm[2], NULL);
            uint64_t n12 = oracle_intersect(c->ie[1], nm[1], c->ie[2], nm[2], NULL);
            return (uint64_t)nm[0] * nm[1] * nm[2] - n01 * nm[2] - n02 * nm[1] - n12 * nm[0] + 2 * n012;
        }
        if (tail) {
            /* group members share one candidate set; their chain is counted by choosec */
            uint32_t below = (1u << d) - 1;
            uint32_t n = candidates(c, d, 0, pl->lower[d] & below, pl->upper[d] & below, pl->check[d] & below);
            return oracle_choosec(n, tail);
        }
        return candidates(c, d, 0, pl->lower[d], pl->upper[d], pl->check[d]);
    }
    candidates(c, d, 1, pl->lower[d], pl->upper[d], pl->check[d]);
    uint64_t sum = 0;
    uint32_t i;
    int descend[NBAGS], bi;
    for (bi = 0; bi < NBAGS; ++bi) {
        const uint32_t bm = bag_levels(pl, bi);
        descend[bi] = c->trie[bi] && ((bm >> d) & 1u) && (bm >> (d + 1));
    }
    for (i = 0; i < c->len[d]; ++i) {
        c->phi[d] = c->buf[d][i];
        for (bi = 0; bi < NBAGS; ++bi)
            if (descend[bi]) bag_enter(c, bi, d);
        sum += dfs(c, d + 1);
    }
    return sum;
}

/* Root edge (phi[0], phi[1]) of a plan with bags: both ends must lie on a path
 * of every bag's trie that holds them. */
static int bag_root_ok(ctx_t *c) {
    int bi;
    for (bi = 0; bi < NBAGS; ++bi) {
        const uint32_t bm = bag_levels(c->pl, bi);
        if (!c->trie[bi]) continue;
        if ((bm & 1u) && !bag_enter(c, bi, 0)) return 0;
        if ((bm & 2u) && !bag_enter(c, bi, 1)) return 0;
    }
    return 1;
}

static int ctx_init(ctx_t *c, const sge_plan *pl, const uint32_t *off, const uint32_t *nbr, uint32_t n) {
    uint32_t v, i;
    memset(c, 0, sizeof(*c));
    c->pl = pl;
    c->off = off;
    c->nbr = nbr;
    c->n = n;
    c->maxdeg = 0;
    for (v = 0; v < n; ++v)
        if (off[v + 1] - off[v] > c->maxdeg) c->maxdeg = off[v + 1] - off[v];
    for (i = 0; i < SGE_MAX_K; ++i) {
        c->buf[i] = (uint32_t *)malloc(((size_t)c->maxdeg + 1) * 4);
        if (!c->buf[i]) return -1;
    }
    c->tmp = (uint32_t *)malloc(((size_t)c->maxdeg + 1) * 4);
    c->tmp2 = (uint32_t *)malloc(((size_t)c->maxdeg + 1) * 4);
    c->tmpcap = c->maxdeg;
    for (i = 0; i < 3; ++i) {
        c->ie[i] = (uint32_t *)malloc(((size_t)c->maxdeg + 1) * 4);
        if (!c->ie[i]) return -1;
    }
    return (c->tmp && c->tmp2) ? 0 : -1;
}

/* Attach bag bi's trie: the working lists must also hold the trie's root values
 * (up to n), which the bag's first level copies before intersecting. */
static void ctx_set_trie(ctx_t *c, int bi, const otrie_t *t) {
    c->trie[bi] = t;
    if (t->vb[1] > c->tmpcap) {
        c->tmpcap = t->vb[1];
        c->tmp = (uint32_t *)realloc(c->tmp, ((size_t)c->tmpcap + 1) * 4);
        c->tmp2 = (uint32_t *)realloc(c->tmp2, ((size_t)c->tmpcap + 1) * 4);
    }
}

static void ctx_free(ctx_t *c) {
    int i;
    for (i = 0; i < SGE_MAX_K; ++i) free(c->buf[i]);
    for (i = 0; i < 3; ++i) free(c->ie[i]);
    free(c->tmp);
    free(c->tmp2);
}

/* source vertex of CSR position e */
static uint32_t src_of(const uint32_t *off, uint32_t n, uint64_t e) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        uint32_t mid = lo + (hi - lo) / 2;
        if ((uint64_t)off[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

uint64_t oracle_enumerate(const sge_plan *pl, const uint32_t *off, const uint32_t *nbr, uint32_t n, uint32_t *emb,
                          uint64_t cap);
uint64_t oracle_count(const sge_plan *pl, const uint32_t *off, const uint32_t *nbr, uint32_t n, int threads);

/* ---- materialised bag: relation and trie */

/* The bag's own plan (nodeJoin over the bag's attributes): the levels of mask
 * B in plan order, masks restricted to the bag, column i = bag level i. */
static void bag_plan(const sge_plan *pl, uint32_t B, sge_plan *sub) {
    int lev[SGE_MAX_K], p = 0, i, j;
    memset(sub, 0, sizeof(*sub));
    for (i = 0; i < (int)pl->k; ++i)
        if ((B >> i) & 1u) lev[p++] = i;
    sub->k = (uint32_t)p;
    sub->mode = SGE_MODE_WCOJ;
    sub->multiplier = 1;
    for (i = 0; i < SGE_MAX_K; ++i) sub->reuse[i] = -1;
    for (i = 0; i < p; ++i) {
        sub->order[i] = (uint32_t)i;
        for (j = 0; j < i; ++j) {
            uint32_t bit = 1u << lev[j];
            if (pl->parent[lev[i]] & bit) sub->parent[i] |= 1u << j;
            if (pl->lower[lev[i]] & bit) sub->lower[i] |= 1u << j;
            if (pl->upper[lev[i]] & bit) sub->upper[i] |= 1u << j;
            if (pl->check[lev[i]] & bit) sub->check[i] |= 1u << j;
        }
    }
}

static uint32_t g_sort_p; /* row width for cmp_rows (trie builds are not concurrent) */

static int cmp_rows(const void *a, const void *b) {
    const uint32_t *x = (const uint32_t *)a, *y = (const uint32_t *)b;
    uint32_t j;
    for (j = 0; j < g_sort_p; ++j)
        if (x[j] != y[j]) return x[j] < y[j] ? -1 : 1;
    return 0;
}

typedef struct {
    uint32_t *v[SGE_MAX_K], *fc[SGE_MAX_K]; /* per depth: node values, first child (depth-local) */
    uint32_t len[SGE_MAX_K];
} groups_t;

/* buildTrieFromSortedMatchesBatch: rows [start, end) share their first `depth`
 * columns; one node per run of equal column-`depth` values, its children built
 * from the run; the deepest column keeps every row. */
static void group_rows(const uint32_t *rows, uint32_t p, uint64_t start, uint64_t end, uint32_t depth, groups_t *g) {
    uint64_t i = start;
    if (depth == p - 1) {
        for (; i < end; ++i) g->v[depth][g->len[depth]++] = rows[i * p + depth];
        return;
    }
    while (i < end) {
        uint32_t key = rows[i * p + depth];
        uint64_t g0 = i;
        while (i < end && rows[i * p + depth] == key) ++i;
        uint32_t a = g->len[depth]++;
        g->v[depth][a] = key;
        g->fc[depth][a] = g->len[depth + 1];
        group_rows(rows, p, g0, i, depth + 1, g);
    }
}

/* Build the trie of the bag with level mask B over the CSR; 0 on success. */
static int bag_trie_build(const sge_plan *pl, uint32_t B, const uint32_t *off, const uint32_t *nbr, uint32_t n,
                          otrie_t *t) {
    sge_plan sub;
    groups_t g;
    uint32_t j, p;
    memset(t, 0, sizeof(*t));
    memset(&g, 0, sizeof(g));
    bag_plan(pl, B, &sub);
    p = sub.k;
    if (p < 3) return -1;
    uint64_t N = oracle_count(&sub, off, nbr, n, 0);
    uint32_t *rows = (uint32_t *)malloc((size_t)(N ? N : 1) * p * 4);
    if (!rows) return -1;
    if (oracle_enumerate(&sub, off, nbr, n, rows, N) != N) {
        free(rows);
        return -1;
    }
    g_sort_p = p;
    qsort(rows, (size_t)N, (size_t)p * 4, cmp_rows);
    for (j = 0; j < p; ++j) {
        g.v[j] = (uint32_t *)malloc((size_t)(N ? N : 1) * 4);
        g.fc[j] = (uint32_t *)malloc((size_t)(N ? N : 1) * 4);
    }
    group_rows(rows, p, 0, N, 0, &g);
    free(rows);
    t->p = p;
    t->vb[0] = 0;
    for (j = 0; j < p; ++j) t->vb[j + 1] = t->vb[j] + g.len[j];
    for (j = p + 1; j <= SGE_MAX_K; ++j) t->vb[j] = t->vb[p];
    t->val = (uint32_t *)malloc((size_t)(t->vb[p] ? t->vb[p] : 1) * 4);
    t->child = (uint32_t *)malloc((size_t)(t->vb[p - 1] + p) * 4);
    for (j = 0; j < p; ++j) {
        uint32_t a;
        for (a = 0; a < g.len[j]; ++a) {
            t->val[t->vb[j] + a] = g.v[j][a];
            if (j + 1 < p) t->child[t->vb[j] + j + a] = t->vb[j + 1] + g.fc[j][a];
        }
        if (j + 1 < p) t->child[t->vb[j] + j + g.len[j]] = t->vb[j + 1] + g.len[j + 1];
        free(g.v[j]);
        free(g.fc[j]);
    }
    return 0;
}

static void bag_trie_free(otrie_t *t) {
    free(t->val);
    free(t->child);
    memset(t, 0, sizeof(*t));
}

/* The tries of every bag of the plan; 0 on success (nothing left allocated on failure). */
static int bags_build(const sge_plan *pl, const uint32_t *off, const uint32_t *nbr, uint32_t n, otrie_t *t) {
    int bi;
    for (bi = 0; bi < NBAGS; ++bi) {
        memset(&t[bi], 0, sizeof(t[bi]));
        if (bag_levels(pl, bi) && bag_trie_build(pl, bag_levels(pl, bi), off, nbr, n, &t[bi])) {
            while (--bi >= 0)
                if (bag_levels(pl, bi)) bag_trie_free(&t[bi]);
            return -1;
        }
    }
    return 0;
}

static void bags_attach(ctx_t *c, const sge_plan *pl, const otrie_t *t) {
    int bi;
    for (bi = 0; bi < NBAGS; ++bi)
        if (bag_levels(pl, bi)) ctx_set_trie(c, bi, &t[bi]);
}

static void bags_free(const sge_plan *pl, otrie_t *t) {
    int bi;
    for (bi = 0; bi < NBAGS; ++bi)
        if (bag_levels(pl, bi)) bag_trie_free(&t[bi]);
}

/* The trie of the plan's first bag: depth returned (-1 on failure),
 * vb[SGE_MAX_K + 1] filled; values / child copied when their capacities suffice
 * (call with 0 to size). */
int oracle_bag_trie(const sge_plan *pl, const uint32_t *off, const uint32_t *nbr, uint32_t n, uint32_t *vb,
                    uint32_t *values, uint64_t values_cap, uint32_t *child, uint64_t child_cap) {
    otrie_t t;
    uint32_t j;
    if (!bag_levels(pl, 0) || bag_trie_build(pl, bag_levels(pl, 0), off, nbr, n, &t)) return -1;
    for (j = 0; j <= SGE_MAX_K; ++j) vb[j] = t.vb[j];
    if (values && values_cap >= t.vb[t.p]) memcpy(values, t.val, (size_t)t.vb[t.p] * 4);
    if (child && child_cap >= (uint64_t)t.vb[t.p - 1] + t.p - 1)
        memcpy(child, t.child, ((size_t)t.vb[t.p - 1] + t.p - 1) * 4);
    j = t.p;
    bag_trie_free(&t);
    return (int)j;
}

/* Count over root tasks [tb, te) (tasks are CSR positions = directed edges).
 * threads <= 0: OpenMP default.  Returns the count (times plan multiplier). */
uint64_t oracle_count_range(const sge_plan *pl, const uint32_t *off, const uint32_t *nbr, uint32_t n,
                            uint64_t tb, uint64_t te, int threads, uint64_t *intersections) {
    uint64_t total = 0, inter = 0;
    int k = (int)pl->k;
    otrie_t trie[NBAGS];
    if (k == 1) return (uint64_t)n * pl->multiplier;
    if (bags_build(pl, off, nbr, n, trie)) return ~0ull;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel reduction(+ : total, inter)
    {
        ctx_t c;
        ctx_init(&c, pl, off, nbr, n);
        bags_attach(&c, pl, trie);
        uint64_t t;
#pragma omp for schedule(dynamic, 64)
        for (t = tb; t < te; ++t) {
            uint32_t v0 = src_of(off, n, t), v1 = nbr[t];
            if (((pl->lower[1] & 1u) && !(v1 > v0)) || ((pl->upper[1] & 1u) && !(v1 < v0))) continue;
            c.phi[0] = v0;
            c.phi[1] = v1;
            if (!bag_root_ok(&c)) continue;
            if (k == 2) {
                total += 1;
                continue;
            }
            total += dfs(&c, 2);
        }
        inter += c.intersections;
        ctx_free(&c);
    }
    bags_free(pl, trie);
    if (intersections) *intersections = inter;
    return total * pl->multiplier;
}

/* Count over the strided root sample {offset, offset+stride, ...}. */
uint64_t oracle_count_strided(const sge_plan *pl, const uint32_t *off, const uint32_t *nbr, uint32_t n,
                              uint64_t stride, uint64_t offset, int threads) {
    uint64_t total = 0, nnz = off[n], ntask = offset < nnz ? (nnz - offset + stride - 1) / stride : 0, i;
    int k = (int)pl->k;
    otrie_t trie[NBAGS];
    if (k < 2) return 0;
    if (bags_build(pl, off, nbr, n, trie)) return ~0ull;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel reduction(+ : total)
    {
        ctx_t c;
        ctx_init(&c, pl, off, nbr, n);
        bags_attach(&c, pl, trie);
#pragma omp for schedule(dynamic, 16)
        for (i = 0; i < ntask; ++i) {
            uint64_t t = offset + i * stride;
            uint32_t v0 = src_of(off, n, t), v1 = nbr[t];
            if (((pl->lower[1] & 1u) && !(v1 > v0)) || ((pl->upper[1] & 1u) && !(v1 < v0))) continue;
            c.phi[0] = v0;
            c.phi[1] = v1;
            if (!bag_root_ok(&c)) continue;
            total += (k == 2) ? 1 : dfs(&c, 2);
        }
        ctx_free(&c);
    }
    bags_free(pl, trie);
    return total * pl->multiplier;
}

uint64_t oracle_count(const sge_plan *pl, const uint32_t *off, const uint32_t *nbr, uint32_t n, int threads) {
    return oracle_count_range(pl, off, nbr, n, 0, off[n], threads, NULL);
}

/* Enumerate one representative per counted subgraph (plan order clamps), images
 * written per pattern vertex into emb[cap*k].  Returns the number found. */
uint64_t oracle_enumerate(const sge_plan *pl, const uint32_t *off, const uint32_t *nbr, uint32_t n, uint32_t *emb,
                          uint64_t cap) {
    uint64_t found = 0;
    sge_plan p2 = *pl;
    p2.tail = 0; /* enumerate every member */
    p2.bag = 0;  /* over CSR rows alone */
    if (p2.k < 3) return 0;
#pragma omp parallel
    {
        ctx_t c;
        ctx_init(&c, &p2, off, nbr, n);
        c.emb = emb;
        c.emb_cap = cap;
        c.emb_n = &found;
        uint64_t t;
#pragma omp for schedule(dynamic, 64)
        for (t = 0; t < off[n]; ++t) {
            uint32_t v0 = src_of(off, n, t), v1 = nbr[t];
            if (((p2.lower[1] & 1u) && !(v1 > v0)) || ((p2.upper[1] & 1u) && !(v1 < v0))) continue;
            c.phi[0] = v0;
            c.phi[1] = v1;
            dfs(&c, 2);
        }
        ctx_free(&c);
    }
    return found;
}

/* Brute force: every injective homomorphism of the k-vertex pattern (adjacency
 * bitmasks) into the graph, by backtracking in vertex order 0..k-1 with a plain
 * adjacency test.  No reference counterpart; independent check for small graphs. */
static int has_edge(const uint32_t *off, const uint32_t *nbr, uint32_t a, uint32_t b) {
    uint32_t lo = off[a], hi = off[a + 1];
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (nbr[mid] < b) lo = mid + 1; else hi = mid;
    }
    return lo < off[a + 1] && nbr[lo] == b;
}

static uint64_t brute_rec(int i, int k, const uint32_t *adj, const uint32_t *off, const uint32_t *nbr, uint32_t n,
                          uint32_t *img) {
    if (i == k) return 1;
    uint64_t s = 0;
    uint32_t v;
    for (v = 0; v < n; ++v) {
        int ok = 1, j;
        for (j = 0; j < i && ok; ++j) {
            if (img[j] == v) ok = 0;
            else if (((adj[i] >> j) & 1u) && !has_edge(off, nbr, img[j], v)) ok = 0;
        }
        if (!ok) continue;
        img[i] = v;
        s += brute_rec(i + 1, k, adj, off, nbr, n, img);
    }
    return s;
}

uint64_t oracle_count_brute(uint32_t k, const uint32_t *adj, const uint32_t *off, const uint32_t *nbr, uint32_t n) {
    uint64_t total = 0;
    uint32_t v0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (v0 = 0; v0 < n; ++v0) {
        uint32_t img[SGE_MAX_K];
        img[0] = v0;
        total += brute_rec(1, (int)k, adj, off, nbr, n, img);
    }
    return total;
}
