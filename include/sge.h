/*
 * sge.h -- C ABI of the MI355X subgraph-enumeration executor (libsge.so).
 *
 * This is the drop-in boundary for the reference's plan-to-join path:
 *   main.cpp:64-69      optCostPlan / selectSymmetry / buildTraverseUnlabeled /
 *                       refineIntersectionInfo  -> sge_plan_build
 *   matching/join.cpp:812       sharedJoin (single thread)      -> sge_count
 *   matching/parallel.cpp:1154  parSharedJoin (WCOJ, OpenMP)    -> sge_count
 *   matching/parallel.cpp:1488  parSharedJoin (exploration)     -> sge_count (SGE_MODE_EXPLORE)
 *   relation/graph.cpp:255      DataGraph::loadDataGraph        -> sge_graph_load
 *   relation/graph.cpp:661      PatternGraph::loadPatternGraph  -> sge_pattern_load
 *
 * Conventions: every function returns 0 on success and a negative SGE_E* code
 * on failure; sge_last_error() gives the message (thread-local).  No C++
 * exception crosses the boundary: a host allocation failure inside a call is
 * SGE_ENOMEM.  Only POD types cross the boundary.  The caller owns every host buffer it passes; the library
 * owns graph/batch/model handles and their device memory (freed by *_destroy).
 * Counts are assigned, never accumulated, and are exact uint64 (the IEP
 * closed form reproduces the reference's wrapping choosec, utility/utils.cpp:80-94).
 * Handles are thread-compatible, not thread-safe.
 */
#ifndef SGE_H
#define SGE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGE_MAX_K 8 /* pattern vertices; the reference's benchmark sets use <= 7 */
#define SGE_MAX_BAGS 2 /* materialised bags per plan (sge_plan.bag) */

/* error codes */
#define SGE_OK 0
#define SGE_EINVAL (-1)   /* bad argument / unsupported shape */
#define SGE_EIO (-2)      /* file open / parse failure (reference: exit(1), graph.cpp:260,296,665,683) */
#define SGE_EDEVICE (-3)  /* HIP runtime failure or no device */
#define SGE_ENOMEM (-4)   /* allocation failure */
#define SGE_ELIMIT (-6)   /* the tuning's node cap stopped the call; its counts are partial */
#define SGE_ERANGE (-5)   /* capacity exceeded (output buffer, sizes) */

/* executor modes (reference flags: default WCOJ, -i exploration, -u IEP) */
#define SGE_MODE_WCOJ 0u
#define SGE_MODE_EXPLORE 1u
#define SGE_MODE_IEP 2u

/* sge_count flags */
#define SGE_COUNT_BATCHED 1u  /* all plans in one persistent launch with a shared root queue */
#define SGE_COUNT_GRAPH 2u    /* replay through a captured hipGraph (with BATCHED) */

/* graph construction flags */
#define SGE_GRAPH_DEGREE_ORDER 1u /* relabel vertices by ascending degree (ties by id) */
#define SGE_GRAPH_BUILD_DEVICE 2u /* deduplicate, order and build the CSR on the GPU (radix sorts);
                                     identical arrays, and the graph is device-resident on return */

/* ------------------------------------------------------------------ errors */
const char *sge_last_error(void);
int sge_version(void);

/* ------------------------------------------------------------------ graphs */
typedef struct sge_graph sge_graph;

typedef struct sge_graph_stats {
    uint32_t n;          /* vertices */
    uint64_t m;          /* undirected edges after dedup / self-loop removal */
    uint32_t max_degree;
    uint32_t degree_ordered;
    uint64_t csr_bytes;  /* offsets + neighbour bytes */
    uint64_t triangles_hint; /* 0 unless computed */
    double build_ms;     /* wall time of the CSR build (normalise, order, CSR), host or device */
} sge_graph_stats;

/* Text "n m" + m lines "u v", or the reference's binary "BINGRPH" + u32 n + u32 m
 * + m x (u32,u32) (relation/graph.cpp:255-314, 357-419).  Duplicate edges and
 * self-loops are dropped. */
int sge_graph_load(const char *path, uint32_t flags, sge_graph **out);
/* Undirected edge list, uv[2*i], uv[2*i+1]. */
int sge_graph_from_edges(uint32_t n, const uint32_t *uv, uint64_t m, uint32_t flags, sge_graph **out);
/* Symmetric CSR with ascending neighbour lists (the reference's Graph layout,
 * relation/graph.h:29-102); nnz = 2m. */
int sge_graph_from_csr(const uint32_t *offsets, const uint32_t *nbrs, uint32_t n, uint64_t nnz, uint32_t flags,
                       sge_graph **out);
/* Synthetic generators (seeded, deterministic, independent of thread count). */
int sge_graph_chung_lu(uint32_t n, uint64_t m, double gamma, double degree_cap, uint64_t seed, uint32_t flags,
                       sge_graph **out);
int sge_graph_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c, uint64_t seed,
                   uint32_t flags, sge_graph **out);
int sge_graph_stats_get(const sge_graph *g, sge_graph_stats *out);
/* Copy the CSR (in the graph's own labelling) into caller buffers of n+1 / nnz entries. */
int sge_graph_csr(const sge_graph *g, uint32_t *offsets, uint32_t *nbrs);
/* perm[new_id] = original id (identity if not degree-ordered). */
int sge_graph_perm(const sge_graph *g, uint32_t *perm);
/* Write "n m" + "u v" (u<v, original ids) text, or BINGRPH when binary != 0. */
int sge_graph_write(const sge_graph *g, const char *path, int binary);
/* Upload the CSR to the current HIP device (idempotent). */
int sge_graph_to_device(sge_graph *g);
void sge_graph_destroy(sge_graph *g);

/* ---------------------------------------------------------------- patterns */
typedef struct sge_pattern {
    uint32_t k;                 /* vertices, 1..SGE_MAX_K */
    uint32_t adj[SGE_MAX_K];    /* adjacency bitmasks */
} sge_pattern;

typedef struct sge_pattern_info {
    uint32_t k, m;
    uint64_t aut_size;          /* |Aut(P)| */
    uint64_t canon;             /* canonical code (min adjacency word over relabellings) */
    uint32_t orbit[SGE_MAX_K];  /* orbit representative (smallest vertex) per vertex */
    uint32_t core_size;         /* 2-core size (relation/graph.cpp:1231) */
    uint32_t num_bags;          /* bags of the chosen minimal FHD incl. peripheral edge bags */
    uint32_t single_bag;        /* 1: the reference prints embeddings, not subgraphs (decomposition.cpp:983) */
    double fhw;                 /* fractional hypertree width of the core */
    uint32_t connected;
} sge_pattern_info;

/* Pattern text file (relation/graph.cpp:661-772): "k m", m edges, optional orbit line (ignored). */
int sge_pattern_load(const char *path, sge_pattern *out);
int sge_pattern_from_edges(uint32_t k, const uint32_t *uv, uint32_t m, sge_pattern *out);
int sge_pattern_analyze(const sge_pattern *p, sge_pattern_info *out);
/* Connected k-vertex patterns with emin..emax edges: uniform edge count, uniform
 * edge subset, reject disconnected and isomorphic duplicates (and codes listed in
 * exclude).  Returns the number written (<= want) in *got. */
int sge_pattern_sample(uint32_t k, uint32_t emin, uint32_t emax, uint32_t want, uint64_t seed,
                       const uint64_t *exclude, uint32_t n_exclude, sge_pattern *out, uint32_t *got);

/* Replace each embedding row (k images per pattern vertex) by its
 * lexicographically smallest relabelling over Aut(P), then sort the rows:
 * a plan-independent canonical embedding list. */
int sge_canonicalize_embeddings(const sge_pattern *p, uint32_t *rows, uint64_t n);

/* ------------------------------------------------------------------- plans */
/* Flat plan descriptor: one matching order over the whole pattern, symmetry
 * breaking as per-level clamps, last level(s) counted in closed form.  Level i
 * matches pattern vertex order[i]; every mask is over LEVELS (bit j = level j). */
typedef struct sge_plan {
    uint32_t k;
    uint32_t mode;                     /* SGE_MODE_* */
    uint32_t order[SGE_MAX_K];         /* level -> pattern vertex */
    uint32_t parent[SGE_MAX_K];        /* earlier levels adjacent in the pattern (backward neighbours) */
    uint32_t lower[SGE_MAX_K];         /* image must be > image of these levels (setBeginPos, utils.cpp:581) */
    uint32_t upper[SGE_MAX_K];         /* image must be < image of these levels (setEndPos, utils.cpp:599) */
    uint32_t check[SGE_MAX_K];         /* image must differ from these levels' images (injectivity) */
    int32_t reuse[SGE_MAX_K];          /* earlier level whose candidate set contains this one's (candidatesBefore), -1 */
    uint32_t ready[SGE_MAX_K];         /* exploration: level after which this level's parent intersection is formed */
    uint32_t tail;                     /* IEP: trailing levels counted in closed form (0 = none) */
    uint32_t tail_kind;                /* 1: symmetric twin group, choosec (join.cpp:1532-1537);
                                          2: independent set, inclusion-exclusion over set partitions */
    uint32_t bag;                      /* levels of up to two materialised bags (0 = none): bag 0's
                                          level mask in bits 0-7, bag 1's in bits 8-15 (the bags of a
                                          hypertree decomposition, HyperTree decomposition.h:59-123,
                                          may share levels).  Each bag's relation is enumerated and
                                          built into a trie before the join, and a bag level then
                                          reads the trie child list of every bag holding it in place
                                          of the rows of those bags' parents (nodeJoin
                                          join.cpp:309-421, buildTrie join.cpp:1713-1717, globalJoin
                                          join.cpp:1202-1420); set with sge_plan_set_bag and
                                          sge_plan_add_bag */
    uint32_t num_bags;                 /* decomposition bag count (reporting / features) */
    uint64_t multiplier;               /* |Aut| when the reference counts embeddings (single bag), else 1 */
    uint64_t aut_size;
    double cost;                       /* cost-model estimate of the order */
} sge_plan;

/* Build a plan.  order == NULL lets the cost model choose (plan/optimizer.cpp:441
 * optCostOrder restated); g may be NULL (uniform-degree cost model). */
int sge_plan_build(const sge_pattern *p, const sge_graph *g, uint32_t mode, const uint32_t *order, sge_plan *out);
/* Cost-model estimate of one order (lower is better). */
int sge_plan_cost(const sge_pattern *p, const sge_graph *g, uint32_t mode, const uint32_t *order, double *cost);
/* The reference's sampled cost model for one order (computeOrderCost,
 * plan/optimizer.cpp:319-439, with cardinalities from cardEstimateUnlabeled,
 * plan/estimator.cpp:562-661: SAMPLE_SIZE 40000, SAMPLE_PORTION 4), for a
 * plan whose one bag holds every pattern vertex, under the plan's symmetry
 * bounds; deterministic for a seed.  The fallback and judge of the learned
 * order model.  g is required. */
int sge_plan_cost_sampled(const sge_pattern *p, const sge_graph *g, uint32_t mode, const uint32_t *order,
                          uint64_t seed, double *cost);
/* All connected orders of the pattern (k! bounded), deduplicated up to automorphism; returns count in *n. */
int sge_plan_orders(const sge_pattern *p, uint32_t *orders, uint32_t cap, uint32_t *n);
/* All connected orders ranked as the planner ranks them for `mode` (longest
 * closed-form tail first in IEP mode, then lowest cost): orders[i*k..], costs[i],
 * tails[i] for i < min(cap, *n); any output may be NULL. */
int sge_plan_rank_orders(const sge_pattern *p, const sge_graph *g, uint32_t mode, uint32_t *orders, double *costs,
                         uint32_t *tails, uint32_t cap, uint32_t *n);
/* Reference-style plan text ("num bags", symmetry rules, ...) into buf. */
int sge_plan_text(const sge_plan *plan, char *buf, size_t cap);
/* Materialise the pattern vertices in vertex_mask as one bag of `plan` (0 clears
 * it).  The bag relation is every assignment of the bag's levels that satisfies
 * the plan's edges, bounds and injectivity checks among them (the reference's
 * bag tuples, join.cpp:309-421); the join then intersects trie child lists with
 * the rows of parents outside the bag (twoIntersection / multiIntersection,
 * join.cpp:9-135).  In exploration mode the non-bag levels form their parent
 * intersections early as usual and a bag level forms its set from the trie list
 * when it is reached (the exploration nodeJoin / globalJoin, join.cpp:423-810).
 * Counts never change.  Valid bags: >= 3 vertices, every bag level below the
 * closed-form level(s), the first two bag levels adjacent and every later bag
 * level adjacent to an earlier bag level (SGE_EINVAL otherwise). */
int sge_plan_set_bag(sge_plan *plan, uint32_t vertex_mask);
/* Add the pattern vertices in vertex_mask as a second bag of a plan that has
 * exactly one (the reference's plans join two to five bags, HyperTree
 * decomposition.h:59-123, globalJoin join.cpp:1202-1420): the same validity
 * rules per bag, different from the first; the bags may share vertices, and a
 * level in both intersects both trie lists. */
int sge_plan_add_bag(sge_plan *plan, uint32_t vertex_mask);
/* The bags of the pattern's minimal fractional hypertree decomposition (vertex
 * masks, as all_td.cpp:483-691 chooses them) that are valid bags of `plan`;
 * count in *n (at most cap written). */
int sge_plan_bags(const sge_pattern *p, const sge_plan *plan, uint32_t *masks, uint32_t cap, uint32_t *n);

/* ---------------------------------------------------------------- executor */
typedef struct sge_tuning {
    uint32_t block_threads;   /* 64..1024, multiple of 64 (one DFS per wavefront) */
    uint32_t chunk;           /* root subtasks per queue item after the head items (0 = default 4) */
    uint32_t stage_max;       /* probe-direction ratio: a sibling's row drives the count while
                                 |row| <= stage_max * |F| and F is staged in LDS as a sorted copy
                                 (0 = default 2) */
    uint32_t blocks_per_cu;   /* 0 = occupancy-derived */
    uint32_t split_span;      /* a root edge whose level-2 set may exceed this is split into
                                 ceil(bound / split_span) subtasks (0 = default 32, 1 << 30 = off) */
    uint32_t donate_budget;   /* wave steps (one per search-tree node, one per 64 candidates it
                                 forms, one per 64 elements of the shortest list per other list of
                                 every intersection, plus the 64-lane iterations of each last-level
                                 count) one work item may spend before its open candidate ranges
                                 are handed out as continuation items (0 = default 1024, scaled to
                                 1024 * 65536 / max degree, at least 256, on graphs with rows of
                                 more than 65,536 ids; 1 << 30 = off) */
    uint32_t bitmap_ratio;    /* the same ratio when F is staged as an LDS bitmap (0 = stage_max) */
    uint32_t node_cap_k;      /* stop a counting call (plain or batched) after about node_cap_k * 1024
                                 wave steps (as donate_budget counts them) and fail with SGE_ELIMIT
                                 (0 = no cap).  Each wavefront checks the cap between search-tree
                                 nodes and between sibling batches of a last-level count, so a capped
                                 call always ends with its kernels exiting normally.  For timing
                                 orders of unknown cost and for bounding runs on large graphs. */
    uint32_t head_items;      /* the heaviest head_items root subtasks of each plan are queue
                                 items of their own, the rest come in chunks (0 = one per
                                 resident wavefront, 1 << 30 = none) */
    uint32_t global_bitmap_min; /* a last-level F set too large for the LDS stage, with at least
                                 this many elements, is staged as a bitmap in a per-wave global
                                 region when its siblings' rows will probe it (0 = default 2048,
                                 1 << 30 = off) */
    uint32_t donate_split;    /* a donated candidate range goes to the next pass as at most this many
                                 items of near-equal length (0 = default 64, at most 256) */
    uint32_t pass_items;      /* 2: waves whose queue is empty take the continuation items their own
                                 pass is writing; 1: those items wait for the next pass; 0 (default):
                                 2 on graphs with a row of 16,384 ids or more, else 1 */
} sge_tuning;

typedef struct sge_stats {
    uint64_t intersections;   /* pairwise sorted-list intersections */
    uint64_t bytes;           /* algorithmic bytes: 4(|A'|+|B'|) + 4|out| + 8 per CSR row */
    uint64_t tasks;           /* root tasks processed */
    uint64_t nodes;           /* materialised search-tree nodes */
    double kernel_ms;         /* device time of the counting launch(es) */
    double wall_ms;           /* host wall time of the call */
    uint64_t wave_cycles;     /* wavefront clock cycles spent on the plan's work (total work,
                                 independent of how well the launch was balanced) */
    double trie_ms;           /* host wall time spent materialising bag tries (sge_batch_run:
                                 the one-off build at sge_batch_create) */
    uint64_t model_bytes;     /* list bytes of the per-pair model, 4(|A'|+|B'|) + 4|out| + 8 per CSR
                                 row for every pairwise intersection, whichever list drove it
                                 (`bytes` charges the last level's probes as executed: 8 per probe) */
} sge_stats;

void sge_tuning_default(sge_tuning *t);

/* Count embeddings of every plan over g.  counts[i] is assigned (not accumulated).
 * flags: 0 = one launch per plan (plain executor); SGE_COUNT_BATCHED = one
 * persistent launch over all plans with a shared root queue.
 * stats may be NULL; per-plan stats when per_plan_stats != NULL (nplans entries). */
int sge_count(sge_graph *g, const sge_plan *plans, uint32_t nplans, uint32_t flags, const sge_tuning *tune,
              uint64_t *counts, sge_stats *stats, sge_stats *per_plan_stats);
/* Count over the root-task subrange [task_begin, task_end) of one plan (counts are
 * additive over disjoint ranges).  Used for sampled work logging.  Root tasks are
 * the directed data edges in CSR order, the reference's edge list
 * (CandidateSpace::getEdges, relation/candidate_space.cpp:787-799) that
 * parSharedJoin splits over its threads (matching/parallel.cpp:1414-1458). */
int sge_count_range(sge_graph *g, const sge_plan *plan, uint64_t task_begin, uint64_t task_end,
                    const sge_tuning *tune, uint64_t *count, sge_stats *stats);
/* Count plans[i] over its own root-task range [task_begin[i], task_end[i]):
 * one launch per plan, or all plans in one launch with SGE_COUNT_BATCHED.
 * Per-plan ranges let a pattern set run on bounded root subsets of a graph
 * whose whole job is out of reach (partial counts depend on the plan's order:
 * compare executors on the same plans). */
int sge_count_ranges(sge_graph *g, const sge_plan *plans, uint32_t nplans, const uint64_t *task_begin,
                     const uint64_t *task_end, uint32_t flags, const sge_tuning *tune, uint64_t *counts,
                     sge_stats *stats, sge_stats *per_plan);
/* Count over the strided root-task sample {offset, offset+stride, ...} of each
 * plan, one launch per plan (the plain executor); per-plan stats.  Used to log
 * (order, measured work) pairs for the learned cost model. */
int sge_count_sampled(sge_graph *g, const sge_plan *plans, uint32_t nplans, uint64_t stride, uint64_t offset,
                      const sge_tuning *tune, uint64_t *counts, sge_stats *per_plan);
/* Number of root tasks (directed edges) of a plan over g. */
int sge_num_tasks(const sge_graph *g, const sge_plan *plan, uint64_t *n);

/* Write every counted embedding (one representative per subgraph, images in
 * pattern-vertex order, original vertex ids) into out[cap*k]; *n_out gets the
 * total (may exceed cap -> SGE_ERANGE, nothing beyond cap written). */
int sge_enumerate(sge_graph *g, const sge_plan *plan, uint32_t *out, uint64_t cap, uint64_t *n_out);

/* Relation of the plan's first bag (level mask plan->bag & 0xff) as a
 * level-order trie on the device: the relation
 * is enumerated by the executor (the bag's levels in plan order), radix sorted
 * lexicographically, and grouped per depth (TrieLevel::buildTrieFromSortedMatchesBatch,
 * relation.cpp:340-379; TrieLevel relation.h:49-69).  Depth j holds nodes[j]
 * values (sorted within each parent's child list), stored after those of the
 * shallower depths; the children of the node at absolute value index a of depth
 * j < depth-1 are the values in [child[a + j], child[a + j + 1]) (each depth's
 * offsets end with a closing sentinel).  Values are the graph's internal
 * (degree-ordered) vertex ids. */
typedef struct sge_trie sge_trie;
int sge_trie_build(sge_graph *g, const sge_plan *plan, sge_trie **out);
/* depth = bag levels, tuples = relation size, nodes[SGE_MAX_K] per depth, build_ms. */
int sge_trie_shape(const sge_trie *t, uint32_t *depth, uint64_t *tuples, uint64_t *nodes, double *build_ms);
/* values[sum nodes], child[sum over j < depth-1 of nodes[j] + 1] */
int sge_trie_copy(const sge_trie *t, uint32_t *values, uint32_t *child);
void sge_trie_destroy(sge_trie *t);

/* Batched executor captured once in a hipGraph, replayed per run.  A batch
 * borrows its graph's device arrays: sge_graph_destroy on a graph with live
 * batches only marks it, and the last sge_batch_destroy releases it. */
typedef struct sge_batch sge_batch;
int sge_batch_create(sge_graph *g, const sge_plan *plans, uint32_t nplans, const sge_tuning *tune, sge_batch **out);
/* Same over the strided root-edge sample {offset, offset+stride, ...}. */
int sge_batch_create_sampled(sge_graph *g, const sge_plan *plans, uint32_t nplans, uint64_t stride, uint64_t offset,
                             const sge_tuning *tune, sge_batch **out);
/* Same with plan i over its own root-task range [task_begin[i], task_end[i]) (as sge_count_ranges). */
int sge_batch_create_ranges(sge_graph *g, const sge_plan *plans, uint32_t nplans, const uint64_t *task_begin,
                            const uint64_t *task_end, const sge_tuning *tune, sge_batch **out);
/* SGE_ELIMIT when the tuning's node_cap_k stopped the run (counts partial). */
int sge_batch_run(sge_batch *b, uint64_t *counts, sge_stats *stats);
void sge_batch_destroy(sge_batch *b);

/* Allocation-free count.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The same counts as sge_count with
 * SGE_COUNT_BATCHED, with every device buffer the kernels need beyond the
 * inputs and outputs taken from one static per-device array of 64 MiB that the
 * library holds until the process ends (control words, per-plan outputs and
 * descriptors, continuation items, and as many wavefronts' scratch slabs as fit
 * -- fewer at high max degree, so large graphs run on part of the GPU).  It
 * calls no hipMalloc/hipFree.  Requirements: the graph already resident on the
 * current device (sge_graph_to_device or SGE_GRAPH_BUILD_DEVICE), at most 256
 * plans, no bag; large last-level sets are searched instead of staged as global
 * bitmaps and hub root edges are not split.  SGE_ERANGE when one wavefront's
 * slab exceeds the array (max degree above about 1.2 million for a 7-vertex IEP
 * set) or nplans is above 256.  Every call on a GPU reuses that GPU's array and
 * leaves the last call's contents in it: calls that overlap on one GPU race.
 * stats (may be NULL): kernel_ms is the call's wall time. */
int sge_count_noalloc(sge_graph *g, const sge_plan *plans, uint32_t nplans, const sge_tuning *tune,
                      uint64_t *counts, sge_stats *stats);

/* ------------------------------------------------------ learned cost model
 * 64 -> 128 -> 128 -> 1 ReLU MLP; hidden layers on BF16 MFMA with FP32
 * accumulation, FP32 master weights trained by Adam.  No reference counterpart:
 * it replaces the order choice of plan/optimizer.cpp:441-577 (optCostOrder),
 * which remains the fallback.  Flat FP32 parameter layout:
 * W1[64][128] | b1[128] | W2[128][128] | b2[128] | w3[128] | b3. */
typedef struct sge_mlp sge_mlp;
int sge_mlp_num_params(void);
int sge_mlp_input_dim(void);
int sge_mlp_create(uint64_t seed, sge_mlp **out);
int sge_mlp_set_params(sge_mlp *m, const float *params, size_t n);
int sge_mlp_get_params(sge_mlp *m, float *params_out, size_t n);
/* y[N] = f(X[N][64]) */
int sge_mlp_forward(sge_mlp *m, const float *X, uint32_t N, float *y);
/* MSE loss gradients w.r.t. the flat parameters (no update). */
int sge_mlp_gradients(sge_mlp *m, const float *X, const float *target, uint32_t N, float *grads, float *loss);
/* forward + backward + Adam step */
int sge_mlp_train_step(sge_mlp *m, const float *X, const float *target, uint32_t N, float lr, float *loss);
void sge_mlp_destroy(sge_mlp *m);
/* The model's 64 input features for one (pattern, order): pattern shape, plan
 * structure per level, graph moments, analytic cost, and a Knuth random-walk
 * estimate (nsamples random root edges, seeded) of nodes and list work per level. */
int sge_order_features(const sge_pattern *p, const sge_graph *g, const uint32_t *order, uint32_t mode,
                       uint32_t nsamples, uint64_t seed, float *feat64);
/* The same features with the random walks rooted in the CSR positions
 * [begin, end) only: the task set of a run over a root-edge range
 * (sge_count_range), whose work can rank orders differently from the whole
 * job's.  The reference samples its cardinalities over the whole candidate
 * space (plan/estimator.cpp:562-661); a range has no counterpart there. */
int sge_order_features_range(const sge_pattern *p, const sge_graph *g, const uint32_t *order, uint32_t mode,
                             uint32_t nsamples, uint64_t seed, uint64_t begin, uint64_t end, float *feat64);

/* Streaming copy of `bytes` (read once, written once) timed on the device: the
 * known-size kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings. */
int sge_calibration_copy(uint64_t bytes, double *ms);

/* Device info for reporting. */
int sge_device_info(int *num_cu, int *num_xcd_hint, char *name, size_t cap);

#ifdef __cplusplus
}
#endif
#endif /* SGE_H */
