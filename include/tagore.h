/* tagore.h -- C ABI of the MI355X (gfx950) Tagore library, libtagore.so.
 *
 * The reference exposes its path only through a pybind11 module (pybind.cu:393-399):
 * GNN_descent, Pruning, Pruning_with_CFS and largeIndex, which allocate device buffers,
 * launch the CUDA kernels and hand raw device pointers back to Python.  This header is the
 * boundary underneath an equivalent module: plain pointers and sizes, an opaque handle that
 * owns every device buffer, integer status codes (0 = success) instead of exit(), and a
 * thread-local message for the last failure.  The pybind11 module tagore_amd/Tagore*.so is
 * a thin layer over these calls; ctypes/cffi callers bind them directly.
 *
 * Threading: a handle is used by one host thread at a time; calls are synchronous (they
 * return after the device work completes) and run on the handle's own HIP stream.
 */
#ifndef TAGORE_H
#define TAGORE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tg_state tg_state;

/* status codes */
#define TG_OK 0
#define TG_ERR_ARG 1      /* invalid argument or combination */
#define TG_ERR_IO 2       /* file missing, short or unwritable */
#define TG_ERR_HIP 3      /* HIP runtime failure */
#define TG_ERR_STATE 4    /* call out of order (e.g. prune before a kNN graph exists) */
#define TG_ERR_CAPACITY 5 /* an internal bounded buffer overflowed */

/* kNN distance precision (the screen; graph decisions are always canonical FP32) */
#define TG_PREC_EXACT 0
#define TG_PREC_FP32 1
#define TG_PREC_BF16 2
#define TG_PREC_FP8 3   /* e4m3 with a per-vector scale (D % 16 == 0) */

/* pruning index kinds (Pruning: NSG/Vamana -> PATH, NSSG -> TWOHOP, CAGRA, DPG) */
#define TG_KIND_PATH 0
#define TG_KIND_TWOHOP 1
#define TG_KIND_CAGRA 2
#define TG_KIND_DPG 3
/* rules */
#define TG_RULE_DIST 0
#define TG_RULE_ANGLE 1
#define TG_RULE_RANK 2

/* initial lists of GNN-Descent */
#define TG_INIT_RANDOM 0  /* seeded random ids (initialize_graph, Tagore_src.cu:158-166) */
#define TG_INIT_LEARNED 1 /* MLP codes -> code-space buckets -> seeded lists (D % 32 == 0) */
#define TG_INIT_GIVEN 2   /* the ids of the current graph (tg_set_knn / a previous build) */

typedef struct {
  uint32_t K;         /* neighbours per node */
  uint32_t iters;     /* iterations; iters/2 local-join + iters/2 expansion (pybind.cu:101);
                         0 returns the initial lists */
  uint64_t seed;      /* random streams of the initial lists */
  int32_t precision;  /* TG_PREC_* */
  int32_t init;       /* TG_INIT_* */
  uint32_t sample;    /* learned init: sample size (0: min(N, 50000)) */
  uint32_t epochs;    /* learned init: training epochs (0: 2) */
  uint32_t tile;      /* local-join MFMA staging depth: 0 = 256-byte rows per LDS round,
                         1 = 64-byte rows (results are identical; see tg_autotune_knn) */
  uint32_t flags;     /* TG_FLAG_*: 0 for the defaults */
} tg_knn_params;
/* learned init: capture the encoder's training loop (about 2,200 small kernels) as a HIP
 * graph and replay it in later builds.  Same result as the default individual launches;
 * measured slower on MI355X with ROCm 7.2 (26.7 ms vs 21.7 ms per 1M-vector build), so
 * off by default. */
#define TG_FLAG_GRAPH 1u
/* Capture the iteration sequence after the initial lists (local-join iterations, reverse
 * lists, expansion iterations, final sort: every kernel of them, all clusters of a segmented
 * handle included) as one HIP graph, and replay it in later builds whose launch arguments
 * are the same (handle buffers, K, iterations, precision, tile, segments).  Same graph as
 * individual launches; the per-kernel section times of tg_stats stay 0 for those builds
 * (the window time is measured). */
#define TG_FLAG_GRAPH_ITERS 2u

typedef struct {
  int32_t kind;      /* TG_KIND_* */
  int32_t fwd_rule;  /* rule of the Filter after Collect (path: DIST, two-hop: ANGLE) */
  int32_t rev_rule;  /* rule of the reverse phase (DIST/ANGLE) or TG_RULE_RANK */
  uint32_t FD;       /* FINAL_DEGREE */
  uint32_t TOPM;     /* beam width m of the path Collect */
  float thre;        /* threshold: alpha for dist, cos(angle) for angle */
  int32_t precision; /* TG_PREC_BF16: read rows from the BF16 copy when every value of the
                        data set is exactly representable in BF16 (bit-identical result);
                        otherwise FP32 rows.  DPG with K <= 64 under TG_PREC_BF16 / FP8:
                        pair distances screened on an MFMA Gram, exact re-score near the
                        threshold (bit-identical result) */
} tg_prune_params;

/* Last error message of the calling thread (never NULL). */
const char* tg_last_error(void);

/* Handle with room for N vectors of dimension D (D % 4 == 0) on HIP device `device`.
 * Replaces the cudaMalloc block of GNN_descent (pybind.cu:64-93). */
int tg_create(uint32_t N, uint32_t D, int device, tg_state** out);
int tg_free(tg_state* s);

/* Vectors: upload host FP32 rows [N*D], or generate a seeded synthetic set on the device
 * (kind 0: integer-valued clipped Gaussian mixture in [0,255] with sigma; 1: unit-norm
 * Gaussian mixture; 2: |Gaussian mixture| scaled to [0,1]). */
int tg_set_data(tg_state* s, const float* host_rows);
int tg_gen_data(tg_state* s, int32_t kind, uint32_t ncomp, uint64_t seed, float sigma);
int tg_get_data(tg_state* s, float* host_rows, uint32_t first, uint32_t count);

/* .fvecs reader (common.cu:6-22): *rows is malloc'ed, free with tg_host_free. */
int tg_read_fvecs(const char* path, float** rows, uint32_t* N, uint32_t* D);
/* .bvecs (u32 dim + dim bytes per vector) and .fbin (u32 N, u32 D, N*D f32) readers, the
 * other vector formats of the ANN benchmarks the reference's loader family reads
 * (common.cu:6-22); rows as f32, malloc'ed like tg_read_fvecs. */
int tg_read_bvecs(const char* path, float** rows, uint32_t* N, uint32_t* D);
int tg_read_fbin(const char* path, float** rows, uint32_t* N, uint32_t* D);
/* Reader of the index layouts StoreNSG / StoreVamana / StoreNSSG / StoreDPG / StoreCAGRA
 * write (common.cu:151-248; tg_store_index writes the same bytes).  N: the vectors the
 * index covers.  *width: row capacity (NSG / NSSG header, Vamana max degree, DPG largest
 * degree, CAGRA d; at most 128); *deg [N] and *ids [N * width] (INV-padded) are malloc'ed
 * (tg_host_free); eps[10] receives the entry points the file names, *n_eps their number. */
int tg_read_index(const char* path, const char* index_type, uint32_t N, uint32_t* width,
                  uint32_t** deg, uint32_t** ids, uint32_t* eps, uint32_t* n_eps);
/* Installs a pruned graph (e.g. one read with tg_read_index) as the handle's current one,
 * for tg_search / tg_store_index. */
int tg_set_pruned(tg_state* s, uint32_t width, const uint32_t* deg, const uint32_t* ids,
                  uint32_t ep);
void tg_host_free(void* p);

/* kNN graph.  tg_knn_build runs GNN-Descent (GNN_descent, pybind.cu:19-142); tg_set_knn
 * installs a given graph (rows sorted by (distance, id)); tg_get_knn copies it back.
 * tg_knn_build takes handles of at most 2^26 vectors with N*K below 2^32 (TG_ERR_ARG
 * otherwise); largeIndex batches larger sets. */
int tg_knn_build(tg_state* s, const tg_knn_params* p);
int tg_set_knn(tg_state* s, uint32_t K, const uint32_t* ids, const float* dists);
int tg_get_knn(tg_state* s, uint32_t* ids, float* dists);

/* CFS pruning of the current kNN graph (Pruning / Pruning_with_CFS, pybind.cu:145-295).
 * The kNN graph is left intact, so several prunings may follow one build. */
int tg_prune(tg_state* s, const tg_prune_params* p);
/* Pruned adjacency: deg[N], ids[N*FD] (row stride FD), entry point (path kinds). */
int tg_get_pruned(tg_state* s, uint32_t* deg, uint32_t* ids, uint32_t* ep);
/* Row width (FINAL_DEGREE) of the last pruned graph; a test accessor beside
 * tg_get_pruned, with no counterpart in the reference (which returns nothing). */
int tg_pruned_width(tg_state* s, uint32_t* FD);
/* Forward-phase adjacency (after Collect+Filter, before the reverse phase). */
int tg_get_forward(tg_state* s, uint32_t* deg, uint32_t* ids, float* dists);
/* Index file of the last prune: "NSG", "Vamana", "NSSG", "DPG", "CAGRA"
 * (StoreNSG..StoreCAGRA, common.cu:151-248). */
int tg_store_index(tg_state* s, const char* path, const char* index_type, uint64_t nssg_seed);

/* Segments: the N rows are nseg independent data sets (clusters) stored back to back,
 * segment c = rows [starts[c], starts[c+1]), starts[0] = 0, starts[nseg] = N.  Builds and
 * prunes then run every segment in the same launches (gpu_construct's per-cluster loop,
 * large_index.cu:173-277); each segment's result equals that of a handle holding only its
 * rows, built with seed ^ ((key_c + 1) * 0xA0761D6478BD642F), with ids shifted by
 * starts[c]; key_c = keys[c] (keys == NULL: c).  Every segment must hold more than K rows.
 * nseg = 0 restores one data set.  Clears the current graph.  Not available with learned
 * init, tg_recall or tg_store_index. */
int tg_set_segments(tg_state* s, uint32_t nseg, const uint32_t* starts, const uint32_t* keys);
/* Autotuner of the kNN kernels: times a 2-iteration build (one local join, one expansion)
 * of the current data for each (precision, tile) candidate -- exact, FP32-MFMA tiles 0/1,
 * BF16-MFMA tiles 0/1 (D % 8 == 0), FP8-MFMA tiles 0/1 (D % 16 == 0) -- and returns the
 * fastest.  Every candidate
 * gives the same graph, so the choice only affects time.  Leaves a 2-iteration graph. */
typedef struct {
  int32_t precision;          /* fastest candidate */
  uint32_t tile;
  uint32_t ncand;
  int32_t cand_precision[8];
  uint32_t cand_tile[8];
  double cand_ms[8];          /* device window of each candidate's build */
} tg_tune_result;
int tg_autotune_knn(tg_state* s, uint32_t K, uint64_t seed, tg_tune_result* out);

/* Entry points of the last path prune, one per segment (one entry without segments). */
int tg_get_entry_points(tg_state* s, uint32_t* eps);

/* largeIndex on one GPU (pybind.cu:297-391, large_index.cu:9-522): clusters from the
 * two-centre file (u32 first, u32 second per point; equal centres are rejected), oversized
 * clusters split round-robin, the Algorithm-4 dispatch order, per-cluster GNN-Descent +
 * Vamana pruning (path/dist, alpha = thre) batched over segmented handles, the streaming
 * union merge through buffer_size cluster graphs with spill files
 * (local_index_path_prefix + cluster id: per point u32 degree + ids) and the merged index
 * (u64 size, u32 max_degree, u32 entry point, u64 0, then per point u32 degree + ids).
 * Where the reference reads undefined values the restatement takes the smallest unvisited
 * cluster / slot 0 (reorder) and clamps buffer_size to the cluster count. */
typedef struct {
  const int32_t* devicelist;
  uint32_t ndevices;               /* must be 1 */
  uint32_t DIM, K, cluster_num, max_points;
  const char* file_prefix;         /* cluster c: file_prefix + decimal c: u32 count, rows */
  const char* local_index_path_prefix;
  const char* index_store_path;
  uint32_t max_degree, buffer_size, all_node_num;
  const char* centers_path;
  float thre;
  uint32_t FINAL_DEGREE, TOPM;
  uint32_t iters;                  /* 0: 16 = 8 local-join + 8 expansion (large_index.cu:225) */
  uint64_t seed;                   /* cluster c is a segment keyed c */
  int32_t precision;               /* TG_PREC_* of the kNN screen */
  uint64_t batch_rows;             /* rows per batched build (0: from free device memory) */
  int32_t verbose;
  int32_t tune;                    /* 1: tg_autotune_knn on the first batch picks precision
                                      and tile for every batch (same graphs, less time) */
  uint32_t tile;                   /* local-join tile of every build when tune == 0 */
} tg_large_params;
typedef struct {
  uint32_t clusters, batches;      /* after splitting; batched builds */
  double build_ms;                 /* reads and builds of every batch (the streaming merge
                                      runs beside them in a thread of its own) */
  double merge_ms;                 /* streaming merge left after the last build, final merge */
  double total_ms;
  int32_t precision;               /* precision and tile the builds used */
  uint32_t tile;
  double device_ms;                /* summed device windows of the builds and prunes */
} tg_large_stats;
int tg_large_index(const tg_large_params* p, tg_large_stats* stats);

/* FP64 recall@k of the kNN graph over nq seeded sample nodes (tie at the k-th distance
 * counts as a hit).  Queries: the first nq of a seeded permutation prefix. */
int tg_recall(tg_state* s, uint32_t nq, uint64_t seed, uint32_t k, double* recall);

/* Largest relative difference |d32 - d64| / d64 between the stored FP32 distances of the
 * kNN rows of the same nq sample nodes as tg_recall (same seed) and their FP64 distances.
 * Measurement only: no counterpart in the reference, whose distances this judges. */
int tg_knn_dist_check(tg_state* s, uint32_t nq, uint64_t seed, double* max_rel_err);

/* Device buffers in the reference's capsule order (pybind.cu:132-141):
 * 0 graph u32[N*K], 1 reverse graph, 2 data f32[N*D], 3 distances f32[N*K],
 * 4 reverse distances, 5 reverse counts, 6 centroid f32[D], 7 sample lists u32[N*128]. */
int tg_buffer(tg_state* s, int32_t idx, void** dev_ptr, size_t* bytes);

/* Timing of the last tg_knn_build / tg_prune (device window, HIP events, ms) and the
 * summed duration of the dominant kernel (ms) and its launch count. */
typedef struct {
  double knn_ms, prune_ms;
  double knn_kernel_ms[8];   /* init, sample, rsample, localjoin, merge, reverse, phase2, final */
  double prune_kernel_ms[4]; /* ep, forward, reverse, other */
  uint64_t knn_evals;        /* distance evaluations of the last build */
  uint64_t knn_rescored;     /* canonical FP32 re-scores after the screen (phase 2 + join) */
  uint64_t prune_counters[15];/* collect, filter, reverse, adjacency rows, ep evaluations,
                                exact re-scores of screened DPG pairs; with profiling 2,
                                summed shader-clock cycles of the path / two-hop forward
                                kernel's collect (6), filter (7) and pool sort + merge (8)
                                phases, of the distance batches inside collect and filter
                                (9), of the beam rounds' candidate step, sort, merge and
                                id-set rebuild (10-13) and of the pool's sorts (14) */
  double learn_ms[5];        /* learned init: sample graph, training, encoding, buckets, seeding
                                (filled when profiling is on) */
  float learn_loss[2];       /* training loss at the end of the first and the last epoch */
  uint32_t learn_buckets;    /* code-space buckets */
  uint32_t learn_steps;      /* Adam steps */
} tg_stats;
int tg_get_stats(tg_state* s, tg_stats* out);
/* 0: only the windows; 1: time each kernel with events (adds a little overhead); 2: as 1,
 * plus the shader-clock phase counters of the prune forward kernel (adds work to it). */
int tg_set_profiling(tg_state* s, int on);

/* Codes [N*32] of the last learned initialisation (FP32). */
int tg_get_codes(tg_state* s, float* codes);

/* Learned-init encoder checks (no counterpart in the reference, which uses random init).
 * tg_mlp_eval: one forward/backward of the encoder D->256->32 on host rows X[M*D],
 * M = B*(1+P) (rows [0,B) anchors, row B+b*P+j positive j of anchor b), normalisation
 * (x-mu)*scale, flat parameters [256*(D+1)] (W1 | b1 per hidden unit) then [32*257]
 * (W2 | b2 per code); returns the loss, its gradient (same layout) and the codes [M*32].
 * tg_mlp_adam: one Adam step t (beta 0.9/0.999, eps 1e-8) of the flat block in place. */
int tg_mlp_eval(int device, uint32_t D, uint32_t B, uint32_t P, const float* X, const float* mu,
                float scale, const float* params, float margin, float* loss, float* grad,
                float* codes);
int tg_mlp_adam(int device, uint32_t D, uint32_t t, float lr, const float* grad, float* params,
                float* m, float* v);

/* ------------------------------------------------------------------ search validator
 * Beam search over the last pruned graph of s, restating get_neighbors (common.cu:65-127),
 * the NSG search the reference carries for its entry point but never runs on its indexes.
 * Initial pool: the first min(L, deg) neighbours of the entry point, then ids drawn from a
 * per-query SplitMix64 stream of `seed` (the reference: rand() % N) that skip visited ids;
 * the pool is kept sorted by (canonical squared distance, id) (the reference: sqrt of an
 * FP32 sum); the first unchecked entry is expanded until none is left; the first k entries
 * are returned.  queries [nq*D] (host), ids / dists [nq*k] (host; dists may be NULL). */
#define TG_INV_ID 0xFFFFFFFFu
typedef struct {
  uint32_t L;      /* pool size, 1..256 */
  uint32_t k;      /* results per query, 1..L */
  uint64_t seed;   /* random fill of the initial pool */
  uint32_t ep;     /* entry point; TG_INV_ID: the pruned graph's (node 0 without one) */
} tg_search_params;
typedef struct {
  double ms;            /* device time of the search kernel (HIP events) */
  uint64_t evals;       /* canonical distance evaluations */
  uint64_t expansions;  /* expanded pool entries */
} tg_search_stats;
int tg_search(tg_state* s, uint32_t nq, const float* queries, const tg_search_params* p,
              uint32_t* ids, float* dists, tg_search_stats* stats);
/* Allocation-free variant of tg_search.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernel and the same results as tg_search, but
 * queries_dev [nq*D] f32, ids_dev [nq*k] u32 and dists_dev [nq*k] f32 are DEVICE buffers
 * the caller owns, and the search's scratch (visited bitmaps, visited-id lists, counters)
 * lives in a 64 MiB static device array that each GPU the library runs on holds until the
 * process ends; the call allocates and frees nothing.  At most
 * (64 MiB - 256) / ((8192 + ceil(N/32)) * 4) workgroups run at once (TG_ERR_CAPACITY when
 * not even one fits, i.e. N above about 5.3e8).  stream: a hipStream_t, or NULL for the
 * handle's stream.  stats == NULL: the call only enqueues work and returns; otherwise it
 * waits for the search and fills stats (evals, expansions; ms is not measured, 0).  Every
 * call on one GPU shares that GPU's array: calls that overlap in time on one GPU (other
 * streams, other host threads, other handles) race and must be serialised by the caller. */
int tg_search_noalloc(tg_state* s, uint32_t nq, const float* queries_dev,
                      const tg_search_params* p, uint32_t* ids_dev, float* dists_dev,
                      void* stream, tg_search_stats* stats);
/* recall@k (k <= 16) of result rows ids [nq*k] for host queries [nq*D] against an FP64
 * brute force over all N vectors of s (a tie at the k-th distance counts as a hit). */
int tg_search_recall(tg_state* s, uint32_t nq, const float* queries, uint32_t k,
                     const uint32_t* ids, double* recall);

/* Self-test of the register sorts the prune kernels use (no counterpart in the reference).
 * keys/out: host [groups * 64 * npl] u64, npl 1, 2 or 8.  mode 0: sort each group
 * ascending; mode 1 (npl 1): each group is a bitonic sequence, sort it; mode 2 (npl 2): the
 * two halves of a group are ascending lists, the first half of the output is the 64 smallest
 * keys of both, ascending (the second half is returned unchanged). */
int tg_selftest_sort(int device, uint32_t npl, uint32_t groups, int32_t mode,
                     const uint64_t* keys, uint64_t* out);

#ifdef __cplusplus
}
#endif

#endif /* TAGORE_H */
