/* efg.h — C ABI of the MI355X (gfx950) EFANNA library, libefg.so.
 *
 * Drop-in boundary for efanna::FIndex<float> with the KDTREE_UB index
 * (efanna.hpp:9-75, algorithm/kdtreeub_index.hpp).  Every entry point below
 * names the reference call it replaces.  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every call returns an int status: EFG_OK (0) or a negative EFG_ERR_*;
 *    efg_last_error() returns the message of the calling thread's last failure.
 *    The library never calls exit() (the reference exits on I/O errors).
 *  - Calls are synchronous: they return when the device work is complete,
 *    except the *_device entry points, which are stream-ordered on `stream`
 *    (a hipStream_t; NULL = the index's own stream).
 *  - An index owns one HIP stream and device copies of the data, trees and
 *    graph; distinct indexes may be used from distinct host threads.
 *  - Ids are 32-bit; n must be below 2^30.
 */
#ifndef EFG_H_
#define EFG_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EFG_OK 0
#define EFG_ERR_ARG -2     /* invalid argument or parameter combination */
#define EFG_ERR_STATE -3   /* e.g. merge level deeper than a tree (kdtreeub_index.hpp:108-111) */
#define EFG_ERR_TREES -4   /* "wrong tree number" (kdtreeub_index.hpp:367-369) */
#define EFG_ERR_METHOD -5  /* "no such searching method" (kdtreeub_index.hpp:357-358) */
#define EFG_ERR_IO -6      /* file cannot be opened/read/written (kdtreeub_index.hpp:141,202) */
#define EFG_ERR_FORMAT -7  /* malformed file (kdtreeub_index.hpp:179,234) */
#define EFG_ERR_DEVICE -8  /* HIP runtime failure */
#define EFG_ERR_MEMORY -9  /* device allocation failed */

typedef struct efg_index efg_index;
typedef struct efg_batch_plan efg_batch_plan;

/* KDTreeUbIndexParams (algorithm/kdtreeub_index.hpp:14-35) + IndexParams
 * (general/params.hpp:19-33).  `seed` keys the counter-based generator that
 * replaces rand()/mt19937 (1998 is the reference's seed). */
typedef struct efg_build_params {
  int trees;        /* tree_num_total  -> TreeNum (kdtreeub_index.hpp:46-58) */
  int trees_build;  /* tree_num_build  -> TreeNumBuild; 0 means `trees` */
  int merge_level;  /* ml (conquer-to-depth level) */
  int epochs;       /* build_epoches (NN-descent iterations) */
  int check;        /* Check_K */
  int L;            /* pool size */
  int K;            /* graph degree */
  int S;            /* initial neighbour-set size */
  int tns;          /* TNS, DFS leaf size (params.hpp:25, default 10) */
  uint64_t seed;
} efg_build_params;

/* Default parameters of KDTreeUbIndexParams(true, trees). */
void efg_default_build_params(efg_build_params* p, int trees);

/* Message of the calling thread's last failure ("" if none). */
const char* efg_last_error(void);

/* FIndex ctor (efanna.hpp:14-20) + KDTreeUbIndex ctor (kdtreeub_index.hpp:42-90).
 * `data` is n rows of d floats, row i at data + i*row_stride (host memory).
 * The library copies it to the device (the reference keeps a non-owning
 * view, general/matrix.hpp:13-38); the caller's buffer may be freed after. */
int efg_create(const float* data, size_t n, size_t d, size_t row_stride, const efg_build_params* params,
               efg_index** out);
/* ~FIndex (the reference never frees its index, efanna.hpp:22-23). */
int efg_destroy(efg_index* ix);

/* FIndex::buildIndex -> KDTreeUbIndex::buildIndexImpl (kdtreeub_index.hpp:92-113):
 * trees, conquer-to-depth merge, random init, NN-descent refine; the graph
 * keeps the first K pool entries (base_index.hpp:469-479). */
int efg_build_index(efg_index* ix);
/* FIndex::buildTrees -> KDTreeUbIndex::buildTrees (kdtreeub_index.hpp:1192-1276). */
int efg_build_trees(efg_index* ix);

/* FIndex::setSearchParams (efanna.hpp:56-58 -> base_index.hpp:153-161):
 * extend_to is clamped to init_num; search_lv is stored and unused as in the
 * reference; search_trees == 0 is rejected (the reference divides by zero). */
int efg_set_search_params(efg_index* ix, int epochs, int init_num, int extend_to, int search_trees, int search_lv,
                          int search_method);
/* Current search parameters (method 1 raises init_num, kdtreeub_index.hpp:510-514). */
int efg_get_search_params(const efg_index* ix, int* epochs, int* init_num, int* extend_to, int* search_trees,
                          int* search_method);

/* FIndex::knnSearch (efanna.hpp:32-34 -> kdtreeub_index.hpp:349-361): host
 * queries (nq rows of the index's d floats, stride q_stride); results are
 * kept in the index (read them with efg_get_results / efg_save_results). */
int efg_knn_search(efg_index* ix, int k, const float* queries, size_t nq, size_t q_stride);
/* Results of the last efg_knn_search: nq*k ids (-1 where fewer were found)
 * and, if dists != NULL, squared distances.  Either output may be NULL to
 * query the shape only. */
int efg_get_results(const efg_index* ix, size_t* nq, int* k, int32_t* ids, float* dists);
/* Device-resident search: queries, ids and dists are device pointers; the
 * call is ordered on `stream` and does not synchronise.  ndist (device,
 * nullable) receives the distance evaluations per query. */
int efg_knn_search_device(efg_index* ix, int k, const float* d_queries, size_t nq, size_t q_stride, int32_t* d_ids,
                          float* d_dists, uint64_t* d_ndist, void* stream);

/* Hashing index (IEH seeds): replaces HASHINGIndex's constructor code loading
 * (algorithm/hashing_index.hpp:66-197, LoadCode32/64 206-321).  Turns the index
 * into a hashing index: `tables` hash tables of codes of `codelen` bits,
 * shifted right by `lenshift` (at most 2^(codelen - lenshift) - 1 each),
 * `upbits` < codelen - lenshift (checked as the reference checks it), Hamming
 * radius `radius` (clamped as the reference clamps it; masks x tables must
 * stay at or below 4096).  base_codes: tables x n codes, table-major;
 * query_codes: tables x nq, query i of a search uses codes column i.  After
 * efg_build_index (the tables, buildIndexImpl 874-889) efg_knn_search
 * seeds every query from the buckets of its code XOR each Hamming-ball mask,
 * mask outer, table inner (getNeighborsIEH32/64_*, 1087-1630):
 *   no graph loaded           -> the K nearest of the first init_num
 *                                candidates (getNeighbors32/64, 929-1085)
 *   graph, search_method == 1 -> nnexp expansion (getNeighborsIEH*_nnexp)
 *   graph, search_method == 0 -> kgraph expansion (getNeighborsIEH*_kgraph)
 * search_trees is not used.  Random seeds beyond the buckets come from the
 * counter-based generator (the reference calls rand()). */
int efg_hash_set_codes(efg_index* ix, int tables, int codelen, int lenshift, int upbits, int radius,
                       const uint64_t* base_codes, const uint64_t* query_codes, size_t nq);
/* The same from the reference's code files "<prefix>_<t>", t = 1..tables:
 * int32 1, int32 code length, int32 count, then count codes of 4 bytes
 * (codelen <= 32) or 8 bytes (HASHINGIndexParams(codelen, TableNum,
 * UpperBits, HashRadius, BaseCodeFile, QueryCodeFile, codelenShift),
 * hashing_index.hpp:19-46). */
int efg_hash_load_codes(efg_index* ix, int tables, int codelen, int lenshift, int upbits, int radius,
                        const char* base_prefix, const char* query_prefix);

/* Search precision (no counterpart in the reference; a form beside the plain
 * FP32 search): 0 = FP32 rows (default, bit-exact with the reference's
 * arithmetic); 1 = BF16 copy of the rows for the pool and expansion
 * distances, the best min(2k, pool) re-ranked by direct FP32 difference.  On
 * integer data up to 256 in magnitude BF16 is exact and both give the same
 * results.  Applies to method 0. */
int efg_set_search_precision(efg_index* ix, int precision);

/* Build precision (no counterpart in the reference; a form beside the plain
 * FP32 build): 0 = FP32 distance tiles (default, bit-exact with the
 * reference's arithmetic); 1 = BF16, 2 = FP16 MFMA tiles with FP32
 * accumulation for the merge's leaf tiles (leaves of at most 64 points) and
 * the join's pair tiles, as |x|^2 + |y|^2 - 2 x.y; the merged sets and the
 * final pools are then re-ranked by the direct FP32 difference.  On rows
 * exact in the 16-bit type whose sums stay below 2^24 (integers 0..255 with
 * d <= 128) the build equals the FP32 build.  Applies to later builds. */
int efg_set_build_precision(efg_index* ix, int precision);

/* Join staging of the build (no counterpart in the reference; sizing only):
 * each NN-descent iteration reserves room for `margin` times the fraction of
 * candidates the previous iteration kept (at most every pair's two), and
 * joins again with room for all of them when they do not fit (counter 7),
 * besides one partly used run of 4096 slots per wavefront.  margin > 0;
 * default 1.25.  Results do not depend on it. */
int efg_set_join_staging_margin(efg_index* ix, double margin);

/* Join candidate routing of the build (no counterpart in the reference; data
 * movement only).  1 (default): each kept candidate goes straight from the
 * join kernels into its target row's bucket, which has room for every pair
 * the row is in (2 x the iteration's pairs slots in all, 12 bytes each),
 * except the precision form's block tasks (nodes of more than 64 rows),
 * whose candidates are staged and scattered into the same buckets; an
 * iteration whose pairs need 2^31 slots or more, or whose buckets cannot be
 * allocated, is staged entirely.  0: always staged (candidates written in
 * per-wavefront runs, then scattered into buckets sized by the kept counts;
 * see efg_set_join_staging_margin).  Results do not depend on it. */
int efg_set_join_routing(efg_index* ix, int direct);

/* Learned stopping (no counterpart in the reference; a form beside the plain
 * search).  A small network (16 pool-state features -> 64 -> 64 -> 1 logit,
 * ReLU, 5,313 parameters laid out W1[16][64] | b1[64] | W2[64][64] | b2[64] |
 * W3[64] | b3, evaluated on BF16 parameters with FP32 accumulation) stands in
 * for the fixed epoch/pool stopping rule of method 0.  A probe search with a
 * small pool evaluates it after the seeds and after every epoch; a query
 * stops when the network's probability reaches the threshold.  Otherwise,
 * in search mode 1 it runs the full plain search (setSearchParams) from
 * scratch; in search mode 2 it continues in place: the rest of the full pool
 * (the next ids of the full search's leaf sequence not yet measured, from
 * where the probe stopped) joins the probe's pool, and the full search's
 * epochs follow with the probe's visited set.  Mode 2 needs the probe's pool
 * and extend_to within the search's. */
int efg_mlp_param_count(void);
/* Initial parameters from the generator (uniform Glorot weights, zero biases). */
int efg_mlp_init(float* params, uint64_t seed);
/* Logits of n feature rows (host buffers); norm = feature mean[16] | 1 / deviation[16]. */
int efg_mlp_logits(const float* params, const float* norm, const float* feat, size_t n, float* z);
/* Gradient of the mean binary cross-entropy over n labelled rows (host buffers). */
int efg_mlp_gradient(const float* params, const float* norm, const float* feat, const uint8_t* label, size_t n,
                     float* grad, float* loss);
/* Adam training on the GPU: `epochs` passes in generator order, `batch` rows
 * per step; params holds the initial values on entry and the trained ones
 * on return; loss_hist (nullable) gets the mean loss of every epoch. */
int efg_mlp_train(const float* feat, const uint8_t* label, size_t n, const float* norm, float* params, int epochs,
                  int batch, float lr, uint64_t seed, float* loss_hist);
/* Probe search parameters (0 = derived: pool max(k, init_num / 6), extend_to
 * max(1, extend_to / 4), the search's epochs). */
int efg_set_probe_params(efg_index* ix, int pool, int extend_to, int epochs);
/* Traces of the probe search for nq host queries: per query and checkpoint
 * (epochs + 1 of them) the 16 features and a label, the number of top-k ids
 * that are also in plain_ids[query] (host, nq x k; capped at 254), 255 for
 * checkpoints the search did not reach.  feat holds nq * (epochs + 1) * 16
 * floats. */
int efg_search_traces(efg_index* ix, int k, const float* queries, size_t nq, size_t q_stride,
                      const int32_t* plain_ids, float* feat, uint8_t* label);
/* Network used when the search mode is 1; a query stops at probability >= threshold. */
int efg_set_stopping_network(efg_index* ix, const float* params, const float* norm, float threshold);
/* Search mode for method 0: 0 = plain (default), 1 = probe search + network + full search from scratch,
 * 2 = probe search + network + continuation in place. */
int efg_set_search_mode(efg_index* ix, int mode);
/* Queries of the last learned search that the network did not stop, and running totals. */
int efg_get_learned_stats(efg_index* ix, uint64_t* last_fallbacks, uint64_t* total_fallbacks,
                          uint64_t* total_queries);
/* Queries of the last learned search that ran the full search from scratch:
 * in mode 1 every query not stopped; in mode 2 those whose visited set
 * outgrew the workgroup's table or with too few rows left to continue. */
int efg_get_learned_restarts(efg_index* ix, uint64_t* last_restarts);

/* Heterogeneous batch (no counterpart in the reference; the reference runs
 * one configuration per knnSearch): query q runs method 0 with init_num
 * pool[q], extend_to extend[q] (nullable: max(1, pool[q] / 6)) and k[q], the
 * epochs and trees of setSearchParams.  Queries are grouped by (pool,
 * extend_to, k) classes.  By default each class runs its own persistent
 * pass (threads per query tuned per class when tune != 0), the classes on
 * four parallel streams; efg_set_batch_form selects the one-pass form
 * instead.  Row q of ids/dists (out_stride entries, >= every k[q]) holds the
 * k[q] results, then -1 / FLT_MAX.  pool, extend, k are host arrays; the
 * _device variant takes device queries and outputs and is ordered on
 * `stream`. */
int efg_knn_search_batch(efg_index* ix, const float* queries, size_t nq, size_t q_stride, const int* pool,
                         const int* extend, const int* k, int out_stride, int32_t* ids, float* dists, int tune);
int efg_knn_search_batch_device(efg_index* ix, const float* d_queries, size_t nq, size_t q_stride, const int* pool,
                                const int* extend, const int* k, int out_stride, int32_t* d_ids, float* d_dists,
                                int tune, void* stream);
/* A heterogeneous batch captured into a HIP graph (no counterpart in the
 * reference), for serving loops that search batch after batch of the same
 * shape: create runs the batch once (tuning first when tune != 0) on
 * `stream`, then captures its launches; run replays them on `stream`.  The
 * graph refers to d_queries, d_ids and d_dists: write the next batch's
 * queries into d_queries before each run.  pool, extend and k are read at
 * creation only.  A plan uses a workspace of its own; it must be destroyed
 * before its index, and runs of one plan must not overlap. */
int efg_batch_plan_create(efg_index* ix, const float* d_queries, size_t nq, size_t q_stride, const int* pool,
                          const int* extend, const int* k, int out_stride, int32_t* d_ids, float* d_dists, int tune,
                          void* stream, efg_batch_plan** plan);
int efg_batch_plan_run(efg_batch_plan* plan, void* stream);
int efg_batch_plan_destroy(efg_batch_plan* plan);
/* Form of the batched search (no counterpart in the reference): 0 = one
 * pass per class, the classes on four parallel streams, each with its own
 * threads per query (default); 1 = one pass over all classes: one descent,
 * then one persistent kernel whose workgroups draw the queries of every
 * class from one order and take each query's configuration from its class
 * (with tune != 0 the tuner picks the shape: one group, or two split by pool
 * size, and the threads per query of each).  Results are the same. */
int efg_set_batch_form(efg_index* ix, int one_pass);
/* The one-pass form's shape in use (no counterpart in the reference): the
 * largest pool of the small group (0: one group), the threads per query of
 * the small and the large group, and the small group's share of its
 * resident workgroups in percent. */
int efg_get_batch_shape(const efg_index* ix, int* split, int* t_small, int* t_large, int* share_pct);

/* Search without device allocation (no counterpart in the reference).  For
 * callers that cannot allocate device memory: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  The same search, with
 * the same results, as efg_knn_search_device (method 0 or 1, FP32 or BF16
 * rows; kd-tree index, plain mode), ordered on `stream`, but every scratch
 * buffer lives in one fixed device array of EFG_NOALLOC_SCRATCH_BYTES per GPU,
 * reserved when the library is loaded and held until the process ends; the
 * call allocates and frees nothing.  Limits: nq * (search_trees * leaves per
 * tree + 4) * 4 bytes plus one global visited table must fit the array (about
 * 30,000 queries at 16 trees and pool 1200); larger batches are refused with
 * EFG_ERR_ARG.  The index must have searched once with the allocating entry
 * points (that lays out the search trees, and the BF16 rows for precision 1),
 * else EFG_ERR_STATE.  Calls on one GPU share the array and its last contents:
 * calls that overlap on one GPU race.  The work counters of
 * efg_get_counters / efg_get_search_work do not see these calls. */
#define EFG_NOALLOC_SCRATCH_BYTES (64u << 20)
int efg_knn_search_device_noalloc(efg_index* ix, int k, const float* d_queries, size_t nq, size_t q_stride,
                                  int32_t* d_ids, float* d_dists, void* stream);

/* Threads per query of the search kernels (no counterpart in the reference;
 * launch tuning only, results do not change): 64, 128 or 256 forces one;
 * 0 (default) uses the value tuned for the configuration, else one
 * wavefront for pools up to 256, two up to 1024, four above. */
int efg_set_search_threads(efg_index* ix, int threads);
/* Times 64, 128 and 256 threads per query on nq host queries for the current
 * search parameters and mode (learned mode: the probe and the full search)
 * and keeps the fastest per configuration.  threads_out[0] (full search) and
 * threads_out[1] (probe, 0 in plain mode) receive the choice (nullable). */
int efg_autotune_search(efg_index* ix, int k, const float* queries, size_t nq, size_t q_stride, int* threads_out);

/* getGraphSize / getGraphRow (base_index.hpp:494-501). */
int efg_graph_size(const efg_index* ix, size_t* n);
int efg_graph_row(const efg_index* ix, size_t row, uint32_t* out, size_t cap, size_t* len);

/* File formats of the reference (byte compatible):
 *  graph   kdtreeub_index.hpp:295-312 (save) / 243-277 (load)
 *  trees   kdtreeub_index.hpp:197-242, 658-671 (save) / 139-196, 672-700 (load)
 *  index   kdtreeub_index.hpp:771-825 (save) / 701-770 (load): trees + graph
 *  results base_index.hpp:136-151 */
int efg_save_graph(const efg_index* ix, const char* path);
int efg_load_graph(efg_index* ix, const char* path);
int efg_save_trees(const efg_index* ix, const char* path);
int efg_load_trees(efg_index* ix, const char* path);
int efg_save_index(const efg_index* ix, const char* path);
int efg_load_index(efg_index* ix, const char* path);
int efg_save_results(const efg_index* ix, const char* path);

/* Index image: saveIndex / loadIndex (kdtreeub_index.hpp:129-137, 701-825)
 * as one device-loadable file.  The image holds the arrays the search reads
 * as they sit in device memory (depth-first search trees, descent nodes, leaf
 * lists, CSR graph) and the build-order trees, in 4 KiB-aligned sections with
 * checksums; loading streams the file through pinned host memory straight
 * into HBM, with no parsing or re-layout, and a device pass range-checks
 * every id and node link before the index takes the arrays (a damaged file
 * gives EFG_ERR_FORMAT and leaves the index unchanged).  The index must hold
 * the same data (rows and dimension) as the one that saved the image.  The
 * reference formats (efg_save_index, efg_save_trees, efg_save_graph) stay the
 * interchange formats; images are for this library's own fast reloads. */
int efg_save_image(efg_index* ix, const char* path);
int efg_load_image(efg_index* ix, const char* path);

/* Exact k nearest neighbours by brute force over device rows (the oracle
 * the reference only has as Matrix::brute_force_search, general/matrix.hpp:57-68):
 * k smallest (distance, id), ties to the lower id; exclude_self skips row q for
 * query q.  Base rows must be padded with zeros to a multiple of 8 floats. */
int efg_exact_knn_device(const float* d_base, size_t n, size_t d, size_t base_stride, const float* d_queries,
                         size_t nq, size_t q_stride, int k, int exclude_self, int32_t* d_ids, float* d_dists,
                         void* stream);

/* Exact kNN, precision form (no counterpart in the reference): dot products
 * on MFMA tiles in BF16 (precision 1) or FP16 (precision 2) with FP32
 * accumulation, distances as |q|^2 + |x|^2 - 2 q.x, the best 2k per query
 * re-ranked by the direct FP32 difference (the reference's order);
 * precision 0 runs efg_exact_knn_device.  Same arguments otherwise; k <= 512. */
int efg_exact_knn_device_prec(const float* d_base, size_t n, size_t d, size_t base_stride, const float* d_queries,
                              size_t nq, size_t q_stride, int k, int exclude_self, int precision, int32_t* d_ids,
                              float* d_dists, void* stream);

/* ---- inspection (tests, tools) ------------------------------------------ */
/* Trees as flat arrays over all nodes (leaves: dim = left = right = -1). */
int efg_tree_shape(const efg_index* ix, int* trees, size_t* nodes);
int efg_get_trees(const efg_index* ix, int32_t* dim, float* val, int32_t* left, int32_t* right, uint32_t* start,
                  uint32_t* end, int32_t* roots, uint32_t* leaf_lists);
int efg_set_trees(efg_index* ix, int trees, size_t nodes, const int32_t* dim, const float* val, const int32_t* left,
                  const int32_t* right, const uint32_t* start, const uint32_t* end, const int32_t* roots,
                  const uint32_t* leaf_lists);
/* Graph as CSR: off[n+1], ids[off[n]]. */
int efg_get_graph(const efg_index* ix, uint64_t* off, uint32_t* ids, size_t cap);
int efg_set_graph(efg_index* ix, const uint64_t* off, const uint32_t* ids);
/* Device pointer of the padded base rows and their stride (floats). */
int efg_device_data(const efg_index* ix, const float** d_data, size_t* stride);
/* Counters: [0] distances in the last search, [1] merge distances, [2] init
 * distances, [3] join distances (pairs), [4] queries of the last search or
 * batch (all its passes) whose visited set outgrew the LDS table (in a pass
 * over a query list they move to a spill table in global memory and go on;
 * otherwise they run again with a table in global memory), [5] those that
 * outgrew the largest global table too (their result rows are -1 /
 * FLT_MAX), [6] join rows, [7] join iterations whose candidates outgrew the
 * staging estimate and were joined again. */
int efg_get_counters(const efg_index* ix, uint64_t* out, int n);
/* Per-iteration join pair counts of the last build. */
int efg_get_join_pairs(const efg_index* ix, uint64_t* out, int cap, int* n);
/* Build stages for stepwise checking: 0 trees, 1 merge+init, 2 one refine
 * iteration, 3 write the graph. */
int efg_build_stage(efg_index* ix, int stage);
/* Build state after a stage: pools as n x L keys (float distance bits << 32 |
 * id, all-ones when empty), flags, fill counts, Range; join lists as CSR
 * (which = 0 new, 1 old). */
int efg_get_pools(const efg_index* ix, uint64_t* keys, uint8_t* flags, uint32_t* fill, uint32_t* range);
int efg_get_lists(const efg_index* ix, int which, uint64_t* off, uint32_t* ids, size_t cap);
/* Phase timings of the last build in milliseconds: [trees, merge+init, join, update]. */
int efg_get_build_times(const efg_index* ix, double* out, int n);
/* Kernel timing of searches (HIP events on the search stream).  While on,
 * every search records its stages; efg_get_kernel_times returns the summed
 * milliseconds [tree descent, search kernel, overflow re-run], the number of
 * searches recorded, and clears the record. */
int efg_set_profiling(efg_index* ix, int on);
int efg_get_kernel_times(efg_index* ix, double* ms, int n, int* searches);
/* Algorithmic work of the searches since the last call: [distances, leaf ids
 * read for seeds, nodes expanded]; clears the counters. */
int efg_get_search_work(efg_index* ix, uint64_t* out, int n);

#ifdef __cplusplus
}
#endif
#endif /* EFG_H_ */
