// Host-side internals of libefg.so: the index object and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <array>
#include <map>
#include <memory>
#include <vector>

#include "efg.h"

namespace efg {

// Device memory cache behind DevBuf.  Freed blocks are kept per device and
// size class instead of going back to the runtime (hipFree synchronises the
// device and unmaps; hipMalloc maps again).  A freed block is reused only
// after the device has been synchronised once since its release, so no block
// is handed out while an earlier kernel may still use it.  On an allocation
// failure the cache is emptied and the allocation retried.
void* dev_alloc(size_t bytes);
void dev_release(void* p, size_t bytes);

// Owning device buffer, or a non-owning view of memory held elsewhere (a
// view never allocates: ensure() only checks that it is large enough).
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  bool owned = true;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p && owned) dev_release(p, n * sizeof(T));
    p = nullptr;
    n = 0;
  }
  void view(T* ptr, size_t count) {
    reset();
    owned = false;
    p = ptr;
    n = count;
  }
  // grow-only (contents not preserved)
  bool ensure(size_t count) {
    if (count <= n && p) return true;
    if (!owned) return false;
    reset();
    if (count == 0) count = 1;
    p = static_cast<T*>(dev_alloc(count * sizeof(T)));
    if (!p) return false;
    n = count;
    return true;
  }
  bool alloc(size_t count) {
    reset();
    return ensure(count);
  }
};

// Tree node in the search layout: nodes of one tree in depth-first pre-order
// (left child first), so a subtree occupies a contiguous block.
struct alignas(32) SNode {
  int32_t dim;
  float val;
  int32_t left, right, parent;
  uint32_t start, end;
  int32_t pad;
};

// The same nodes, same positions, reduced to what the descent reads: the
// left child of an internal node is the next node (pre-order), dim < 0 marks
// a leaf.
struct alignas(16) DNode {
  int32_t dim;
  float val;
  int32_t right, parent;
};

// Search launch configuration (all sizes in elements).
struct SearchCfg {
  int n;        // base rows
  int d, D;     // dimension, padded row stride (multiple of 8)
  int T;        // trees searched
  int lsize;    // leaves per tree (kdtreeub_index.hpp:392)
  int nleaf;    // T * lsize
  int init_num; // seed pool
  int E;        // extend_to
  int epochs;
  int K;        // results per query
  int R;        // pool kept: max(extend_to, K)
  int Rp;       // R rounded up to a power of two
  int kb;       // key buffer (seeds, then pool / new / merge areas)
  int stop;     // 0 plain; 1 probe search stopped by the network; 2 probe search writing traces
  int threads;  // threads per query (one workgroup): 64, 128 or 256
  int out_stride;  // results per query row in the output (>= K)
  int logcap;   // visited-set capacity (log2)
  uint64_t seed;
  int method;   // 0 nnexp, 1 kgraph
  int sorted;   // method 1: seeds reduced to the best epochs * extend_to
  int deg;      // graph rows all of this degree at offsets row * deg (0: read the offsets)
};

// Device state of the NN-descent build (see k_graph.hip).
struct BuildState {
  int L = 0;                    // pool capacity
  DevBuf<uint64_t> pool;        // n x L keys
  DevBuf<uint8_t> flag;         // n x L
  DevBuf<uint32_t> fill, range; // n
  DevBuf<float> radiusM;        // n
  DevBuf<uint8_t> found;        // n
  DevBuf<uint64_t> new_off, old_off;  // CSR join lists
  DevBuf<uint32_t> new_ids, old_ids;
  size_t new_n = 0, old_n = 0;
  // join scratch, grown as needed and released with the finished graph:
  // task headers, task row ids (64 per task), staged candidates, buckets
  DevBuf<uint4> jhdr;
  DevBuf<uint32_t> jtids;
  DevBuf<unsigned char> jstage;
  DevBuf<uint64_t> jbkey;
  DevBuf<uint32_t> jbsrc;
  // per-iteration scratch of the join and the list update (n or n + 1
  // entries, the reverse lists as needed), kept across iterations so that
  // no iteration allocates or frees device memory once they have grown
  struct IterScratch {
    DevBuf<uint64_t> worst;
    DevBuf<uint32_t> tcount, cursor;
    DevBuf<int32_t> toff, sums, scal, tcnt, tskoff;
    DevBuf<unsigned long long> ctr;
    DevBuf<uint32_t> fnew, fold, rcn, rco, rszn, rszo, ridn, rido;
    DevBuf<uint32_t> rslot;  // n x L: slot of each pool entry in its reverse list (see k_lists)
    DevBuf<int32_t> roffn, roffo, snew, sold, onew, oold, bign, bigo;
    DevBuf<int> nbig;
    void release() {
      for (DevBuf<uint32_t>* b :
           {&tcount, &cursor, &fnew, &fold, &rcn, &rco, &rszn, &rszo, &ridn, &rido, &rslot})
        b->reset();
      for (DevBuf<int32_t>* b :
           {&toff, &sums, &scal, &tcnt, &tskoff, &roffn, &roffo, &snew, &sold, &onew, &oold, &bign, &bigo})
        b->reset();
      nbig.reset();
      worst.reset();
      ctr.reset();
    }
  };
  IterScratch it;
  int iter = 0;
  double keep_frac = 1.0;  // join candidates kept / 2 x pair bound, last iteration (sizes the staging)
  bool ready = false;
};

struct Workspace {
  DevBuf<int32_t> leaves, overflow, qorder, qkey, qoff, scan_sums;
  DevBuf<uint32_t> qhist;
  DevBuf<int> nover;
  DevBuf<int> nover_sum;  // [overflow queries, second overflows] summed over the passes of one search or batch
  DevBuf<unsigned long long> ndist;
  DevBuf<uint32_t> gtab;
  DevBuf<uint32_t> spill;  // per-workgroup tables that queries outgrowing the LDS table move into
  DevBuf<float> qbuf;
  DevBuf<int32_t> ids;
  DevBuf<float> dists;
  DevBuf<int32_t> blist;  // batched search: queries per class, concatenated
  DevBuf<int> bcount;     // batched search: queries per class
  // batched search in one pass: class configurations, the class and the
  // leaves per tree of every query, the order the workgroups draw queries in
  DevBuf<SearchCfg> bcfg;
  DevBuf<uint8_t> bcls;
  DevBuf<int> blsize;
  DevBuf<int32_t> border;
};

}  // namespace efg

struct efg_index {
  size_t n = 0, d = 0, D = 0;
  efg_build_params bp{};
  // search parameters (base_index.hpp:153-161)
  int s_epochs = 0, s_init = 0, s_extend = 0, s_trees = 0, s_lv = -1, s_method = 0;
  bool s_set = false;
  hipStream_t stream = nullptr;
  efg::DevBuf<float> data;  // n x D, zero padded
  efg::DevBuf<uint16_t> data_bf16;  // BF16 copy, lane-split rows (precision form)
  int s_precision = 0;              // 0 FP32 rows, 1 BF16 rows + FP32 re-rank
  int b_precision = 0;              // build: 0 FP32 tiles, 1 BF16 / 2 FP16 MFMA tiles + FP32 re-rank
  double b_stage_margin = 1.25;     // join staging: margin over the last iteration's kept fraction
  bool b_join_direct = true;        // join candidates straight into their targets' buckets when they fit
  efg::DevBuf<uint16_t> split_hi, split_lo;  // build precision form: rows as hi + lo 16-bit parts (n x Dk)
  efg::DevBuf<float> split_norm;             // FP32 squared norms of the rows
  int split_Dk = 0, split_kind = 0;
  bool split_lo_used = true;                 // some low part is non-zero (else the tiles skip them)
  bool split_exact = false;                  // every tile distance is exact (integer rows, small norms)
  // trees
  int T = 0;
  size_t M = 0;
  efg::DevBuf<int4> node4;       // (dim, val bits, left, right)
  efg::DevBuf<int32_t> parent;
  efg::DevBuf<uint2> range;      // [start, end)
  efg::DevBuf<int32_t> roots;
  efg::DevBuf<uint32_t> leaf;    // T x n
  bool leaf0_perm = false;       // tree 0's leaf list holds every row once
  std::vector<int8_t> leaf_dim;  // value written as DivDim of a leaf (0 or -1)
  // after a device tree build leaf_dim is derived when the trees are read
  // back (saving): node levels, the BFS->DFS switch level and S
  bool leaf_dim_lazy = false;
  std::vector<int8_t> node_level;
  int tree_lb = -1, tree_S = 0;
  efg::DevBuf<efg::SNode> snode;  // search layout of the trees
  efg::DevBuf<efg::DNode> dnode;  // descent copy of snode
  efg::DevBuf<int32_t> sroots;
  bool snode_valid = false;
  // hashing index (IEH seeds, hashing_index.hpp): the seed sequence of a
  // query is the buckets of its code XOR each Hamming-ball mask, mask outer,
  // table inner.  Each table's rows are grouped by code (codes ascending, ids
  // ascending inside a bucket); bucket b of table t is a range [start, end)
  // of that table's list, kept in an SNode so the search kernel reads it as
  // it reads a leaf.
  bool hashing = false;
  int h_T = 0, h_bits = 0, h_radius = 0, h_nmask = 0;
  size_t h_nq = 0;                      // queries with codes
  std::vector<uint64_t> h_base;         // T x n codes, shifted (until the tables are built)
  std::vector<uint64_t> h_query;        // T x h_nq codes, shifted
  bool h_built = false;
  efg::DevBuf<uint64_t> h_ucode;        // per table the distinct codes, ascending, tables concatenated
  efg::DevBuf<uint32_t> h_uoff;         // T + 1: first bucket of each table
  efg::DevBuf<efg::SNode> h_nodes;      // bucket ranges (start, end)
  efg::DevBuf<uint32_t> h_lists;        // T x n rows grouped by bucket
  efg::DevBuf<uint64_t> h_masks;        // h_nmask masks in the reference's order
  efg::DevBuf<uint64_t> h_qcodes;       // T x h_nq query codes
  efg::DevBuf<int32_t> h_dense;         // codes of at most kHashDenseBits bits: T x 2^bits bucket index or -1
  // graph (CSR)
  size_t nnz = 0;
  int max_deg = 0;
  int reg_deg = 0;  // degree of every row when the offsets are row * degree, else 0
  efg::DevBuf<uint64_t> goff;
  efg::DevBuf<uint32_t> gids;
  // build
  efg::BuildState bs;
  std::vector<uint64_t> join_pairs;
  double times[4] = {0, 0, 0, 0};
  uint64_t counters[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  // results of the last host search
  std::vector<int32_t> res_ids;
  std::vector<float> res_d;
  size_t res_nq = 0;
  int res_k = 0;
  efg::Workspace ws;
  // search profiling: events per recorded search (4 each)
  bool profiling = false;
  std::vector<hipEvent_t> prof_events;
  efg::DevBuf<unsigned long long> work;  // [distances, seed ids read, expanded]
  // learned stopping: probe search parameters, network, mode
  int lp_pool = 0, lp_extend = 0, lp_epochs = 0;  // 0 = derived from the search parameters
  int s_mode = 0;                                // 0 plain, 1 probe search + network + fallback
  bool net_set = false;
  float net_zthr = 0.f;                          // stop when the logit reaches this value
  efg::DevBuf<uint32_t> net_pack;                // BF16 parameters packed for the search kernel
  efg::DevBuf<float> net_norm;                   // feature mean | 1 / deviation
  efg::DevBuf<int32_t> fb_list;                  // queries handed to the full search
  efg::DevBuf<int> fb_count;                     // [queries handed to the full search by the last search, total not
                                                 //  stopped, total learned queries, not stopped in the last search]
  bool fb_count_init = false;
  // threads per query: forced (64/128/256), else tuned per configuration, else a rule of thumb
  int s_threads = 0;
  std::map<std::array<int, 6>, int> tuned;  // (init_num, extend_to, epochs, k, trees, stop) -> threads
  // batched search: the classes run on parallel streams (default priority),
  // forked from and joined to the caller's stream with events; one
  // workspace per class for the eager batches (plans keep their own)
  std::vector<hipStream_t> bstreams;
  hipEvent_t bfork = nullptr;
  std::vector<hipEvent_t> bjoin;
  std::vector<std::unique_ptr<efg::Workspace>> bws;
  // batched search form: 0 one pass per class on the parallel streams
  // (default), 1 one pass over all classes; the one-pass shape chosen by the
  // tuner: classes with pools up to b_split in a group at b_tsmall threads
  // per query, the rest at b_tlarge (b_split 0: one group)
  int b_mode = 0;
  int b_split = 0, b_tsmall = 128, b_tlarge = 256;
  float b_share = 1.f;  // two groups: the small group's part of its resident workgroups
};

namespace efg {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

// k_hash.hip
constexpr int kHashSlots = 4096;     // largest masks x tables per query
constexpr int kHashDenseBits = 22;   // direct bucket lookup up to this code length
int hash_set_codes(efg_index* ix, int tables, int codelen, int lenshift, int upbits, int radius, const uint64_t* base,
                   const uint64_t* query, size_t nq);
int hash_build(efg_index* ix);
int hash_probe(efg_index* ix, int nleaf, size_t nq, int32_t* leaves, const int32_t* qlist, const int* nlist,
               hipStream_t st);
// efg_image.cpp: the index image (device arrays of a built index in one file)
int save_image(efg_index* ix, const char* path);
int load_image(efg_index* ix, const char* path);
// k_search.hip
int search_device(efg_index* ix, int k, const float* dq, size_t nq, size_t q_stride, int32_t* d_ids, float* d_dists,
                  uint64_t* d_ndist, hipStream_t st);
// The same search with every scratch buffer in the library's fixed device
// array (no device allocation); see efg_knn_search_device_noalloc.
int search_device_noalloc(efg_index* ix, int k, const float* dq, size_t nq, size_t q_stride, int32_t* d_ids,
                          float* d_dists, hipStream_t st);
// k_exact.hip
int exact_knn(const float* X, size_t n, int d, size_t x_stride, const float* Q, int nq, size_t q_stride, int k,
              int exclude_self, int32_t* ids, float* dists, hipStream_t st, std::string* err);
// precision 1 = BF16, 2 = FP16 MFMA tiles, then FP32 re-rank of the best 2k
int exact_knn_mfma(const float* X, size_t n, int d, size_t x_stride, const float* Q, int nq, size_t q_stride, int k,
                   int exclude_self, int precision, int32_t* ids, float* dists, hipStream_t st, std::string* err);
// Heterogeneous batch: per-query pool (init_num), extend_to (nullable: pool / 6)
// and k (host arrays); results in rows of out_stride.  With tune, every
// (pool, extend_to, k) class first times 64/128/256 threads per query.
int search_batch(efg_index* ix, const float* dq, size_t nq, size_t q_stride, const int* pool, const int* ext,
                 const int* kq, int out_stride, int32_t* d_ids, float* d_dists, bool tune, hipStream_t st);
// The same batch captured into a HIP graph (after one eager run, and tuning
// when asked); the graph refers to dq, d_ids and d_dists and to the plan's
// own workspace.
int batch_plan_create(efg_index* ix, const float* dq, size_t nq, size_t q_stride, const int* pool, const int* ext,
                      const int* kq, int out_stride, int32_t* d_ids, float* d_dists, bool tune, hipStream_t st,
                      efg_batch_plan** out);
int batch_plan_run(efg_batch_plan* pl, hipStream_t st);
void batch_plan_destroy(efg_batch_plan* pl);
// Times the search with 64, 128 and 256 threads per query on the host queries
// (a sample of them) for the current parameters and mode; keeps the fastest.
int search_autotune(efg_index* ix, int k, const float* dq, size_t nq, size_t q_stride, int* chosen,
                    hipStream_t st);
// Probe search traces: per query and checkpoint, features and whether the
// top k then equals plain_ids (device buffers).
int search_traces(efg_index* ix, int k, const float* dq, size_t nq, size_t q_stride, const int32_t* d_plain,
                  float* d_feat, uint8_t* d_label, hipStream_t st);
// k_mlp.hip
int mlp_logits(const float* params, const float* norm, const float* feat, size_t n, float* z, hipStream_t st);
int mlp_gradient(const float* params, const float* norm, const float* feat, const uint8_t* label, size_t n,
                 float* grad, float* loss, hipStream_t st);
int mlp_train(const float* feat, const uint8_t* label, size_t n, const float* norm, float* params, int epochs,
              int batch, float lr, uint64_t seed, float* loss_hist, hipStream_t st);
void mlp_init(float* params, uint64_t seed);
// k_tree.hip
int gpu_build_trees(efg_index* ix);
// Exclusive scan of m ints into out[0..m]; out[m] and *total_dev get the total.
int device_exclusive_scan(const int32_t* in, int m, int32_t* out, int32_t* total_dev, DevBuf<int32_t>& sums,
                          hipStream_t st);
// k_graph.hip
// The split rows of the build's precision form (hi + lo 16-bit parts, n x
// split_Dk), made once per build; no-op for the FP32 form.
int ensure_split(efg_index* ix, hipStream_t st);
int gpu_merge_init(efg_index* ix);
int gpu_refine_iter(efg_index* ix);
int gpu_finalize_graph(efg_index* ix);

}  // namespace efg
