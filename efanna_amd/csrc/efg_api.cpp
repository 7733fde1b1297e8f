// C ABI of libefg.so: index lifetime, parameters, search entry points and the
// reference's file formats.  Kernels live in k_*.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <fstream>
#include <string>
#include <vector>

#include "efg.h"
#include "efg_internal.hpp"
#include "efg_mlp.hpp"

namespace efg {

thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return EFG_OK;
  return fail(EFG_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
struct MemCache {
  std::mutex mu;
  // per device: free blocks by size class, and blocks released since the
  // device was last synchronised
  std::map<int, std::multimap<size_t, void*>> ready;
  std::map<int, std::vector<std::pair<size_t, void*>>> pending;
};
MemCache& mem_cache() {
  static MemCache* c = new MemCache();  // never destroyed: blocks live until the process ends
  return *c;
}
// size classes: multiples of 256 B up to 1 MiB, then eight classes per power of two
size_t size_class(size_t b) {
  if (b <= (size_t(1) << 20)) return (b + 255) & ~size_t(255);
  int e = 63 - __builtin_clzll(b);
  size_t step = size_t(1) << (e - 3);
  return (b + step - 1) & ~(step - 1);
}
}  // namespace

void* dev_alloc(size_t bytes) {
  const size_t cls = size_class(bytes);
  int dev = 0;
  hipGetDevice(&dev);
  MemCache& c = mem_cache();
  std::lock_guard<std::mutex> g(c.mu);
  auto& rd = c.ready[dev];
  auto& pd = c.pending[dev];
  for (int round = 0; round < 2; ++round) {
    auto it = rd.find(cls);
    if (it != rd.end()) {
      void* p = it->second;
      rd.erase(it);
      return p;
    }
    bool fits = false;
    for (auto& b : pd) fits = fits || b.first == cls;
    if (!fits || round) break;
    // released blocks of this class exist: once the device has finished what
    // was submitted before their release, they are free to reuse
    if (hipDeviceSynchronize() != hipSuccess) return nullptr;
    for (auto& b : pd) rd.emplace(b.first, b.second);
    pd.clear();
  }
  void* p = nullptr;
  if (hipMalloc(&p, cls) == hipSuccess) return p;
  // out of memory: give every cached block back and try again
  (void)hipGetLastError();
  if (hipDeviceSynchronize() != hipSuccess) return nullptr;
  for (auto& b : pd) hipFree(b.second);
  pd.clear();
  for (auto& b : rd) hipFree(b.second);
  rd.clear();
  if (hipMalloc(&p, cls) == hipSuccess) return p;
  (void)hipGetLastError();
  return nullptr;
}

void dev_release(void* p, size_t bytes) {
  if (!p) return;
  int dev = 0;
  hipGetDevice(&dev);
  MemCache& c = mem_cache();
  std::lock_guard<std::mutex> g(c.mu);
  c.pending[dev].emplace_back(size_class(bytes), p);
}

namespace {

// Host copy of the trees: flat node arrays.
struct HostTrees {
  int T = 0;
  std::vector<int32_t> dim, left, right, roots;
  std::vector<float> val;
  std::vector<uint32_t> start, end, leaf;
  std::vector<int8_t> leaf_dim;
};

int upload_trees(efg_index* ix, const HostTrees& h) {
  const size_t M = h.dim.size();
  std::vector<int4> n4(M);
  std::vector<int32_t> par(M, -1);
  std::vector<uint2> rg(M);
  for (size_t i = 0; i < M; ++i) {
    int32_t vb;
    std::memcpy(&vb, &h.val[i], 4);
    n4[i] = make_int4(h.dim[i], vb, h.left[i], h.right[i]);
    rg[i] = make_uint2(h.start[i], h.end[i]);
  }
  for (size_t i = 0; i < M; ++i) {
    if (h.left[i] >= 0) par[h.left[i]] = int32_t(i);
    if (h.right[i] >= 0) par[h.right[i]] = int32_t(i);
  }
  if (!ix->node4.alloc(M) || !ix->parent.alloc(M) || !ix->range.alloc(M) || !ix->roots.alloc(h.T) ||
      !ix->leaf.alloc(size_t(h.T) * ix->n))
    return fail(EFG_ERR_MEMORY, "out of device memory (trees)");
  hipStream_t st = ix->stream;
  int rc = 0;
  rc |= check_hip(hipMemcpyAsync(ix->node4.p, n4.data(), M * sizeof(int4), hipMemcpyHostToDevice, st), "copy");
  rc |= check_hip(hipMemcpyAsync(ix->parent.p, par.data(), M * 4, hipMemcpyHostToDevice, st), "copy");
  rc |= check_hip(hipMemcpyAsync(ix->range.p, rg.data(), M * sizeof(uint2), hipMemcpyHostToDevice, st), "copy");
  rc |= check_hip(hipMemcpyAsync(ix->roots.p, h.roots.data(), h.T * 4, hipMemcpyHostToDevice, st), "copy");
  rc |= check_hip(hipMemcpyAsync(ix->leaf.p, h.leaf.data(), h.leaf.size() * 4, hipMemcpyHostToDevice, st), "copy");
  rc |= check_hip(hipStreamSynchronize(st), "sync");
  if (rc) return EFG_ERR_DEVICE;
  ix->T = h.T;
  ix->M = M;
  ix->snode_valid = false;
  {  // tree 0's leaf list holds every row once (the join's task order)
    std::vector<uint8_t> seen(ix->n, 0);
    bool perm = true;
    for (size_t i = 0; i < ix->n && perm; ++i) {
      const uint32_t x = h.leaf[i];
      perm = x < ix->n && !seen[x];
      if (perm) seen[x] = 1;
    }
    ix->leaf0_perm = perm;
  }
  ix->leaf_dim = h.leaf_dim;
  if (ix->leaf_dim.size() != M) ix->leaf_dim.assign(M, -1);
  ix->leaf_dim_lazy = false;
  return EFG_OK;
}

int download_trees(const efg_index* ix, HostTrees& h) {
  const size_t M = ix->M;
  std::vector<int4> n4(M);
  std::vector<uint2> rg(M);
  h.T = ix->T;
  h.roots.resize(ix->T);
  h.leaf.resize(size_t(ix->T) * ix->n);
  int rc = 0;
  rc |= check_hip(hipMemcpy(n4.data(), ix->node4.p, M * sizeof(int4), hipMemcpyDeviceToHost), "copy");
  rc |= check_hip(hipMemcpy(rg.data(), ix->range.p, M * sizeof(uint2), hipMemcpyDeviceToHost), "copy");
  rc |= check_hip(hipMemcpy(h.roots.data(), ix->roots.p, ix->T * 4, hipMemcpyDeviceToHost), "copy");
  rc |= check_hip(hipMemcpy(h.leaf.data(), ix->leaf.p, h.leaf.size() * 4, hipMemcpyDeviceToHost), "copy");
  if (rc) return EFG_ERR_DEVICE;
  h.dim.resize(M);
  h.val.resize(M);
  h.left.resize(M);
  h.right.resize(M);
  h.start.resize(M);
  h.end.resize(M);
  for (size_t i = 0; i < M; ++i) {
    h.dim[i] = n4[i].x;
    std::memcpy(&h.val[i], &n4[i].y, 4);
    h.left[i] = n4[i].z;
    h.right[i] = n4[i].w;
    h.start[i] = rg[i].x;
    h.end[i] = rg[i].y;
  }
  if (ix->leaf_dim_lazy) {
    // DivDim written for leaves (kdtreeub_index.hpp:1249, 1142-1143): leaves
    // the DFS phase made (and roots) carry -1, leaves left behind by the BFS
    // phase carry 0
    h.leaf_dim.assign(M, -1);
    const int Lb = ix->tree_lb;
    for (size_t i = size_t(ix->T); i < M; ++i) {
      const int lv = ix->node_level[i];
      const bool dfs_made = Lb >= 0 && (lv > Lb || (lv == Lb && rg[i].y - rg[i].x > uint32_t(ix->tree_S)));
      h.leaf_dim[i] = dfs_made ? -1 : 0;
    }
  } else {
    h.leaf_dim = ix->leaf_dim;
  }
  return EFG_OK;
}

int upload_graph(efg_index* ix, const std::vector<uint64_t>& off, const std::vector<uint32_t>& ids) {
  if (off.size() != ix->n + 1) return fail(EFG_ERR_FORMAT, "graph row count differs from the data set");
  for (uint32_t x : ids)
    if (x >= ix->n) return fail(EFG_ERR_FORMAT, "graph holds an id outside the data set");
  int maxd = 0;
  for (size_t i = 0; i < ix->n; ++i) maxd = std::max<int>(maxd, int(off[i + 1] - off[i]));
  if (!ix->goff.alloc(off.size()) || !ix->gids.alloc(ids.size()))
    return fail(EFG_ERR_MEMORY, "out of device memory (graph)");
  int rc = 0;
  rc |= check_hip(hipMemcpyAsync(ix->goff.p, off.data(), off.size() * 8, hipMemcpyHostToDevice, ix->stream), "copy");
  if (!ids.empty())
    rc |= check_hip(hipMemcpyAsync(ix->gids.p, ids.data(), ids.size() * 4, hipMemcpyHostToDevice, ix->stream), "copy");
  rc |= check_hip(hipStreamSynchronize(ix->stream), "sync");
  if (rc) return EFG_ERR_DEVICE;
  ix->nnz = ids.size();
  ix->max_deg = maxd;
  bool regular = true;
  for (size_t i = 0; i <= ix->n && regular; ++i) regular = off[i] == uint64_t(i) * uint64_t(maxd);
  ix->reg_deg = regular ? maxd : 0;
  return EFG_OK;
}

int download_graph(const efg_index* ix, std::vector<uint64_t>& off, std::vector<uint32_t>& ids) {
  off.assign(ix->n + 1, 0);
  ids.assign(ix->nnz, 0);
  if (ix->nnz == 0) return EFG_OK;
  int rc = 0;
  rc |= check_hip(hipMemcpy(off.data(), ix->goff.p, off.size() * 8, hipMemcpyDeviceToHost), "copy");
  rc |= check_hip(hipMemcpy(ids.data(), ix->gids.p, ids.size() * 4, hipMemcpyDeviceToHost), "copy");
  return rc ? EFG_ERR_DEVICE : EFG_OK;
}

// ---- trees file (kdtreeub_index.hpp:197-242 / 139-196) -----------------------
// Per tree: int node_num, int node_size (sizeof(Node) = 48 in the reference),
// node_num post-order records {int DivDim, float DivVal, size_t Start,
// size_t End, Node* L, Node* R}; the pointer fields carry no meaning for the
// loader (it rebuilds links from Start/End adjacency, 672-700) and are written
// as child index + 1 (0 for leaves).
void write_tree_nodes(std::ofstream& out, const HostTrees& h, int32_t root, int& count) {
  // iterative post-order
  std::vector<std::pair<int32_t, int>> st{{root, 0}};
  while (!st.empty()) {
    auto& top = st.back();
    int32_t x = top.first;
    if (h.left[x] >= 0 && top.second == 0) {
      top.second = 1;
      st.push_back({h.left[x], 0});
      continue;
    }
    if (h.right[x] >= 0 && top.second == 1) {
      top.second = 2;
      st.push_back({h.right[x], 0});
      continue;
    }
    st.pop_back();
    bool leaf = h.left[x] < 0;
    int32_t dd = leaf ? int32_t(h.leaf_dim.empty() ? -1 : h.leaf_dim[x]) : h.dim[x];
    float dv = leaf ? 0.f : h.val[x];
    uint64_t s = h.start[x], e = h.end[x];
    uint64_t lp = leaf ? 0 : uint64_t(h.left[x]) + 1, rp = leaf ? 0 : uint64_t(h.right[x]) + 1;
    out.write(reinterpret_cast<const char*>(&dd), 4);
    out.write(reinterpret_cast<const char*>(&dv), 4);
    out.write(reinterpret_cast<const char*>(&s), 8);
    out.write(reinterpret_cast<const char*>(&e), 8);
    out.write(reinterpret_cast<const char*>(&lp), 8);
    out.write(reinterpret_cast<const char*>(&rp), 8);
    ++count;
  }
}

int write_trees(const efg_index* ix, std::ofstream& out, uint32_t K) {
  HostTrees h;
  if (int rc = download_trees(ix, h)) return rc;
  uint32_t tn = uint32_t(h.T);
  uint64_t num = ix->n, dim = ix->d;
  out.write(reinterpret_cast<const char*>(&K), 4);
  out.write(reinterpret_cast<const char*>(&tn), 4);
  out.write(reinterpret_cast<const char*>(&num), 8);
  out.write(reinterpret_cast<const char*>(&dim), 8);
  for (int t = 0; t < h.T; ++t) {
    std::streampos at = out.tellp();
    int32_t zero = 0, node_size = 48;
    out.write(reinterpret_cast<const char*>(&zero), 4);
    out.write(reinterpret_cast<const char*>(&node_size), 4);
    int cnt = 0;
    write_tree_nodes(out, h, h.roots[t], cnt);
    std::streampos endp = out.tellp();
    out.seekp(at);
    out.write(reinterpret_cast<const char*>(&cnt), 4);
    out.seekp(endp);
  }
  for (int t = 0; t < h.T; ++t) out.write(reinterpret_cast<const char*>(&h.leaf[size_t(t) * ix->n]), ix->n * 4);
  return out ? EFG_OK : fail(EFG_ERR_IO, "write failed");
}

int read_trees(efg_index* ix, std::ifstream& in, uint32_t* K_out) {
  uint32_t K = 0, tn = 0;
  uint64_t num = 0, dim = 0;
  in.read(reinterpret_cast<char*>(&K), 4);
  in.read(reinterpret_cast<char*>(&tn), 4);
  in.read(reinterpret_cast<char*>(&num), 8);
  in.read(reinterpret_cast<char*>(&dim), 8);
  if (!in) return fail(EFG_ERR_FORMAT, "truncated trees header");
  if (num != ix->n || dim != ix->d) return fail(EFG_ERR_FORMAT, "trees file was built for another data set");
  if (tn == 0 || tn > 4096) return fail(EFG_ERR_FORMAT, "bad tree count in trees file");
  HostTrees h;
  h.T = int(tn);
  for (uint32_t t = 0; t < tn; ++t) {
    int32_t node_num = 0, node_size = 0;
    in.read(reinterpret_cast<char*>(&node_num), 4);
    in.read(reinterpret_cast<char*>(&node_size), 4);
    if (!in || node_num <= 0 || uint64_t(node_num) > 4 * num + 8) return fail(EFG_ERR_FORMAT, "bad node count in trees file");
    const int32_t base = int32_t(h.dim.size());
    std::vector<int32_t> st;  // DepthFirstBuildTree (672-700)
    for (int32_t j = 0; j < node_num; ++j) {
      int32_t dd;
      float dv;
      uint64_t s, e, lp, rp;
      in.read(reinterpret_cast<char*>(&dd), 4);
      in.read(reinterpret_cast<char*>(&dv), 4);
      in.read(reinterpret_cast<char*>(&s), 8);
      in.read(reinterpret_cast<char*>(&e), 8);
      in.read(reinterpret_cast<char*>(&lp), 8);
      in.read(reinterpret_cast<char*>(&rp), 8);
      if (!in) return fail(EFG_ERR_FORMAT, "truncated trees file");
      if (s > e || e > num) return fail(EFG_ERR_FORMAT, "node range outside the data set");
      int32_t id = base + j;
      h.dim.push_back(dd);
      h.val.push_back(dv);
      h.start.push_back(uint32_t(s));
      h.end.push_back(uint32_t(e));
      h.left.push_back(-1);
      h.right.push_back(-1);
      h.leaf_dim.push_back(int8_t(dd < 0 ? -1 : 0));
      size_t rs = st.size();
      if (rs >= 2) {
        int32_t l1 = st[rs - 1], l2 = st[rs - 2];
        if (h.end[l1] == e && h.start[l2] == s) {
          h.right[id] = l1;
          h.left[id] = l2;
          st.pop_back();
          st.pop_back();
        }
      }
      st.push_back(id);
    }
    if (st.size() != 1) return fail(EFG_ERR_FORMAT, "Error constructing trees");
    h.roots.push_back(st[0]);
  }
  for (size_t i = 0; i < h.dim.size(); ++i) {
    if (h.left[i] >= 0) {
      if (h.dim[i] < 0 || uint64_t(h.dim[i]) >= dim) return fail(EFG_ERR_FORMAT, "split dimension out of range");
      h.leaf_dim[i] = 0;
    }
  }
  h.leaf.resize(size_t(tn) * num);
  in.read(reinterpret_cast<char*>(h.leaf.data()), h.leaf.size() * 4);
  if (!in) return fail(EFG_ERR_FORMAT, "truncated leaf lists");
  for (uint32_t x : h.leaf)
    if (x >= num) return fail(EFG_ERR_FORMAT, "leaf list holds an id outside the data set");
  if (K_out) *K_out = K;
  int rc = upload_trees(ix, h);
  if (rc) return rc;
  ix->s_trees = int(tn);  // SP.tree_num = tree_num (kdtreeub_index.hpp:151)
  return EFG_OK;
}

int sync_results_to_host(efg_index* ix, size_t nq, int k) {
  ix->res_ids.resize(nq * size_t(k));
  ix->res_d.resize(nq * size_t(k));
  int rc = check_hip(hipMemcpyAsync(ix->res_ids.data(), ix->ws.ids.p, nq * k * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
  rc |= check_hip(hipMemcpyAsync(ix->res_d.data(), ix->ws.dists.p, nq * k * 4, hipMemcpyDeviceToHost, ix->stream), "copy");
  rc |= check_hip(hipStreamSynchronize(ix->stream), "sync");
  ix->res_nq = nq;
  ix->res_k = k;
  return rc ? EFG_ERR_DEVICE : EFG_OK;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace
}  // namespace efg

using namespace efg;

extern "C" {

const char* efg_last_error(void) { return g_err.c_str(); }

void efg_default_build_params(efg_build_params* p, int trees) {
  // KDTreeUbIndexParams(rnn_used, tree_num_total, ml=4, epoches=4, check=25,
  // L=30, K=10, tree_num_build=0, S=10) + IndexParams::TNS = 10
  p->trees = trees;
  p->trees_build = 0;
  p->merge_level = 4;
  p->epochs = 4;
  p->check = 25;
  p->L = 30;
  p->K = 10;
  p->S = 10;
  p->tns = 10;
  p->seed = 1998;
}

int efg_create(const float* data, size_t n, size_t d, size_t row_stride, const efg_build_params* params,
               efg_index** out) {
  if (!out) return fail(EFG_ERR_ARG, "null output");
  *out = nullptr;
  if (!data || n == 0 || d == 0 || row_stride < d) return fail(EFG_ERR_ARG, "invalid data matrix");
  if (n >= (size_t(1) << 30)) return fail(EFG_ERR_ARG, "at most 2^30 - 1 rows");
  efg_index* ix = new efg_index();
  ix->n = n;
  ix->d = d;
  ix->D = (d + 7) & ~size_t(7);
  if (params) ix->bp = *params;
  else efg_default_build_params(&ix->bp, 4);
  if (ix->bp.trees_build <= 0) ix->bp.trees_build = ix->bp.trees;
  if (ix->bp.tns <= 0) ix->bp.tns = 10;
  ix->s_trees = ix->bp.trees;  // SP.tree_num = TreeNum (kdtreeub_index.hpp:59)
  if (int rc = check_hip(hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking), "stream")) {
    delete ix;
    return rc;
  }
  if (!ix->data.alloc(n * ix->D)) {
    efg_destroy(ix);
    return fail(EFG_ERR_MEMORY, "out of device memory (data)");
  }
  std::vector<float> pad;
  const size_t rows_per = std::max<size_t>(1, (size_t(64) << 20) / (ix->D * 4));
  for (size_t r0 = 0; r0 < n; r0 += rows_per) {
    size_t r1 = std::min(n, r0 + rows_per);
    pad.assign((r1 - r0) * ix->D, 0.f);
    for (size_t i = r0; i < r1; ++i) std::memcpy(&pad[(i - r0) * ix->D], data + i * row_stride, d * 4);
    int rc = check_hip(hipMemcpyAsync(ix->data.p + r0 * ix->D, pad.data(), pad.size() * 4, hipMemcpyHostToDevice, ix->stream), "copy");
    rc |= check_hip(hipStreamSynchronize(ix->stream), "sync");
    if (rc) {
      efg_destroy(ix);
      return EFG_ERR_DEVICE;
    }
  }
  *out = ix;
  return EFG_OK;
}

int efg_destroy(efg_index* ix) {
  if (!ix) return EFG_OK;
  if (ix->stream) hipStreamSynchronize(ix->stream);
  for (hipStream_t s : ix->bstreams) hipStreamSynchronize(s);
  for (hipEvent_t e : ix->prof_events) hipEventDestroy(e);
  if (ix->bfork) hipEventDestroy(ix->bfork);
  for (hipEvent_t e : ix->bjoin) hipEventDestroy(e);
  std::vector<hipStream_t> aux = ix->bstreams;
  hipStream_t st = ix->stream;
  delete ix;
  for (hipStream_t s : aux) hipStreamDestroy(s);
  if (st) hipStreamDestroy(st);
  return EFG_OK;
}

int efg_build_trees(efg_index* ix) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  double t0 = now_ms();
  int rc = gpu_build_trees(ix);
  ix->times[0] = now_ms() - t0;
  return rc;
}

int efg_build_stage(efg_index* ix, int stage) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  switch (stage) {
    case 0: return efg_build_trees(ix);
    case 1: return gpu_merge_init(ix);
    case 2: return gpu_refine_iter(ix);
    case 3: return gpu_finalize_graph(ix);
    default: return fail(EFG_ERR_ARG, "no such build stage");
  }
}

int efg_hash_set_codes(efg_index* ix, int tables, int codelen, int lenshift, int upbits, int radius,
                       const uint64_t* base_codes, const uint64_t* query_codes, size_t nq) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  return hash_set_codes(ix, tables, codelen, lenshift, upbits, radius, base_codes, query_codes, nq);
}

// Code files as LoadCode32/64 read them (hashing_index.hpp:206-321): file
// "<prefix>_<t>" for t = 1..tables holds int32 1, int32 code length, int32
// count, then count codes of 4 bytes (code length <= 32) or 8 bytes.  The
// code length given here picks the width, the file's length is used after.
static int read_codes(const char* prefix, int tables, int width_bits, int* file_len, std::vector<uint64_t>& out,
                      size_t* count) {
  out.clear();
  for (int t = 0; t < tables; ++t) {
    const std::string path = std::string(prefix) + "_" + std::to_string(t + 1);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return fail(EFG_ERR_IO, ("open file " + std::string(prefix) + " error").c_str());
    int32_t hdr[3];
    if (std::fread(hdr, 4, 3, f) != 3) {
      std::fclose(f);
      return fail(EFG_ERR_IO, ("Codefile  " + std::to_string(t) + " error!").c_str());
    }
    if (hdr[0] != 1 || hdr[2] < 0 || (t > 0 && size_t(hdr[2]) != *count)) {
      std::fclose(f);
      return fail(EFG_ERR_IO, ("Codefile  " + std::to_string(t) + " error!").c_str());
    }
    *file_len = hdr[1];
    *count = size_t(hdr[2]);
    const size_t m = size_t(hdr[2]);
    const size_t w = width_bits <= 32 ? 4 : 8;
    std::vector<unsigned char> raw(m * w);
    const size_t got = m ? std::fread(raw.data(), w, m, f) : 0;
    std::fclose(f);
    if (got != m) return fail(EFG_ERR_IO, ("Codefile  " + std::to_string(t) + " error!").c_str());
    for (size_t i = 0; i < m; ++i) {
      uint64_t v = 0;
      if (w == 4) {
        uint32_t u;
        std::memcpy(&u, &raw[i * 4], 4);
        v = u;
      } else {
        std::memcpy(&v, &raw[i * 8], 8);
      }
      out.push_back(v);
    }
  }
  return EFG_OK;
}

int efg_hash_load_codes(efg_index* ix, int tables, int codelen, int lenshift, int upbits, int radius,
                        const char* base_prefix, const char* query_prefix) {
  if (!ix || !base_prefix || !query_prefix) return fail(EFG_ERR_ARG, "null argument");
  if (tables < 1) return fail(EFG_ERR_ARG, "Total hash table num error!");
  if (codelen > 64 || codelen < 1) return fail(EFG_ERR_ARG, "code length not supported yet!");
  std::vector<uint64_t> base, query;
  int len_b = codelen, len_q = codelen;
  size_t nb = 0, nq = 0;
  if (int rc = read_codes(base_prefix, tables, codelen, &len_b, base, &nb)) return rc;
  if (nb != ix->n) return fail(EFG_ERR_IO, "base code count differs from the data rows");
  // the query file's width follows the base file's code length, as the
  // reference reads it after the base file has set codelength
  if (int rc = read_codes(query_prefix, tables, len_b, &len_q, query, &nq)) return rc;
  return hash_set_codes(ix, tables, len_q, lenshift, upbits, radius, base.data(), query.data(), nq);
}

int efg_build_index(efg_index* ix) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (ix->hashing) return hash_build(ix);  // hashing index: the tables (hashing_index.hpp:874-889)
  const efg_build_params& p = ix->bp;
  if (p.K <= 0 || p.S <= 0 || p.L <= 0 || p.check <= 0 || p.trees <= 0 || p.epochs < 0 || p.merge_level < 0)
    return fail(EFG_ERR_ARG, "invalid build parameters");
  if (p.K > p.L) return fail(EFG_ERR_ARG, "K must not exceed L");
  ix->join_pairs.clear();
  double t0 = now_ms();
  if (int rc = gpu_build_trees(ix)) return rc;
  double t1 = now_ms();
  if (int rc = gpu_merge_init(ix)) return rc;
  double t2 = now_ms();
  ix->times[0] = t1 - t0;
  ix->times[1] = t2 - t1;
  ix->times[2] = ix->times[3] = 0;
  for (int e = 0; e < p.epochs; ++e)
    if (int rc = gpu_refine_iter(ix)) return rc;
  return gpu_finalize_graph(ix);
}

int efg_set_search_params(efg_index* ix, int epochs, int init_num, int extend_to, int search_trees, int search_lv,
                          int search_method) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  ix->s_epochs = epochs;
  ix->s_init = init_num;
  ix->s_extend = extend_to > init_num ? init_num : extend_to;
  ix->s_lv = search_lv;
  ix->s_trees = search_trees;
  ix->s_method = search_method;
  ix->s_set = true;
  return EFG_OK;
}

int efg_set_search_precision(efg_index* ix, int precision) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (precision != 0 && precision != 1) return fail(EFG_ERR_ARG, "precision must be 0 (FP32) or 1 (BF16)");
  ix->s_precision = precision;
  return EFG_OK;
}

int efg_set_build_precision(efg_index* ix, int precision) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (precision < 0 || precision > 2) return fail(EFG_ERR_ARG, "precision must be 0 (FP32), 1 (BF16) or 2 (FP16)");
  ix->b_precision = precision;
  return EFG_OK;
}

int efg_set_join_staging_margin(efg_index* ix, double margin) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (!(margin > 0.0)) return fail(EFG_ERR_ARG, "staging margin must be positive");
  ix->b_stage_margin = margin;
  return EFG_OK;
}

int efg_set_join_routing(efg_index* ix, int direct) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (direct != 0 && direct != 1) return fail(EFG_ERR_ARG, "join routing must be 0 (staged) or 1 (direct)");
  ix->b_join_direct = direct != 0;
  return EFG_OK;
}

int efg_knn_search_batch_device(efg_index* ix, const float* d_queries, size_t nq, size_t q_stride, const int* pool,
                                const int* extend, const int* k, int out_stride, int32_t* d_ids, float* d_dists,
                                int tune, void* stream) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (nq && (!d_queries || !pool || !k || !d_ids)) return fail(EFG_ERR_ARG, "null buffer");
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ix->stream;
  return efg::search_batch(ix, d_queries, nq, q_stride, pool, extend, k, out_stride, d_ids, d_dists, tune != 0, st);
}

int efg_batch_plan_create(efg_index* ix, const float* d_queries, size_t nq, size_t q_stride, const int* pool,
                          const int* extend, const int* k, int out_stride, int32_t* d_ids, float* d_dists, int tune,
                          void* stream, efg_batch_plan** plan) {
  if (!ix || !plan) return fail(EFG_ERR_ARG, "null argument");
  *plan = nullptr;
  if (!d_queries || !pool || !k || !d_ids) return fail(EFG_ERR_ARG, "null buffer");
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ix->stream;
  return efg::batch_plan_create(ix, d_queries, nq, q_stride, pool, extend, k, out_stride, d_ids, d_dists, tune != 0,
                                st, plan);
}

int efg_batch_plan_run(efg_batch_plan* plan, void* stream) {
  if (!plan) return fail(EFG_ERR_ARG, "null plan");
  return efg::batch_plan_run(plan, static_cast<hipStream_t>(stream));
}

int efg_batch_plan_destroy(efg_batch_plan* plan) {
  efg::batch_plan_destroy(plan);
  return EFG_OK;
}

int efg_get_batch_shape(const efg_index* ix, int* split, int* t_small, int* t_large, int* share_pct) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (split) *split = ix->b_split;
  if (t_small) *t_small = ix->b_tsmall;
  if (t_large) *t_large = ix->b_tlarge;
  if (share_pct) *share_pct = int(ix->b_share * 100.f + 0.5f);
  return EFG_OK;
}

int efg_set_batch_form(efg_index* ix, int one_pass) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (one_pass != 0 && one_pass != 1) return fail(EFG_ERR_ARG, "batch form must be 0 or 1");
  ix->b_mode = one_pass;
  return EFG_OK;
}

int efg_knn_search_batch(efg_index* ix, const float* queries, size_t nq, size_t q_stride, const int* pool,
                         const int* extend, const int* k, int out_stride, int32_t* ids, float* dists, int tune) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (nq == 0) return EFG_OK;
  if (!queries || !pool || !k || !ids || q_stride < ix->d || out_stride <= 0) return fail(EFG_ERR_ARG, "invalid arguments");
  efg::DevBuf<float> dq, dd;
  efg::DevBuf<int32_t> di;
  const size_t D = ix->D, m = nq * size_t(out_stride);
  if (!dq.alloc(nq * D) || !di.alloc(m) || !dd.alloc(m)) return fail(EFG_ERR_MEMORY, "out of device memory");
  hipStream_t st = ix->stream;
  hipMemsetAsync(dq.p, 0, nq * D * 4, st);
  if (int rc = check_hip(hipMemcpy2DAsync(dq.p, D * 4, queries, q_stride * 4, ix->d * 4, nq, hipMemcpyHostToDevice, st),
                         "copy"))
    return rc;
  if (int rc = efg::search_batch(ix, dq.p, nq, D, pool, extend, k, out_stride, di.p, dd.p, tune != 0, st)) return rc;
  hipMemcpyAsync(ids, di.p, m * 4, hipMemcpyDeviceToHost, st);
  if (dists) hipMemcpyAsync(dists, dd.p, m * 4, hipMemcpyDeviceToHost, st);
  return check_hip(hipStreamSynchronize(st), "batch search");
}

int efg_set_search_threads(efg_index* ix, int threads) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (threads != 0 && threads != 64 && threads != 128 && threads != 256)
    return fail(EFG_ERR_ARG, "threads per query must be 0, 64, 128 or 256");
  ix->s_threads = threads;
  return EFG_OK;
}

int efg_autotune_search(efg_index* ix, int k, const float* queries, size_t nq, size_t q_stride, int* threads_out) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (!queries || nq == 0) return fail(EFG_ERR_ARG, "no queries");
  if (q_stride < ix->d) return fail(EFG_ERR_ARG, "query stride below the dimension");
  efg::DevBuf<float> dq;
  const size_t qs = ix->D;
  if (!dq.alloc(nq * qs)) return fail(EFG_ERR_MEMORY, "out of device memory");
  hipStream_t st = ix->stream;
  hipMemsetAsync(dq.p, 0, nq * qs * 4, st);
  if (int rc = efg::check_hip(hipMemcpy2DAsync(dq.p, qs * 4, queries, q_stride * 4, ix->d * 4, nq,
                                               hipMemcpyHostToDevice, st),
                              "copy"))
    return rc;
  int chosen[2] = {0, 0};
  if (int rc = efg::search_autotune(ix, k, dq.p, nq, qs, chosen, st)) return rc;
  if (threads_out) {
    threads_out[0] = chosen[0];
    threads_out[1] = chosen[1];
  }
  return efg::check_hip(hipStreamSynchronize(st), "autotune");
}

int efg_mlp_param_count(void) { return efg::kMlpParams; }

int efg_mlp_init(float* params, uint64_t seed) {
  if (!params) return fail(EFG_ERR_ARG, "null parameters");
  efg::mlp_init(params, seed);
  return EFG_OK;
}

int efg_mlp_logits(const float* params, const float* norm, const float* feat, size_t n, float* z) {
  if (!params || !norm || (n && (!feat || !z))) return fail(EFG_ERR_ARG, "null buffer");
  return efg::mlp_logits(params, norm, feat, n, z, nullptr);
}

int efg_mlp_gradient(const float* params, const float* norm, const float* feat, const uint8_t* label, size_t n,
                     float* grad, float* loss) {
  if (!params || !norm || !feat || !label || !grad) return fail(EFG_ERR_ARG, "null buffer");
  return efg::mlp_gradient(params, norm, feat, label, n, grad, loss, nullptr);
}

int efg_mlp_train(const float* feat, const uint8_t* label, size_t n, const float* norm, float* params, int epochs,
                  int batch, float lr, uint64_t seed, float* loss_hist) {
  if (!feat || !label || !norm || !params) return fail(EFG_ERR_ARG, "null buffer");
  return efg::mlp_train(feat, label, n, norm, params, epochs, batch, lr, seed, loss_hist, nullptr);
}

int efg_set_probe_params(efg_index* ix, int pool, int extend_to, int epochs) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (pool < 0 || extend_to < 0 || epochs < 0) return fail(EFG_ERR_ARG, "invalid probe parameters");
  ix->lp_pool = pool;
  ix->lp_extend = extend_to;
  ix->lp_epochs = epochs;
  return EFG_OK;
}

int efg_search_traces(efg_index* ix, int k, const float* queries, size_t nq, size_t q_stride,
                      const int32_t* plain_ids, float* feat, uint8_t* label) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (nq && (!queries || !plain_ids || !feat || !label)) return fail(EFG_ERR_ARG, "null buffer");
  if (k <= 0) return fail(EFG_ERR_ARG, "k must be positive");
  const int ep = ix->lp_epochs > 0 ? ix->lp_epochs : ix->s_epochs;
  const size_t C = size_t(ep) + 1;
  efg::DevBuf<float> dq, df;
  efg::DevBuf<int32_t> dp;
  efg::DevBuf<uint8_t> dl;
  const size_t qs = ix->d;
  if (!dq.alloc(nq * qs + 1) || !dp.alloc(nq * size_t(k) + 1) || !df.alloc(nq * C * efg::kFeat + 1) ||
      !dl.alloc(nq * C + 1))
    return fail(EFG_ERR_MEMORY, "out of device memory (traces)");
  hipStream_t st = ix->stream;
  if (nq) {
    if (int rc = efg::check_hip(hipMemcpy2DAsync(dq.p, qs * 4, queries, q_stride * 4, ix->d * 4, nq,
                                                 hipMemcpyHostToDevice, st),
                                "copy"))
      return rc;
    hipMemcpyAsync(dp.p, plain_ids, nq * size_t(k) * 4, hipMemcpyHostToDevice, st);
  }
  if (int rc = efg::search_traces(ix, k, dq.p, nq, qs, dp.p, df.p, dl.p, st)) return rc;
  if (nq) {
    hipMemcpyAsync(feat, df.p, nq * C * efg::kFeat * 4, hipMemcpyDeviceToHost, st);
    hipMemcpyAsync(label, dl.p, nq * C, hipMemcpyDeviceToHost, st);
  }
  return efg::check_hip(hipStreamSynchronize(st), "traces");
}

int efg_set_stopping_network(efg_index* ix, const float* params, const float* norm, float threshold) {
  if (!ix || !params || !norm) return fail(EFG_ERR_ARG, "null argument");
  if (!(threshold > 0.f && threshold < 1.f)) return fail(EFG_ERR_ARG, "threshold must lie in (0, 1)");
  using namespace efg;
  // packed columns (efg_mlp.hpp), then W3 | b3 as floats, all BF16-rounded
  std::vector<uint32_t> pack(kPackLen + kHid + 1);
  auto bf = [&](int i) { return uint32_t(float_to_bf16(params[i])); };
  for (int j = 0; j < kHid; ++j) {
    for (int i = 0; i < kFeat / 2; ++i)
      pack[i * kHid + j] = bf(kOffW1 + (2 * i) * kHid + j) | (bf(kOffW1 + (2 * i + 1) * kHid + j) << 16);
    for (int i = 0; i < kHid / 2; ++i)
      pack[(kFeat / 2 + i) * kHid + j] = bf(kOffW2 + (2 * i) * kHid + j) | (bf(kOffW2 + (2 * i + 1) * kHid + j) << 16);
    pack[(kPackWords - 1) * kHid + j] = bf(kOffB1 + j) | (bf(kOffB2 + j) << 16);
  }
  for (int i = 0; i <= kHid; ++i) {
    const float v = bf16_to_float(float_to_bf16(params[kOffW3 + i]));
    std::memcpy(&pack[kPackLen + i], &v, 4);
  }
  if (!ix->net_pack.ensure(pack.size()) || !ix->net_norm.ensure(2 * kFeat))
    return fail(EFG_ERR_MEMORY, "out of device memory");
  hipStream_t st = ix->stream;
  hipMemcpyAsync(ix->net_pack.p, pack.data(), pack.size() * 4, hipMemcpyHostToDevice, st);
  hipMemcpyAsync(ix->net_norm.p, norm, 2 * kFeat * 4, hipMemcpyHostToDevice, st);
  if (int rc = check_hip(hipStreamSynchronize(st), "copy")) return rc;
  // probability >= threshold  <=>  logit >= log(threshold / (1 - threshold))
  ix->net_zthr = float(std::log(double(threshold) / (1.0 - double(threshold))));
  ix->net_set = true;
  return EFG_OK;
}

int efg_set_search_mode(efg_index* ix, int mode) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (mode < 0 || mode > 2)
    return fail(EFG_ERR_ARG, "search mode must be 0 (plain), 1 (learned stopping, restart) or 2 (learned stopping, "
                             "continuation)");
  if (mode != 0 && !ix->net_set) return fail(EFG_ERR_STATE, "no stopping network set");
  ix->s_mode = mode;
  return EFG_OK;
}

int efg_get_learned_stats(efg_index* ix, uint64_t* last_fallbacks, uint64_t* total_fallbacks,
                          uint64_t* total_queries) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  int h[4] = {0, 0, 0, 0};
  if (ix->fb_count.p) {
    hipMemcpyAsync(h, ix->fb_count.p, sizeof(h), hipMemcpyDeviceToHost, ix->stream);
    if (int rc = efg::check_hip(hipStreamSynchronize(ix->stream), "copy")) return rc;
  }
  if (last_fallbacks) *last_fallbacks = uint64_t(h[3]);
  if (total_fallbacks) *total_fallbacks = uint64_t(h[1]);
  if (total_queries) *total_queries = uint64_t(h[2]);
  return EFG_OK;
}

int efg_get_learned_restarts(efg_index* ix, uint64_t* last_restarts) {
  if (!ix || !last_restarts) return fail(EFG_ERR_ARG, "null argument");
  int h = 0;
  if (ix->fb_count.p) {
    hipMemcpyAsync(&h, ix->fb_count.p, sizeof(h), hipMemcpyDeviceToHost, ix->stream);
    if (int rc = efg::check_hip(hipStreamSynchronize(ix->stream), "copy")) return rc;
  }
  *last_restarts = uint64_t(h);
  return EFG_OK;
}

int efg_get_search_params(const efg_index* ix, int* epochs, int* init_num, int* extend_to, int* search_trees,
                          int* search_method) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (epochs) *epochs = ix->s_epochs;
  if (init_num) *init_num = ix->s_init;
  if (extend_to) *extend_to = ix->s_extend;
  if (search_trees) *search_trees = ix->s_trees;
  if (search_method) *search_method = ix->s_method;
  return EFG_OK;
}

int efg_knn_search_device(efg_index* ix, int k, const float* d_queries, size_t nq, size_t q_stride, int32_t* d_ids,
                          float* d_dists, uint64_t* d_ndist, void* stream) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ix->stream;
  return search_device(ix, k, d_queries, nq, q_stride, d_ids, d_dists, d_ndist, st);
}

int efg_knn_search_device_noalloc(efg_index* ix, int k, const float* d_queries, size_t nq, size_t q_stride,
                                  int32_t* d_ids, float* d_dists, void* stream) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ix->stream;
  return search_device_noalloc(ix, k, d_queries, nq, q_stride, d_ids, d_dists, st);
}

int efg_knn_search(efg_index* ix, int k, const float* queries, size_t nq, size_t q_stride) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (nq && (!queries || q_stride < ix->d)) return fail(EFG_ERR_ARG, "invalid query matrix");
  if (k <= 0) return fail(EFG_ERR_ARG, "k must be positive");
  efg::Workspace& w = ix->ws;
  const size_t D = ix->D;
  if (!w.qbuf.ensure(std::max<size_t>(1, nq * D)) || !w.ids.ensure(std::max<size_t>(1, nq * k)) ||
      !w.dists.ensure(std::max<size_t>(1, nq * k)))
    return fail(EFG_ERR_MEMORY, "out of device memory (queries)");
  std::vector<float> pad(nq * D, 0.f);
  for (size_t i = 0; i < nq; ++i) std::memcpy(&pad[i * D], queries + i * q_stride, ix->d * 4);
  if (nq) {
    if (int rc = check_hip(hipMemcpyAsync(w.qbuf.p, pad.data(), pad.size() * 4, hipMemcpyHostToDevice, ix->stream), "copy"))
      return rc;
  }
  int rc = search_device(ix, k, w.qbuf.p, nq, D, w.ids.p, w.dists.p, nullptr, ix->stream);
  if (rc) return rc;
  return sync_results_to_host(ix, nq, k);
}

int efg_get_results(const efg_index* ix, size_t* nq, int* k, int32_t* ids, float* dists) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (nq) *nq = ix->res_nq;
  if (k) *k = ix->res_k;
  if (ids) std::memcpy(ids, ix->res_ids.data(), ix->res_ids.size() * 4);
  if (dists) std::memcpy(dists, ix->res_d.data(), ix->res_d.size() * 4);
  return EFG_OK;
}

int efg_graph_size(const efg_index* ix, size_t* n) {
  if (!ix || !n) return fail(EFG_ERR_ARG, "null argument");
  *n = ix->nnz ? ix->n : 0;
  return EFG_OK;
}

int efg_graph_row(const efg_index* ix, size_t row, uint32_t* out, size_t cap, size_t* len) {
  if (!ix || !len) return fail(EFG_ERR_ARG, "null argument");
  *len = 0;
  if (ix->nnz == 0 || row >= ix->n) return EFG_OK;  // empty row (base_index.hpp:495-501)
  uint64_t o[2];
  if (int rc = check_hip(hipMemcpy(o, ix->goff.p + row, 16, hipMemcpyDeviceToHost), "copy")) return rc;
  size_t m = size_t(o[1] - o[0]);
  *len = m;
  if (out && cap) {
    size_t c = std::min(cap, m);
    if (c) {
      if (int rc = check_hip(hipMemcpy(out, ix->gids.p + o[0], c * 4, hipMemcpyDeviceToHost), "copy")) return rc;
    }
  }
  return EFG_OK;
}

int efg_save_graph(const efg_index* ix, const char* path) {
  if (!ix || !path) return fail(EFG_ERR_ARG, "null argument");
  std::vector<uint64_t> off;
  std::vector<uint32_t> ids;
  if (int rc = download_graph(ix, off, ids)) return rc;
  std::ofstream out(path, std::ios::binary);
  if (!out.is_open()) return fail(EFG_ERR_IO, "open file error");
  const size_t rows = ix->nnz ? ix->n : 0;
  for (size_t i = 0; i < rows; ++i) {
    uint32_t k = uint32_t(off[i + 1] - off[i]);
    out.write(reinterpret_cast<const char*>(&k), 4);
    out.write(reinterpret_cast<const char*>(&ids[off[i]]), size_t(k) * 4);
  }
  return out ? EFG_OK : fail(EFG_ERR_IO, "write failed");
}

int efg_load_graph(efg_index* ix, const char* path) {
  if (!ix || !path) return fail(EFG_ERR_ARG, "null argument");
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) return fail(EFG_ERR_IO, "open file error");
  in.seekg(0, std::ios::end);
  size_t fsize = size_t(in.tellg());
  in.seekg(0, std::ios::beg);
  int32_t k0 = 0;
  in.read(reinterpret_cast<char*>(&k0), 4);
  if (!in || k0 < 0) return fail(EFG_ERR_FORMAT, "bad graph header");
  size_t N = fsize / (size_t(k0) + 1) / 4;  // kdtreeub_index.hpp:253
  if (N != ix->n) return fail(EFG_ERR_FORMAT, "graph row count differs from the data set");
  in.seekg(0, std::ios::beg);
  std::vector<uint64_t> off(N + 1, 0);
  std::vector<uint32_t> ids;
  ids.reserve(N * size_t(k0));
  for (size_t i = 0; i < N; ++i) {
    uint32_t k = 0;
    in.read(reinterpret_cast<char*>(&k), 4);
    if (!in || k > (1u << 20)) return fail(EFG_ERR_FORMAT, "truncated graph file");
    size_t at = ids.size();
    ids.resize(at + k);
    in.read(reinterpret_cast<char*>(&ids[at]), size_t(k) * 4);
    if (!in) return fail(EFG_ERR_FORMAT, "truncated graph file");
    off[i + 1] = ids.size();
  }
  return upload_graph(ix, off, ids);
}

int efg_save_trees(const efg_index* ix, const char* path) {
  if (!ix || !path) return fail(EFG_ERR_ARG, "null argument");
  if (ix->T == 0) return fail(EFG_ERR_STATE, "no trees to save");
  std::ofstream out(path, std::ios::binary);
  if (!out.is_open()) return fail(EFG_ERR_IO, "open file error");
  return write_trees(ix, out, uint32_t(ix->bp.K));
}

int efg_load_trees(efg_index* ix, const char* path) {
  if (!ix || !path) return fail(EFG_ERR_ARG, "null argument");
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) return fail(EFG_ERR_IO, "open file error");
  return read_trees(ix, in, nullptr);
}

// saveIndex (kdtreeub_index.hpp:771-825): trees file followed by n rows of K
// int ids.  The reference writes its merge heaps here, which are cleared after
// a build (1424) so it exits; the graph (first K pool ids) is written instead.
int efg_save_index(const efg_index* ix, const char* path) {
  if (!ix || !path) return fail(EFG_ERR_ARG, "null argument");
  if (ix->T == 0 || ix->nnz == 0) return fail(EFG_ERR_STATE, "index needs trees and a graph");
  std::vector<uint64_t> off;
  std::vector<uint32_t> ids;
  if (int rc = download_graph(ix, off, ids)) return rc;
  uint32_t K = uint32_t(off[1] - off[0]);
  for (size_t i = 0; i < ix->n; ++i)
    if (off[i + 1] - off[i] != K) return fail(EFG_ERR_STATE, "index file needs a graph of uniform degree");
  std::ofstream out(path, std::ios::binary);
  if (!out.is_open()) return fail(EFG_ERR_IO, "open file error");
  if (int rc = write_trees(ix, out, K)) return rc;
  out.write(reinterpret_cast<const char*>(ids.data()), ids.size() * 4);
  return out ? EFG_OK : fail(EFG_ERR_IO, "write failed");
}

// loadIndex (kdtreeub_index.hpp:701-770): the K ids per row become the search
// graph (the reference fills heaps the search never reads).
int efg_load_index(efg_index* ix, const char* path) {
  if (!ix || !path) return fail(EFG_ERR_ARG, "null argument");
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open()) return fail(EFG_ERR_IO, "open file error");
  uint32_t K = 0;
  if (int rc = read_trees(ix, in, &K)) return rc;
  std::vector<uint64_t> off(ix->n + 1);
  std::vector<uint32_t> ids(ix->n * size_t(K));
  in.read(reinterpret_cast<char*>(ids.data()), ids.size() * 4);
  if (!in) return fail(EFG_ERR_FORMAT, "truncated graph section of the index file");
  for (size_t i = 0; i <= ix->n; ++i) off[i] = i * K;
  return upload_graph(ix, off, ids);
}

int efg_save_image(efg_index* ix, const char* path) {
  if (!ix || !path) return fail(EFG_ERR_ARG, "null argument");
  return save_image(ix, path);
}

int efg_load_image(efg_index* ix, const char* path) {
  if (!ix || !path) return fail(EFG_ERR_ARG, "null argument");
  return load_image(ix, path);
}

// saveResults (base_index.hpp:136-151): per query int k, then k int ids.
// Rows report their valid entries only (method 1 can find fewer than k).
int efg_save_results(const efg_index* ix, const char* path) {
  if (!ix || !path) return fail(EFG_ERR_ARG, "null argument");
  std::ofstream out(path, std::ios::binary);
  if (!out.is_open()) return fail(EFG_ERR_IO, "open file error");
  for (size_t q = 0; q < ix->res_nq; ++q) {
    const int32_t* r = &ix->res_ids[q * ix->res_k];
    int32_t m = 0;
    while (m < ix->res_k && r[m] >= 0) ++m;
    out.write(reinterpret_cast<const char*>(&m), 4);
    out.write(reinterpret_cast<const char*>(r), size_t(m) * 4);
  }
  return out ? EFG_OK : fail(EFG_ERR_IO, "write failed");
}

int efg_exact_knn_device(const float* d_base, size_t n, size_t d, size_t base_stride, const float* d_queries,
                         size_t nq, size_t q_stride, int k, int exclude_self, int32_t* d_ids, float* d_dists,
                         void* stream) {
  if (nq == 0) return EFG_OK;
  if (!d_base || !d_queries || !d_ids || q_stride < d) return fail(EFG_ERR_ARG, "invalid arguments");
  if (nq >= (size_t(1) << 31) || n >= (size_t(1) << 31)) return fail(EFG_ERR_ARG, "too large");
  std::string err;
  int rc = exact_knn(d_base, n, int(d), base_stride, d_queries, int(nq), q_stride, k, exclude_self, d_ids, d_dists,
                     static_cast<hipStream_t>(stream), &err);
  if (rc) return fail(rc, err);
  return EFG_OK;
}

int efg_exact_knn_device_prec(const float* d_base, size_t n, size_t d, size_t base_stride, const float* d_queries,
                              size_t nq, size_t q_stride, int k, int exclude_self, int precision, int32_t* d_ids,
                              float* d_dists, void* stream) {
  if (precision == 0)
    return efg_exact_knn_device(d_base, n, d, base_stride, d_queries, nq, q_stride, k, exclude_self, d_ids, d_dists,
                                stream);
  if (precision != 1 && precision != 2) return fail(EFG_ERR_ARG, "precision must be 0 (FP32), 1 (BF16) or 2 (FP16)");
  if (nq == 0) return EFG_OK;
  if (!d_base || !d_queries || !d_ids || q_stride < d) return fail(EFG_ERR_ARG, "invalid arguments");
  if (nq >= (size_t(1) << 31) || n >= (size_t(1) << 31)) return fail(EFG_ERR_ARG, "too large");
  std::string err;
  int rc = exact_knn_mfma(d_base, n, int(d), base_stride, d_queries, int(nq), q_stride, k, exclude_self, precision,
                          d_ids, d_dists, static_cast<hipStream_t>(stream), &err);
  if (rc) return fail(rc, err);
  return EFG_OK;
}

int efg_tree_shape(const efg_index* ix, int* trees, size_t* nodes) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (trees) *trees = ix->T;
  if (nodes) *nodes = ix->M;
  return EFG_OK;
}

int efg_get_trees(const efg_index* ix, int32_t* dim, float* val, int32_t* left, int32_t* right, uint32_t* start,
                  uint32_t* end, int32_t* roots, uint32_t* leaf_lists) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  HostTrees h;
  if (int rc = download_trees(ix, h)) return rc;
  const size_t M = ix->M;
  if (dim) std::memcpy(dim, h.dim.data(), M * 4);
  if (val) std::memcpy(val, h.val.data(), M * 4);
  if (left) std::memcpy(left, h.left.data(), M * 4);
  if (right) std::memcpy(right, h.right.data(), M * 4);
  if (start) std::memcpy(start, h.start.data(), M * 4);
  if (end) std::memcpy(end, h.end.data(), M * 4);
  if (roots) std::memcpy(roots, h.roots.data(), size_t(h.T) * 4);
  if (leaf_lists) std::memcpy(leaf_lists, h.leaf.data(), h.leaf.size() * 4);
  return EFG_OK;
}

int efg_set_trees(efg_index* ix, int trees, size_t nodes, const int32_t* dim, const float* val, const int32_t* left,
                  const int32_t* right, const uint32_t* start, const uint32_t* end, const int32_t* roots,
                  const uint32_t* leaf_lists) {
  if (!ix || trees <= 0 || nodes == 0) return fail(EFG_ERR_ARG, "invalid trees");
  HostTrees h;
  h.T = trees;
  h.dim.assign(dim, dim + nodes);
  h.val.assign(val, val + nodes);
  h.left.assign(left, left + nodes);
  h.right.assign(right, right + nodes);
  h.start.assign(start, start + nodes);
  h.end.assign(end, end + nodes);
  h.roots.assign(roots, roots + trees);
  h.leaf.assign(leaf_lists, leaf_lists + size_t(trees) * ix->n);
  for (size_t i = 0; i < nodes; ++i) {
    bool inner = h.left[i] >= 0;
    if (inner != (h.right[i] >= 0)) return fail(EFG_ERR_ARG, "node with one child");
    if (inner && (h.left[i] >= int64_t(nodes) || h.right[i] >= int64_t(nodes) || h.dim[i] < 0 ||
                  size_t(h.dim[i]) >= ix->d))
      return fail(EFG_ERR_ARG, "bad node link or split dimension");
    if (h.end[i] > ix->n || h.start[i] > h.end[i]) return fail(EFG_ERR_ARG, "bad node range");
  }
  for (uint32_t x : h.leaf)
    if (x >= ix->n) return fail(EFG_ERR_ARG, "leaf list id out of range");
  h.leaf_dim.assign(nodes, -1);
  int rc = upload_trees(ix, h);
  if (!rc) ix->s_trees = trees;
  return rc;
}

int efg_get_graph(const efg_index* ix, uint64_t* off, uint32_t* ids, size_t cap) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  std::vector<uint64_t> o;
  std::vector<uint32_t> g;
  if (int rc = download_graph(ix, o, g)) return rc;
  if (off) std::memcpy(off, o.data(), o.size() * 8);
  if (ids) std::memcpy(ids, g.data(), std::min(cap, g.size()) * 4);
  return EFG_OK;
}

int efg_set_graph(efg_index* ix, const uint64_t* off, const uint32_t* ids) {
  if (!ix || !off) return fail(EFG_ERR_ARG, "null argument");
  std::vector<uint64_t> o(off, off + ix->n + 1);
  for (size_t i = 0; i < ix->n; ++i)
    if (o[i + 1] < o[i]) return fail(EFG_ERR_ARG, "offsets must be non-decreasing");
  std::vector<uint32_t> g(ids, ids + o[ix->n]);
  return upload_graph(ix, o, g);
}

int efg_device_data(const efg_index* ix, const float** d_data, size_t* stride) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (d_data) *d_data = ix->data.p;
  if (stride) *stride = ix->D;
  return EFG_OK;
}

int efg_get_counters(const efg_index* ix, uint64_t* out, int n) {
  if (!ix || !out) return fail(EFG_ERR_ARG, "null argument");
  uint64_t c[8];
  std::memcpy(c, ix->counters, sizeof c);
  if (ix->ws.ndist.p) {
    hipStreamSynchronize(ix->stream);
    unsigned long long nd = 0;
    int nv[2] = {0, 0};
    hipMemcpy(&nd, ix->ws.ndist.p, 8, hipMemcpyDeviceToHost);
    if (ix->ws.nover_sum.p) hipMemcpy(nv, ix->ws.nover_sum.p, 8, hipMemcpyDeviceToHost);
    c[0] = nd;
    c[4] = uint64_t(nv[0]);
    c[5] = uint64_t(nv[1]);
  }
  for (int i = 0; i < n && i < 8; ++i) out[i] = c[i];
  return EFG_OK;
}

int efg_get_join_pairs(const efg_index* ix, uint64_t* out, int cap, int* n) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  if (n) *n = int(ix->join_pairs.size());
  for (int i = 0; i < cap && i < int(ix->join_pairs.size()); ++i) out[i] = ix->join_pairs[i];
  return EFG_OK;
}

int efg_set_profiling(efg_index* ix, int on) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  ix->profiling = on != 0;
  return EFG_OK;
}

int efg_get_kernel_times(efg_index* ix, double* ms, int n, int* searches) {
  if (!ix) return fail(EFG_ERR_ARG, "null index");
  double acc[3] = {0, 0, 0};
  const size_t m = ix->prof_events.size() / 4;
  for (size_t i = 0; i < m; ++i) {
    hipEvent_t* e = &ix->prof_events[4 * i];
    hipEventSynchronize(e[3]);
    for (int j = 0; j < 3; ++j) {
      float t = 0.f;
      hipEventElapsedTime(&t, e[j], e[j + 1]);
      acc[j] += t;
    }
    for (int j = 0; j < 4; ++j) hipEventDestroy(e[j]);
  }
  ix->prof_events.clear();
  for (int j = 0; j < n && j < 3; ++j) ms[j] = acc[j];
  if (searches) *searches = int(m);
  return EFG_OK;
}

int efg_get_search_work(efg_index* ix, uint64_t* out, int n) {
  if (!ix || !out) return fail(EFG_ERR_ARG, "null argument");
  unsigned long long w[3] = {0, 0, 0};
  if (ix->work.p) {
    hipStreamSynchronize(ix->stream);
    if (int rc = check_hip(hipDeviceSynchronize(), "sync")) return rc;
    if (int rc = check_hip(hipMemcpy(w, ix->work.p, sizeof w, hipMemcpyDeviceToHost), "copy")) return rc;
    if (int rc = check_hip(hipMemset(ix->work.p, 0, sizeof w), "memset")) return rc;
  }
  for (int i = 0; i < n && i < 3; ++i) out[i] = w[i];
  return EFG_OK;
}

int efg_get_build_times(const efg_index* ix, double* out, int n) {
  if (!ix || !out) return fail(EFG_ERR_ARG, "null argument");
  for (int i = 0; i < n && i < 4; ++i) out[i] = ix->times[i];
  return EFG_OK;
}

}  // extern "C"
