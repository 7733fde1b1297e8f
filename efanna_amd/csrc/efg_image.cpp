// Index image: the device arrays of a built index in one file, loaded
// straight into HBM (efg_save_image / efg_load_image).  The reference's
// saveIndex/loadIndex (kdtreeub_index.hpp:129-137, 701-825) keep trees and
// graph in its host layout, which a loader has to parse and re-lay out; the
// image holds the arrays the search reads as they sit in device memory (the
// depth-first search trees, the descent nodes, leaf lists, CSR graph) plus
// the build-order trees, so loading is file -> pinned staging -> device.
//
// Layout (little endian):
//   header   "EFGIMG01", u32 version, u32 sections, u64 n, d, D, M, nnz,
//            u32 T, i32 reg_deg, i32 max_deg, i32 leaf0_perm, i32 tree_lb,
//            i32 tree_S, i32 lazy, i32 pad, u64 seed, u64 header checksum
//   table    per section: u32 id, u32 element size, u64 count, u64 offset,
//            u64 checksum
//   data     sections at 4096-byte aligned offsets
// Checksums are a 64-bit multiply-xor hash over the section's bytes.  After
// the upload a device kernel range-checks every id and node link, so a
// damaged file is refused before any search reads it.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "efg.h"
#include "efg_internal.hpp"

namespace efg {
int build_search_layout(efg_index* ix);

namespace {

constexpr char kMagic[8] = {'E', 'F', 'G', 'I', 'M', 'G', '0', '1'};
constexpr uint32_t kVersion = 1;
constexpr size_t kAlign = 4096;
constexpr size_t kStage = size_t(64) << 20;  // bytes per staging buffer (two of them)

enum SectionId : uint32_t {
  kNode4 = 1, kParent, kRange, kRoots, kLeaf, kSNode, kDNode, kSRoots, kGOff, kGIds, kLevel, kLeafDim
};

struct Header {
  char magic[8];
  uint32_t version, sections;
  uint64_t n, d, D, M, nnz;
  uint32_t T;
  int32_t reg_deg, max_deg, leaf0_perm, tree_lb, tree_S, lazy, pad;
  uint64_t seed;
  uint64_t check;
};
struct Section {
  uint32_t id, elem;
  uint64_t count, offset, check;
};

// 64-bit multiply-xor hash, streamed over chunks.
struct Hasher {
  uint64_t h = 0x9E3779B97F4A7C15ull;
  void add(const unsigned char* p, size_t len) {
    size_t i = 0;
    for (; i + 8 <= len; i += 8) {
      uint64_t w;
      std::memcpy(&w, p + i, 8);
      h = (h ^ w) * 0xff51afd7ed558ccdull;
      h ^= h >> 29;
    }
    for (; i < len; ++i) h = (h ^ p[i]) * 0x100000001b3ull;
  }
};

uint64_t header_check(Header h) {
  h.check = 0;
  Hasher x;
  x.add(reinterpret_cast<const unsigned char*>(&h), sizeof h);
  return x.h;
}

// A pinned staging pair for streaming between file and device.
struct Staging {
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  ~Staging() {
    for (int i = 0; i < 2; ++i) {
      if (done[i]) hipEventDestroy(done[i]);
      if (buf[i]) hipHostFree(buf[i]);
    }
  }
  int init() {
    for (int i = 0; i < 2; ++i) {
      if (hipHostMalloc(&buf[i], kStage, hipHostMallocDefault) != hipSuccess)
        return fail(EFG_ERR_MEMORY, "out of pinned host memory (image staging)");
      if (hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess)
        return fail(EFG_ERR_DEVICE, "event create failed");
    }
    return EFG_OK;
  }
};

struct SrcArray {
  uint32_t id, elem;
  const void* dev;  // device array, or
  const void* host; // host array
  uint64_t count;
};

// Range checks of an uploaded image: every node link inside [0, M) or -1,
// ranges inside [0, n], leaf-list and graph ids inside [0, n), CSR offsets
// non-decreasing up to nnz, roots inside [0, M).  err[0] counts violations.
__global__ void k_check_image(const int4* __restrict__ node4, const int32_t* __restrict__ parent,
                              const uint2* __restrict__ range, const int32_t* __restrict__ roots,
                              const SNode* __restrict__ sn, const DNode* __restrict__ dn,
                              const int32_t* __restrict__ sroots, const uint32_t* __restrict__ leaf,
                              const uint64_t* __restrict__ goff, const uint32_t* __restrict__ gids, uint64_t n,
                              uint64_t M, uint32_t T, uint64_t nnz, unsigned int* __restrict__ err) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  unsigned bad = 0;
  auto link = [&](int32_t x) { return x == -1 || (x >= 0 && uint64_t(x) < M); };
  for (uint64_t i = g; i < M; i += stride) {
    const int4 a = node4[i];
    bad += !(link(a.z) && link(a.w) && link(parent[i]));
    const uint2 r = range[i];
    bad += !(r.x <= r.y && r.y <= n);
    const SNode s = sn[i];
    bad += !(link(s.left) && link(s.right) && link(s.parent) && s.start <= s.end && s.end <= n);
    bad += !((s.left < 0) == (s.right < 0));
    if (s.left >= 0) bad += !(s.dim >= 0 && uint64_t(s.left) == i + 1);  // pre-order: left child is next
    const DNode e = dn[i];
    bad += !(link(e.right) && link(e.parent));
  }
  for (uint64_t i = g; i < T; i += stride) bad += !(roots[i] >= 0 && uint64_t(roots[i]) < M && sroots[i] >= 0 &&
                                                    uint64_t(sroots[i]) < M);
  for (uint64_t i = g; i < uint64_t(T) * n; i += stride) bad += !(leaf[i] < n);
  for (uint64_t i = g; i < nnz; i += stride) bad += !(gids[i] < n);
  for (uint64_t i = g; i <= n; i += stride) {
    const uint64_t o = goff[i];
    bad += !(o <= nnz && (i == 0 ? o == 0 : o >= goff[i - 1]) && (i < n || o == nnz));
  }
  if (bad) atomicAdd(err, bad);
}

}  // namespace

int save_image(efg_index* ix, const char* path) {
  if (ix->hashing) return fail(EFG_ERR_STATE, "images hold kd-tree indexes");
  if (ix->T == 0 || ix->nnz == 0) return fail(EFG_ERR_STATE, "index needs trees and a graph");
  if (!ix->snode_valid) {
    if (int rc = check_hip(hipStreamSynchronize(ix->stream), "sync")) return rc;
    if (int rc = build_search_layout(ix)) return rc;
  }
  const size_t M = ix->M, T = size_t(ix->T), n = ix->n;
  std::vector<int8_t> lazy_level;
  if (ix->leaf_dim_lazy) lazy_level = ix->node_level;
  std::vector<SrcArray> src = {
      {kNode4, sizeof(int4), ix->node4.p, nullptr, M},
      {kParent, 4, ix->parent.p, nullptr, M},
      {kRange, sizeof(uint2), ix->range.p, nullptr, M},
      {kRoots, 4, ix->roots.p, nullptr, T},
      {kLeaf, 4, ix->leaf.p, nullptr, T * n},
      {kSNode, sizeof(SNode), ix->snode.p, nullptr, M},
      {kDNode, sizeof(DNode), ix->dnode.p, nullptr, M},
      {kSRoots, 4, ix->sroots.p, nullptr, T},
      {kGOff, 8, ix->goff.p, nullptr, n + 1},
      {kGIds, 4, ix->gids.p, nullptr, ix->nnz},
      {kLevel, 1, nullptr, lazy_level.data(), lazy_level.size()},
      {kLeafDim, 1, nullptr, ix->leaf_dim.data(), ix->leaf_dim_lazy ? 0 : ix->leaf_dim.size()},
  };
  Header h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.sections = uint32_t(src.size());
  h.n = n;
  h.d = ix->d;
  h.D = ix->D;
  h.M = M;
  h.nnz = ix->nnz;
  h.T = uint32_t(T);
  h.reg_deg = ix->reg_deg;
  h.max_deg = ix->max_deg;
  h.leaf0_perm = ix->leaf0_perm ? 1 : 0;
  h.tree_lb = ix->tree_lb;
  h.tree_S = ix->tree_S;
  h.lazy = ix->leaf_dim_lazy ? 1 : 0;
  h.seed = ix->bp.seed;
  std::vector<Section> sec(src.size());
  size_t off = (sizeof(Header) + sec.size() * sizeof(Section) + kAlign - 1) / kAlign * kAlign;
  for (size_t i = 0; i < src.size(); ++i) {
    sec[i] = Section{src[i].id, src[i].elem, src[i].count, off, 0};
    off += (src[i].count * src[i].elem + kAlign - 1) / kAlign * kAlign;
  }
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(EFG_ERR_IO, "open file error");
  Staging stg;
  int rc = stg.init();
  // sections: device -> pinned staging -> file, hashed on the way
  for (size_t i = 0; i < src.size() && rc == EFG_OK; ++i) {
    Hasher hs;
    const size_t bytes = src[i].count * src[i].elem;
    if (std::fseek(f, long(sec[i].offset), SEEK_SET) != 0) rc = fail(EFG_ERR_IO, "seek failed");
    for (size_t at = 0; at < bytes && rc == EFG_OK; at += kStage) {
      const size_t len = std::min(kStage, bytes - at);
      const unsigned char* chunk;
      if (src[i].dev) {
        rc = check_hip(hipMemcpy(stg.buf[0], static_cast<const unsigned char*>(src[i].dev) + at, len,
                                 hipMemcpyDeviceToHost), "image download");
        chunk = static_cast<const unsigned char*>(stg.buf[0]);
      } else {
        chunk = static_cast<const unsigned char*>(src[i].host) + at;
      }
      if (rc) break;
      hs.add(chunk, len);
      if (std::fwrite(chunk, 1, len, f) != len) rc = fail(EFG_ERR_IO, "write failed");
    }
    sec[i].check = hs.h;
  }
  if (rc == EFG_OK) {
    h.check = header_check(h);
    // the table is hashed into the header check through the sections' own checks
    Hasher ht;
    ht.add(reinterpret_cast<const unsigned char*>(sec.data()), sec.size() * sizeof(Section));
    h.check ^= ht.h;
    // pad the file to the end of the last section
    if (std::fseek(f, 0, SEEK_SET) != 0 || std::fwrite(&h, sizeof h, 1, f) != 1 ||
        std::fwrite(sec.data(), sizeof(Section), sec.size(), f) != sec.size())
      rc = fail(EFG_ERR_IO, "write failed");
    if (rc == EFG_OK && off > 0) {
      const char z = 0;
      if (std::fseek(f, long(off) - 1, SEEK_SET) != 0 || std::fwrite(&z, 1, 1, f) != 1)
        rc = fail(EFG_ERR_IO, "write failed");
    }
  }
  if (std::fclose(f) != 0 && rc == EFG_OK) rc = fail(EFG_ERR_IO, "write failed");
  return rc;
}

int load_image(efg_index* ix, const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(EFG_ERR_IO, "open file error");
  struct Closer {
    FILE* f;
    ~Closer() { std::fclose(f); }
  } closer{f};
  std::fseek(f, 0, SEEK_END);
  const long fsize = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  Header h{};
  if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, kMagic, 8) != 0)
    return fail(EFG_ERR_FORMAT, "not an index image");
  if (h.version != kVersion) return fail(EFG_ERR_FORMAT, "unsupported index image version");
  if (h.sections != 12) return fail(EFG_ERR_FORMAT, "bad index image section table");
  std::vector<Section> sec(h.sections);
  if (std::fread(sec.data(), sizeof(Section), sec.size(), f) != sec.size())
    return fail(EFG_ERR_FORMAT, "truncated index image");
  {
    Hasher ht;
    ht.add(reinterpret_cast<const unsigned char*>(sec.data()), sec.size() * sizeof(Section));
    if ((header_check(h) ^ ht.h) != h.check) return fail(EFG_ERR_FORMAT, "index image header checksum mismatch");
  }
  if (h.n != ix->n || h.d != ix->d || h.D != ix->D)
    return fail(EFG_ERR_FORMAT, "index image was built for other data (rows or dimension differ)");
  const uint64_t n = h.n, M = h.M, T = h.T;
  if (T == 0 || M < T || M > T * (2 * n + 1) || h.nnz == 0 || h.nnz > (uint64_t(1) << 40) || h.max_deg <= 0 ||
      (h.reg_deg != 0 && h.reg_deg != h.max_deg))
    return fail(EFG_ERR_FORMAT, "index image header out of range");
  const uint64_t want[12][2] = {{sizeof(int4), M}, {4, M}, {sizeof(uint2), M}, {4, T}, {4, T * n},
                                {sizeof(SNode), M}, {sizeof(DNode), M}, {4, T}, {8, n + 1}, {4, h.nnz},
                                {1, h.lazy ? M : 0}, {1, h.lazy ? 0 : M}};
  for (size_t i = 0; i < sec.size(); ++i) {
    const Section& s = sec[i];
    const bool dim_ok = s.id == i + 1 && s.elem == want[i][0] &&
                        (s.count == want[i][1] || (i == 11 && s.count == 0));
    if (!dim_ok || s.offset % kAlign != 0 || s.offset + s.count * s.elem > uint64_t(fsize))
      return fail(EFG_ERR_FORMAT, "index image section table inconsistent with its header");
  }
  // device buffers of the image (the index keeps its own until the image is checked)
  DevBuf<int4> node4;
  DevBuf<int32_t> parent, roots, sroots;
  DevBuf<uint2> range;
  DevBuf<uint32_t> leaf, gids;
  DevBuf<SNode> snode;
  DevBuf<DNode> dnode;
  DevBuf<uint64_t> goff;
  if (!node4.alloc(M) || !parent.alloc(M) || !range.alloc(M) || !roots.alloc(T) || !leaf.alloc(T * n) ||
      !snode.alloc(M) || !dnode.alloc(M) || !sroots.alloc(T) || !goff.alloc(n + 1) || !gids.alloc(h.nnz))
    return fail(EFG_ERR_MEMORY, "out of device memory (index image)");
  void* dst[10] = {node4.p, parent.p, range.p, roots.p, leaf.p, snode.p, dnode.p, sroots.p, goff.p, gids.p};
  std::vector<int8_t> level(sec[10].count), leaf_dim(sec[11].count);
  hipStream_t st = ix->stream;
  Staging stg;
  if (int rc = stg.init()) return rc;
  int slot = 0;
  for (size_t i = 0; i < sec.size(); ++i) {
    const Section& s = sec[i];
    const size_t bytes = s.count * s.elem;
    if (std::fseek(f, long(s.offset), SEEK_SET) != 0) return fail(EFG_ERR_IO, "seek failed");
    Hasher hs;
    if (i >= 10) {  // host sections
      unsigned char* out = reinterpret_cast<unsigned char*>(i == 10 ? level.data() : leaf_dim.data());
      if (bytes && std::fread(out, 1, bytes, f) != bytes) return fail(EFG_ERR_FORMAT, "truncated index image");
      hs.add(out, bytes);
    } else {
      // file -> staging[slot] while the other staging buffer's copy runs
      for (size_t at = 0; at < bytes; at += kStage) {
        const size_t len = std::min(kStage, bytes - at);
        if (int rc = check_hip(hipEventSynchronize(stg.done[slot]), "image staging")) return rc;
        if (std::fread(stg.buf[slot], 1, len, f) != len) return fail(EFG_ERR_FORMAT, "truncated index image");
        hs.add(static_cast<const unsigned char*>(stg.buf[slot]), len);
        if (int rc = check_hip(hipMemcpyAsync(static_cast<unsigned char*>(dst[i]) + at, stg.buf[slot], len,
                                              hipMemcpyHostToDevice, st), "image upload"))
          return rc;
        hipEventRecord(stg.done[slot], st);
        slot ^= 1;
      }
    }
    if (hs.h != s.check) return fail(EFG_ERR_FORMAT, "index image section checksum mismatch");
  }
  // range checks on the device before the index takes the arrays
  DevBuf<unsigned int> err;
  if (!err.alloc(1)) return fail(EFG_ERR_MEMORY, "out of device memory");
  if (int rc = check_hip(hipMemsetAsync(err.p, 0, 4, st), "memset")) return rc;
  k_check_image<<<1024, 256, 0, st>>>(node4.p, parent.p, range.p, roots.p, snode.p, dnode.p, sroots.p, leaf.p, goff.p,
                                      gids.p, n, M, uint32_t(T), h.nnz, err.p);
  unsigned int bad = 0;
  if (int rc = check_hip(hipMemcpyAsync(&bad, err.p, 4, hipMemcpyDeviceToHost, st), "image check")) return rc;
  if (int rc = check_hip(hipStreamSynchronize(st), "image check")) return rc;
  if (bad) return fail(EFG_ERR_FORMAT, "index image holds ids or node links out of range");
  // take the arrays
  auto take = [](auto& to, auto& from) {
    to.reset();
    to.p = from.p;
    to.n = from.n;
    to.owned = true;
    from.p = nullptr;
    from.n = 0;
  };
  take(ix->node4, node4);
  take(ix->parent, parent);
  take(ix->range, range);
  take(ix->roots, roots);
  take(ix->leaf, leaf);
  take(ix->snode, snode);
  take(ix->dnode, dnode);
  take(ix->sroots, sroots);
  take(ix->goff, goff);
  take(ix->gids, gids);
  ix->T = int(T);
  ix->M = size_t(M);
  ix->nnz = size_t(h.nnz);
  ix->reg_deg = h.reg_deg;
  ix->max_deg = h.max_deg;
  ix->leaf0_perm = h.leaf0_perm != 0;
  ix->tree_lb = h.tree_lb;
  ix->tree_S = h.tree_S;
  ix->leaf_dim_lazy = h.lazy != 0;
  ix->node_level = std::move(level);
  ix->leaf_dim = std::move(leaf_dim);
  ix->snode_valid = true;
  ix->s_trees = int(T);
  return EFG_OK;
}

}  // namespace efg
