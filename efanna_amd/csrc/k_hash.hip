// Hashing index seeds (IEH, hashing_index.hpp): the reference's second
// initializer.  A query's seeds are the members of the buckets of its code
// XOR each Hamming-ball mask, mask outer, table inner (getNeighborsIEH32/64_*
// 1108-1130, getNeighbors32/64 943-965); from there the search is the
// kd-tree index's (nnexp, kgraph) or, without a graph, the hash-only K
// nearest of the candidates.
//
// Layout: each table's rows grouped by code (codes ascending, ids ascending
// in a bucket, as BuildHashTable32/64 323-347 / 629-655 append them), the
// distinct codes of each table in ascending order, and one SNode (start,
// end) per bucket, so the search kernel reads a bucket as it reads a leaf of
// a tree: slot j * T + t of a query is its bucket in table t for mask j.
// k_hash_probe finds the buckets in a dense table indexed by code (codes of
// at most 22 bits) or by binary search over the table's codes.
#include <algorithm>
#include <numeric>
#include <vector>

#include "efg_device.hpp"
#include "efg_internal.hpp"

namespace efg {

namespace {

// Hamming-ball masks (generateMask32/64, 349-627 / 657-872): 0, then for
// r = 1..radius the r-bit combinations of bits 0..bits-1 in lexicographic
// order of (i < j < ...), the order of the reference's nested loops.
void gen_masks(int bits, int radius, std::vector<uint64_t>& out) {
  out.assign(1, 0ull);
  for (int r = 1; r <= radius && r <= bits; ++r) {
    std::vector<int> idx(r);
    std::iota(idx.begin(), idx.end(), 0);
    for (;;) {
      uint64_t m = 0;
      for (int i : idx) m |= 1ull << i;
      out.push_back(m);
      int i = r - 1;
      while (i >= 0 && idx[i] == bits - r + i) --i;
      if (i < 0) break;
      ++idx[i];
      for (int j = i + 1; j < r; ++j) idx[j] = idx[j - 1] + 1;
    }
  }
}

// Largest Hamming radius the reference accepts for a code length (118-150).
int radius_cap(int bits) {
  if (bits <= 32) return 13;
  if (bits <= 36) return 11;
  if (bits <= 40) return 10;
  if (bits <= 48) return 9;
  if (bits <= 60) return 8;
  return 7;
}

// Slot li = j * T + t of query list position w: the bucket of table t holding
// code(query, t) XOR mask j, or -1.  Codes of at most kHashDenseBits bits
// index a dense table (one load); longer ones are found by binary search.
__global__ void k_hash_probe(const uint64_t* __restrict__ qcodes, size_t qn, int T,
                             const uint64_t* __restrict__ masks, const uint64_t* __restrict__ ucode,
                             const uint32_t* __restrict__ uoff, const int32_t* __restrict__ dense, int bits,
                             int nleaf, int nq, int32_t* __restrict__ leaves, const int32_t* __restrict__ qlist,
                             const int* __restrict__ nlist) {
  const size_t gid = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int nw = qlist ? *nlist : nq;
  if (gid >= size_t(nw) * nleaf) return;
  const int w = int(gid / nleaf), li = int(gid % nleaf);
  const int qi = qlist ? qlist[w] : w;
  const int j = li / T, t = li - j * T;
  const uint64_t code = qcodes[size_t(t) * qn + qi] ^ masks[j];
  if (dense) {  // code < 2^bits: query codes are checked against the same bound
    leaves[size_t(qi) * nleaf + li] = dense[(size_t(t) << bits) + code];
    return;
  }
  uint32_t lo = uoff[t], hi = uoff[t + 1];
  const uint32_t end = hi;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ucode[mid] < code) lo = mid + 1;
    else hi = mid;
  }
  leaves[size_t(qi) * nleaf + li] = (lo < end && ucode[lo] == code) ? int32_t(lo) : -1;
}

}  // namespace

// Codes as the reference loads them (LoadCode32/64, 206-321): shifted right
// by lenshift, each at most 2^(codelen - lenshift) - 1; the radius is
// clamped as the constructor clamps it.
int hash_set_codes(efg_index* ix, int tables, int codelen, int lenshift, int upbits, int radius, const uint64_t* base,
                   const uint64_t* query, size_t nq) {
  if (tables < 1) return fail(EFG_ERR_ARG, "Total hash table num error!");
  if (codelen < 1 || codelen > 64) return fail(EFG_ERR_ARG, "code length not supported yet!");
  const int bits = codelen - lenshift;
  if (lenshift < 0 || bits <= 0) return fail(EFG_ERR_ARG, "lenShift error: could not be larger than the code length!");
  if (upbits >= bits) return fail(EFG_ERR_ARG, "upbits should be smaller than the actual codelength!");
  if (radius < 0) return fail(EFG_ERR_ARG, "invalid Hamming radius");
  radius = std::min(radius, radius_cap(bits));
  if (!base || (nq && !query)) return fail(EFG_ERR_ARG, "null code array");
  const uint64_t maxv = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
  std::vector<uint64_t> b(size_t(tables) * ix->n), q(size_t(tables) * nq);
  for (size_t i = 0; i < b.size(); ++i) {
    b[i] = base[i] >> lenshift;
    if (b[i] > maxv) return fail(EFG_ERR_ARG, "codefile error! Exceed maximum value");
  }
  for (size_t i = 0; i < q.size(); ++i) {
    q[i] = query[i] >> lenshift;
    if (q[i] > maxv) return fail(EFG_ERR_ARG, "codefile error! Exceed maximum value");
  }
  std::vector<uint64_t> masks;
  gen_masks(bits, radius, masks);
  if (masks.size() * size_t(tables) > size_t(kHashSlots))
    return fail(EFG_ERR_ARG, "Hamming ball too large: masks x tables above 4096 seed slots per query");
  ix->h_base.swap(b);
  ix->h_query.swap(q);
  ix->h_T = tables;
  ix->h_bits = bits;
  ix->h_radius = radius;
  ix->h_nmask = int(masks.size());
  ix->h_nq = nq;
  ix->h_built = false;
  ix->hashing = true;
  if (!ix->h_masks.alloc(masks.size())) return fail(EFG_ERR_MEMORY, "out of device memory (hash masks)");
  return check_hip(hipMemcpy(ix->h_masks.p, masks.data(), masks.size() * 8, hipMemcpyHostToDevice), "copy");
}

// The tables (buildIndexImpl on a hashing index, 874-889).
int hash_build(efg_index* ix) {
  if (!ix->hashing) return fail(EFG_ERR_STATE, "no hash codes");
  const size_t n = ix->n, T = size_t(ix->h_T);
  std::vector<uint32_t> lists(T * n), uoff(T + 1, 0);
  std::vector<uint64_t> ucode;
  std::vector<SNode> nodes;
  for (size_t t = 0; t < T; ++t) {
    const uint64_t* code = ix->h_base.data() + t * n;
    uint32_t* l = lists.data() + t * n;
    std::iota(l, l + n, 0u);
    std::stable_sort(l, l + n, [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    uoff[t] = uint32_t(ucode.size());
    for (size_t i = 0; i < n;) {
      size_t e = i + 1;
      while (e < n && code[l[e]] == code[l[i]]) ++e;
      SNode nd{};
      nd.dim = -1;
      nd.left = nd.right = nd.parent = -1;
      nd.start = uint32_t(i);
      nd.end = uint32_t(e);
      ucode.push_back(code[l[i]]);
      nodes.push_back(nd);
      i = e;
    }
  }
  uoff[T] = uint32_t(ucode.size());
  std::vector<int32_t> dense;
  if (ix->h_bits <= kHashDenseBits) {  // bucket index by code
    dense.assign(T << ix->h_bits, -1);
    for (size_t t = 0; t < T; ++t)
      for (uint32_t b = uoff[t]; b < uoff[t + 1]; ++b) dense[(t << ix->h_bits) + ucode[b]] = int32_t(b);
    if (!ix->h_dense.alloc(dense.size())) return fail(EFG_ERR_MEMORY, "out of device memory (hash tables)");
    if (int rc = check_hip(hipMemcpy(ix->h_dense.p, dense.data(), dense.size() * 4, hipMemcpyHostToDevice), "copy"))
      return rc;
  } else {
    ix->h_dense.reset();
  }
  if (!ix->h_lists.alloc(lists.size()) || !ix->h_ucode.alloc(ucode.size()) || !ix->h_uoff.alloc(uoff.size()) ||
      !ix->h_nodes.alloc(nodes.size()) || !ix->h_qcodes.alloc(std::max<size_t>(ix->h_query.size(), 1)))
    return fail(EFG_ERR_MEMORY, "out of device memory (hash tables)");
  int rc = 0;
  rc |= check_hip(hipMemcpy(ix->h_lists.p, lists.data(), lists.size() * 4, hipMemcpyHostToDevice), "copy");
  rc |= check_hip(hipMemcpy(ix->h_ucode.p, ucode.data(), ucode.size() * 8, hipMemcpyHostToDevice), "copy");
  rc |= check_hip(hipMemcpy(ix->h_uoff.p, uoff.data(), uoff.size() * 4, hipMemcpyHostToDevice), "copy");
  rc |= check_hip(hipMemcpy(ix->h_nodes.p, nodes.data(), nodes.size() * sizeof(SNode), hipMemcpyHostToDevice),
                  "copy");
  if (!ix->h_query.empty())
    rc |= check_hip(hipMemcpy(ix->h_qcodes.p, ix->h_query.data(), ix->h_query.size() * 8, hipMemcpyHostToDevice),
                    "copy");
  if (rc) return EFG_ERR_DEVICE;
  ix->h_built = true;
  return EFG_OK;
}

int hash_probe(efg_index* ix, int nleaf, size_t nq, int32_t* leaves, const int32_t* qlist, const int* nlist,
               hipStream_t st) {
  const size_t threads = nq * size_t(nleaf);
  k_hash_probe<<<unsigned((threads + 255) / 256), 256, 0, st>>>(
      ix->h_qcodes.p, ix->h_nq, ix->h_T, ix->h_masks.p, ix->h_ucode.p, ix->h_uoff.p,
      ix->h_bits <= kHashDenseBits ? ix->h_dense.p : nullptr, ix->h_bits, nleaf, int(nq), leaves, qlist, nlist);
  return EFG_OK;
}

}  // namespace efg
