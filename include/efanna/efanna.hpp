// efanna::FIndex on the MI355X library (reference: efanna.hpp:9-75).
//
// Source compatible with the reference's samples: the same class, methods,
// parameter types and argument orders.  The index lives on the GPU (libefg.so,
// C ABI in efg.h).  Messages the reference prints are printed the same way;
// where the reference calls exit() (unreadable files) this header throws
// std::runtime_error instead.  Only DataType = float is supported.
#ifndef EFANNA
#define EFANNA
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../efg.h"
#include "algorithm/hashing_index.hpp"
#include "algorithm/kdtreeub_index.hpp"
#include "general/distance.hpp"
#include "general/matrix.hpp"
#include "general/params.hpp"

namespace efanna {

template <typename DataType>
class FIndex {
  static_assert(std::is_same<DataType, float>::value, "the MI355X library indexes float data");

 public:
  FIndex(const Matrix<DataType>& features, Distance<DataType>* d, const IndexParams& params) : params_(params) {
    // The four L2 classes (L2Distance, L2DistanceSSE, L2DistanceAVX,
    // L2DistanceAVXr4) map to the device's squared L2; any other distance
    // has no device kernel and is refused.
    if (dynamic_cast<L2DistanceAVX<DataType>*>(d) == nullptr) {
      status_ = EFG_ERR_ARG;
      throw std::invalid_argument("unsupported distance: the MI355X library computes squared L2 only");
    }
    efg_build_params bp;
    int trees = extra(params, "trees", 4);
    efg_default_build_params(&bp, trees);
    bp.trees_build = extra(params, "treesb", trees);
    bp.merge_level = extra(params, "ml", -1);
    bp.epochs = params.build_epoches;
    bp.check = int(params.Check_K);
    bp.L = int(params.L);
    bp.K = int(params.K);
    bp.S = int(params.S);
    bp.tns = int(params.TNS);
    bp.seed = params.seed;
    check(efg_create(features.data(), features.get_rows(), features.get_cols(), features.get_stride(), &bp, &ix_),
          true);
    if (params.init_index_type == HASHING) {  // code files (HASHINGIndex's constructor)
      ExtraParamsMap::const_iterator b = params.extra_params.find("bcfile"), q = params.extra_params.find("qcfile");
      if (b == params.extra_params.end() || q == params.extra_params.end())
        throw std::runtime_error("error: no base or query code file");
      check(efg_hash_load_codes(ix_, extra(params, "tablenum", 0), extra(params, "codelen", 0),
                                extra(params, "lenshift", 0), extra(params, "upbits", 0), extra(params, "radius", 0),
                                b->second.str_pt, q->second.str_pt),
            true);
    }
  }
  virtual ~FIndex() { efg_destroy(ix_); }
  FIndex(const FIndex&) = delete;
  FIndex& operator=(const FIndex&) = delete;

  void buildIndex() { check(efg_build_index(ix_)); }
  void buildTrees() { check(efg_build_trees(ix_)); }
  void knnSearch(int k, const Matrix<DataType>& query) {
    check(efg_knn_search(ix_, k, query.data(), query.get_rows(), query.get_stride()));
  }
  void saveIndex(char* filename) { check(efg_save_index(ix_, filename), true); }
  void loadIndex(char* filename) { check(efg_load_index(ix_, filename), true); }
  void saveTrees(char* filename) { check(efg_save_trees(ix_, filename), true); }
  void loadTrees(char* filename) { check(efg_load_trees(ix_, filename), true); }
  void saveGraph(char* filename) { check(efg_save_graph(ix_, filename), true); }
  void loadGraph(char* filename) { check(efg_load_graph(ix_, filename), true); }
  void saveResults(char* filename) { check(efg_save_results(ix_, filename), true); }
  void setSearchParams(int epochs, int init_num, int extend_to, int search_trees = 0, int search_lv = -1,
                       int search_method = 0) {
    check(efg_set_search_params(ix_, epochs, init_num, extend_to, search_trees, search_lv, search_method));
  }
  size_t getGraphSize() {
    size_t n = 0;
    check(efg_graph_size(ix_, &n));
    return n;
  }
  std::vector<unsigned> getGraphRow(unsigned row_id) {
    size_t len = 0;
    check(efg_graph_row(ix_, row_id, nullptr, 0, &len));
    std::vector<unsigned> row(len);
    if (len) check(efg_graph_row(ix_, row_id, row.data(), len, &len));
    return row;
  }
  void outputVisitBucketNum() {}
  // status of the last call (0 = ok) and the search results as arrays
  int lastStatus() const { return status_; }
  void getResults(std::vector<int>& ids, std::vector<float>& dists, size_t& nq, int& k) {
    check(efg_get_results(ix_, &nq, &k, nullptr, nullptr));
    ids.resize(nq * size_t(k));
    dists.resize(nq * size_t(k));
    check(efg_get_results(ix_, nullptr, nullptr, ids.data(), dists.data()));
  }
  efg_index* handle() { return ix_; }

 private:
  static int extra(const IndexParams& p, const char* key, int dflt) {
    ExtraParamsMap::const_iterator it = p.extra_params.find(key);
    return it == p.extra_params.end() ? dflt : it->second.int_val;
  }
  void check(int rc, bool fatal = false) {
    status_ = rc;
    if (rc == EFG_OK) return;
    std::string msg = efg_last_error();
    if (fatal || rc == EFG_ERR_MEMORY || rc == EFG_ERR_DEVICE) throw std::runtime_error(msg);
    std::cout << msg << std::endl;  // the reference prints and returns
  }
  efg_index* ix_ = nullptr;
  IndexParams params_;
  int status_ = 0;
};

}  // namespace efanna
#endif
