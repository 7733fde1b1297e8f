// KDTreeUbIndexParams (reference: algorithm/kdtreeub_index.hpp:14-35).
#ifndef EFANNA_KDTREE_UB_INDEX_H_
#define EFANNA_KDTREE_UB_INDEX_H_
#include <utility>

#include "../general/params.hpp"

namespace efanna {
struct KDTreeUbIndexParams : public IndexParams {
  KDTreeUbIndexParams(bool rnn_used, int tree_num_total, int merge_level = 4, int epoches = 4, int check = 25,
                      int myL = 30, int building_use_k = 10, int tree_num_build = 0, int myS = 10) {
    reverse_nn_used = rnn_used;
    init_index_type = KDTREE_UB;
    K = building_use_k;
    build_epoches = epoches;
    S = myS;
    ValueType v;
    v.int_val = tree_num_total;
    extra_params.insert(std::make_pair("trees", v));
    v.int_val = tree_num_build > 0 ? tree_num_build : tree_num_total;
    extra_params.insert(std::make_pair("treesb", v));
    v.int_val = merge_level;
    extra_params.insert(std::make_pair("ml", v));
    L = myL;
    Check_K = check;
  }
};
}  // namespace efanna
#endif
