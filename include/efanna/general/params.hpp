// Parameter types of the efanna API (reference: general/params.hpp:1-45).
#ifndef EFANNA_PARAMS_H_
#define EFANNA_PARAMS_H_
#include <cstddef>
#include <map>
#include <string>

namespace efanna {
enum init_algorithm { KDTREE_UB, HASHING };

union ValueType {
  int int_val;
  float float_val;
  char* str_pt;
};

typedef std::map<std::string, ValueType> ExtraParamsMap;

struct IndexParams {
  init_algorithm init_index_type = KDTREE_UB;
  size_t K = 10;         // graph degree
  size_t S = 10;         // initial neighbour-set size
  size_t L = 30;         // pool size
  size_t TNS = 10;       // tree leaf size
  size_t Check_K = 40;
  int build_epoches = 4;
  int extend_num = 0;
  size_t pool_size = 100;
  size_t init_num = 100;
  ExtraParamsMap extra_params;
  bool reverse_nn_used = true;
  unsigned long long seed = 1998;  // key of the library's counter-based generator
};

struct SearchParams {
  int search_init_num = 0;
  int search_epoches = 0;
  int search_method = 0;
  unsigned extend_to = 0;
  unsigned tree_num = 0;
  unsigned search_depth = 0;
};
}  // namespace efanna
#endif
