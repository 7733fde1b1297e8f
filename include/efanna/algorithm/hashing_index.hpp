// HASHINGIndexParams (reference: algorithm/hashing_index.hpp:19-46).
//
// The hashing index (IEH seeds) on the MI355X library: FIndex built with
// these parameters loads the base and query code files "<file>_<t>",
// t = 1..TableNum, and buildIndex() builds the hash tables.  knnSearch then
// seeds every query from the buckets of its code XOR each Hamming-ball mask
// (mask outer, table inner) and, with a graph loaded, expands it by nnexp
// (search_method 1) or kgraph (search_method 0); without a graph it returns
// the K nearest of the bucket candidates.
#ifndef EFANNA_HASHING_INDEX_H_
#define EFANNA_HASHING_INDEX_H_
#include <utility>

#include "../general/params.hpp"

namespace efanna {
struct HASHINGIndexParams : public IndexParams {
  HASHINGIndexParams(int codelen, int TableNum, int UpperBits, int HashRadius, char*& BaseCodeFile,
                     char*& QueryCodeFile, int codelenShift = 0) {
    init_index_type = HASHING;
    ValueType v;
    v.int_val = codelen;
    extra_params.insert(std::make_pair("codelen", v));
    v.int_val = TableNum;
    extra_params.insert(std::make_pair("tablenum", v));
    v.int_val = UpperBits;
    extra_params.insert(std::make_pair("upbits", v));
    v.int_val = HashRadius;
    extra_params.insert(std::make_pair("radius", v));
    v.str_pt = BaseCodeFile;
    extra_params.insert(std::make_pair("bcfile", v));
    v.str_pt = QueryCodeFile;
    extra_params.insert(std::make_pair("qcfile", v));
    v.int_val = codelenShift;
    extra_params.insert(std::make_pair("lenshift", v));
  }
};
}  // namespace efanna
#endif
