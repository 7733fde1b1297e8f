// Search with saved trees and graph (argument order of samples/efanna_search.cc).
#include <chrono>
#include <efanna.hpp>

#include "fvecs.hpp"
using namespace efanna;

int main(int argc, char** argv) {
  if (argc != 11 && argc != 12) {
    std::cout << argv[0]
              << " data_file tree_index graph_index query_file result_file tree epoch initsz extend querNN "
                 "search_method(optional)"
              << std::endl;
    return -1;
  }
  std::vector<float> data, qdata;
  size_t n, nq;
  int dim, qdim;
  if (!load_vecs(argv[1], data, n, dim) || !load_vecs(argv[4], qdata, nq, qdim)) return -1;
  Matrix<float> dataset(n, dim, data.data());
  Matrix<float> query(nq, qdim, qdata.data());
  L2DistanceAVX<float> dist;
  FIndex<float> index(dataset, &dist, KDTreeUbIndexParams(true, 8, 8, 10, 25, 30, 10));
  index.loadTrees(argv[2]);
  index.loadGraph(argv[3]);
  int trees = atoi(argv[6]), epochs = atoi(argv[7]), pool = atoi(argv[8]), extend = atoi(argv[9]);
  int method = argc == 12 ? atoi(argv[11]) : 0;
  // the reference passes five arguments here, so `method` lands in search_lv
  // and method 0 always runs (samples/efanna_search.cc:114); method is
  // honoured as a sixth argument instead
  index.setSearchParams(epochs, pool, extend, trees, -1, method);
  auto t0 = std::chrono::steady_clock::now();
  index.knnSearch(atoi(argv[10]), query);
  double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cout << s << std::endl;
  if (index.lastStatus()) return 1;
  index.saveResults(argv[5]);
  return 0;
}
