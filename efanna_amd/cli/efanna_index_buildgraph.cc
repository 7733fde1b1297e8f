// Build the kNN graph and save it (argument order of samples/efanna_index_buildgraph.cc).
#include <chrono>
#include <efanna.hpp>

#include "fvecs.hpp"
using namespace efanna;

int main(int argc, char** argv) {
  if (argc != 10) {
    std::cout << argv[0] << " data_file save_graph_file trees level epoch L K kNN S" << std::endl;
    return -1;
  }
  std::vector<float> data;
  size_t n;
  int dim;
  if (!load_vecs(argv[1], data, n, dim)) return -1;
  Matrix<float> dataset(n, dim, data.data());
  int trees = atoi(argv[3]), ml = atoi(argv[4]), epochs = atoi(argv[5]), L = atoi(argv[6]), check = atoi(argv[7]),
      k = atoi(argv[8]), S = atoi(argv[9]);
  L2DistanceAVX<float> dist;
  FIndex<float> index(dataset, &dist, KDTreeUbIndexParams(true, trees, ml, epochs, check, L, k, trees, S));
  auto t0 = std::chrono::steady_clock::now();
  index.buildIndex();
  double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cout << s << std::endl;
  if (index.lastStatus()) return 1;
  index.saveGraph(argv[2]);
  return 0;
}
