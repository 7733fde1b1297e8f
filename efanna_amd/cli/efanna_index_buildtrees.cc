// Build the trees only and save them (argument order of samples/efanna_index_buildtrees.cc).
#include <chrono>
#include <efanna.hpp>

#include "fvecs.hpp"
using namespace efanna;

int main(int argc, char** argv) {
  if (argc != 4) {
    std::cout << argv[0] << " data_file save_trees_file trees" << std::endl;
    return -1;
  }
  std::vector<float> data;
  size_t n;
  int dim;
  if (!load_vecs(argv[1], data, n, dim)) return -1;
  Matrix<float> dataset(n, dim, data.data());
  int trees = atoi(argv[3]);
  L2DistanceAVX<float> dist;
  FIndex<float> index(dataset, &dist, KDTreeUbIndexParams(true, trees, 8, 8, 25, 30, 10, 8, 10));
  auto t0 = std::chrono::steady_clock::now();
  index.buildTrees();
  double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::cout << "Index building time : " << s << " seconds" << std::endl;
  index.saveTrees(argv[2]);
  return 0;
}
