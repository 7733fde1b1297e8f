// recall@k of a results file against a ground-truth ivecs file (samples/evaluate.cc:24-65).
#include <set>

#include "fvecs.hpp"

int main(int argc, char** argv) {
  if (argc != 4) {
    std::cout << argv[0] << " resultfile gound_truth kNN" << std::endl;
    return -1;
  }
  std::vector<int> gt, gt_;
  int p1, p2, dim1, dim2;
  if (!load_ivecs(argv[1], gt, p1, dim1) || !load_ivecs(argv[2], gt_, p2, dim2)) return -1;
  if (p1 != p2) {
    std::cout << "result file and groundtruth don't match" << std::endl;
    return -1;
  }
  int k = atoi(argv[3]);
  if (k > dim1) {
    std::cout << "result file not enough for k=" << k << std::endl;
    return -1;
  }
  for (int i = 0; i < p1; ++i) {
    std::set<int> s(gt.begin() + size_t(i) * dim1, gt.begin() + size_t(i) * dim1 + k);
    if (s.size() < size_t(k)) std::cout << "query " << i << " result index not unique!" << std::endl;
  }
  long miss = 0;
  for (int i = 0; i < p1; ++i)
    for (int j = 0; j < k; ++j) {
      int x = gt[size_t(i) * dim1 + j];
      bool hit = false;
      for (int m = 0; m < k; ++m)
        if (x == gt_[size_t(i) * dim2 + m]) { hit = true; break; }
      if (!hit) ++miss;
    }
  std::cout << k << "NN accuracy: " << 1 - float(miss) / float(size_t(p1) * k) << std::endl;
  return 0;
}
