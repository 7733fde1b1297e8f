// fvecs / ivecs loading for the command-line tools: each record is an int
// dimension followed by that many 4-byte values (samples/*.cc load_data).
// Rows are stored with a stride rounded up to a multiple of 8 values, zero
// padded, which is the layout efanna::Matrix expects.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <vector>

inline bool load_vecs(const char* filename, std::vector<float>& data, size_t& num, int& dim) {
  std::ifstream in(filename, std::ios::binary);
  if (!in.is_open()) {
    std::cout << "open file error" << std::endl;
    return false;
  }
  in.read(reinterpret_cast<char*>(&dim), 4);
  in.seekg(0, std::ios::end);
  size_t fsize = size_t(in.tellg());
  num = fsize / (size_t(dim) + 1) / 4;
  size_t cols = (size_t(dim) + 7) / 8 * 8;
  data.assign(num * cols, 0.f);
  in.seekg(0, std::ios::beg);
  for (size_t i = 0; i < num; ++i) {
    in.seekg(4, std::ios::cur);
    in.read(reinterpret_cast<char*>(&data[i * cols]), size_t(dim) * 4);
  }
  std::cout << "data dimension: " << dim << std::endl;
  return bool(in);
}

inline bool load_ivecs(const char* filename, std::vector<int>& data, int& num, int& dim) {
  std::ifstream in(filename, std::ios::binary);
  if (!in.is_open()) {
    std::cout << "open file error" << std::endl;
    return false;
  }
  in.read(reinterpret_cast<char*>(&dim), 4);
  in.seekg(0, std::ios::end);
  size_t fsize = size_t(in.tellg());
  num = int(fsize / (size_t(dim) + 1) / 4);
  data.assign(size_t(num) * dim, 0);
  in.seekg(0, std::ios::beg);
  for (int i = 0; i < num; ++i) {
    in.seekg(4, std::ios::cur);
    in.read(reinterpret_cast<char*>(&data[size_t(i) * dim]), size_t(dim) * 4);
  }
  return true;
}
