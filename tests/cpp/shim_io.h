// File helpers for the overload-layer test programs: raw little-endian
// int32 / float32 arrays and a "key value" meta file written by the tests.
#pragma once
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

template <class T>
std::vector<T> load(const std::string &path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot open " + path);
  const size_t bytes = (size_t)f.tellg();
  std::vector<T> v(bytes / sizeof(T));
  f.seekg(0);
  f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)bytes);
  return v;
}

template <class T>
void save(const std::string &path, const T *p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(p), (std::streamsize)(n * sizeof(T)));
}

inline std::map<std::string, double> load_meta(const std::string &path) {
  std::ifstream f(path);
  std::map<std::string, double> m;
  std::string k;
  double v;
  while (f >> k >> v) m[k] = v;
  return m;
}

template <class Info>
void save_info(const std::string &path, const Info &i) {
  std::ofstream f(path);
  f.precision(17);
  f << i.outer_iters << " " << i.total_iters << " " << (double)i.rel_chg << " "
    << (double)i.rel_error << "\n";
}

template <class F>
void expect_error(const char *name, F f) {
  try {
    f();
    std::printf("%s: NO ERROR\n", name);
  } catch (const tval3_exception &e) {
    std::printf("%s: %s\n", name, e.what());
  }
}
