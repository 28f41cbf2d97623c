#include "io.hpp"

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <numeric>

#include "aknn_common.hpp"

namespace aknn {

namespace {

// trim (euclideandistancematrixgpu.cu:1987-1995): spaces only
std::string trim_spaces(const std::string& s) {
  if (s.empty()) return s;
  const size_t first = s.find_first_not_of(' ');
  const size_t f = first == std::string::npos ? s.size() : first;
  const size_t last = s.find_last_not_of(' ');
  return s.substr(f, last - f + 1);
}

// strtok(line, ","): maximal runs of non-comma characters
void split_fields(const std::string& line, std::vector<std::string>& out) {
  out.clear();
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && line[i] == ',') ++i;
    if (i >= line.size()) break;
    const size_t j = line.find(',', i);
    const size_t e = j == std::string::npos ? line.size() : j;
    out.push_back(trim_spaces(line.substr(i, e - i)));
    i = e;
  }
}

}  // namespace

bool read_csv_header(const std::string& path, int* n, int* d) {
  std::ifstream f(path.c_str());
  if (!f.is_open()) return false;
  std::string line;
  std::getline(f, line);
  *n = std::atoi(line.c_str());
  std::getline(f, line);
  *d = std::atoi(line.c_str());
  return true;
}

std::string load_csv(const std::string& path, int64_t n, int d, bool klaster, std::vector<float>& x,
                     std::vector<int32_t>& labels) {
  std::ifstream f(path.c_str());
  if (!f.is_open()) return "Error opening the file " + path;
  std::string line;
  std::getline(f, line);
  std::getline(f, line);
  x.assign(static_cast<size_t>(n) * d, 0.0f);
  labels.assign(klaster ? n : 0, 0);
  std::vector<std::string> fields;
  const size_t expected = static_cast<size_t>(d) + (klaster ? 1 : 0);
  int64_t row = 0;
  while (row < n && std::getline(f, line)) {
    split_fields(line, fields);
    if (fields.size() != expected)
      return "data row " + std::to_string(row) + " has " + std::to_string(fields.size()) + " fields, expected " +
             std::to_string(expected);
    float* dst = x.data() + static_cast<size_t>(row) * d;
    for (int j = 0; j < d; ++j) dst[j] = static_cast<float>(std::atof(fields[j].c_str()));
    if (klaster) labels[row] = std::atoi(fields.back().c_str());
    ++row;
  }
  if (row < n) return "the file holds " + std::to_string(row) + " data rows, expected " + std::to_string(n);
  return std::string();
}

std::string write_result(const std::string& path, bool debug, int metric, int64_t n, int k, const int32_t* ids,
                         const float* vals) {
  std::ofstream ofs(path.c_str(), std::ofstream::trunc | std::ofstream::binary);
  std::ofstream dbg;
  if (debug) dbg.open((path + "DEBUG").c_str(), std::ofstream::trunc | std::ofstream::binary);
  std::ofstream val((path + "VALIDATION").c_str(), std::ofstream::trunc | std::ofstream::binary);
  if (!ofs.is_open()) return "Can not open the file for saving result.";
  ofs << n << ";" << k << ";" << sizeof(long) << "\n";
  const long marker = 0x01020304;
  ofs.write(reinterpret_cast<const char*>(&marker), sizeof(long));
  std::vector<int> order(k);
  for (int64_t p = 0; p < n; ++p) {
    const int32_t* ri = ids + p * k;
    const float* rv = vals + p * k;
    std::iota(order.begin(), order.end(), 0);
    // .cu:2935-2939: ascending by value (COSINE descending); ties by lower id
    std::sort(order.begin(), order.end(), [&](int a, int b) {
      if (rv[a] != rv[b]) return metric == kCosine ? rv[a] > rv[b] : rv[a] < rv[b];
      return ri[a] < ri[b];
    });
    for (int e : order) {
      ofs.write(reinterpret_cast<const char*>(&ri[e]), sizeof(int32_t));
      ofs.write(reinterpret_cast<const char*>(&rv[e]), sizeof(float));
      if (debug && dbg.is_open()) dbg << "NEAR: <" << p << ", " << ri[e] << ">(" << rv[e] << ") ";
    }
    if (debug && dbg.is_open()) dbg << ";\n";
  }
  ofs.close();
  if (debug && dbg.is_open()) dbg.close();
  if (val.is_open()) {
    val << "Everything is OK.";
    val.close();
  }
  if (!ofs) return "error while writing " + path;
  return std::string();
}

}  // namespace aknn
