#include "io.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <numeric>
#include <thread>

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

namespace {

// One data line [b, e) parsed in place with the single-threaded loader's
// semantics: fields are maximal runs of non-comma bytes (strtok), each
// trimmed of spaces; a value is atof of the trimmed field (strtod stops at the
// field's end in place, since a number cannot continue into ' ', ',' or the
// line end); the label is atoi of the last field.  Returns the field count.
size_t parse_line(const char* b, const char* e, int d, bool klaster, float* dst, int32_t* label) {
  const size_t expected = static_cast<size_t>(d) + (klaster ? 1 : 0);
  size_t nf = 0;
  const char* p = b;
  char tmp[64];
  while (p < e) {
    while (p < e && *p == ',') ++p;
    if (p >= e) break;
    const char* q = static_cast<const char*>(std::memchr(p, ',', static_cast<size_t>(e - p)));
    const char* fe = q ? q : e;
    // trim spaces
    const char* fb = p;
    while (fb < fe && *fb == ' ') ++fb;
    const char* fl = fe;
    while (fl > fb && fl[-1] == ' ') --fl;
    if (nf < expected) {
      // copy the trimmed field so the parse cannot see past it (bounded copy;
      // a longer field is parsed from a heap copy)
      const size_t len = static_cast<size_t>(fl - fb);
      std::string big;
      const char* src;
      if (len < sizeof(tmp)) {
        std::memcpy(tmp, fb, len);
        tmp[len] = '\0';
        src = tmp;
      } else {
        big.assign(fb, fl);
        src = big.c_str();
      }
      if (static_cast<int>(nf) < d)
        dst[nf] = static_cast<float>(std::atof(src));
      else if (klaster)
        *label = std::atoi(src);
    }
    ++nf;
    p = fe;
  }
  return nf;
}

}  // namespace

std::string load_csv_parallel(const std::string& path, int64_t n, int d, bool klaster, int threads,
                              std::vector<float>& x, std::vector<int32_t>& labels) {
  std::ifstream f(path.c_str(), std::ios::binary);
  if (!f.is_open()) return "Error opening the file " + path;
  f.seekg(0, std::ios::end);
  const std::streamoff size = f.tellg();
  f.seekg(0, std::ios::beg);
  std::vector<char> buf(static_cast<size_t>(size > 0 ? size : 0));
  if (size > 0) f.read(buf.data(), size);
  if (!f && size > 0) return "error while reading " + path;
  const char* base = buf.data();
  const size_t total = buf.size();
  if (threads < 1) threads = static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
  // pass 1: newlines per byte chunk, in parallel
  const size_t chunk = (total + threads - 1) / std::max(threads, 1) + 1;
  std::vector<std::vector<size_t>> nl(threads);
  {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
      pool.emplace_back([&, t] {
        const size_t b = std::min(total, t * chunk), e = std::min(total, b + chunk);
        for (const char* p = base + b; p < base + e;) {
          const char* q = static_cast<const char*>(std::memchr(p, '\n', static_cast<size_t>(base + e - p)));
          if (!q) break;
          nl[t].push_back(static_cast<size_t>(q - base));
          p = q + 1;
        }
      });
    for (auto& th : pool) th.join();
  }
  // line starts: 0 and every byte after a newline (getline semantics: a last
  // line without a newline still counts; an empty tail after the last newline
  // does not)
  std::vector<size_t> starts;
  starts.push_back(0);
  for (auto& v : nl)
    for (size_t q : v) starts.push_back(q + 1);
  if (starts.back() >= total && starts.size() > 1) starts.pop_back();
  const int64_t lines = static_cast<int64_t>(starts.size());
  x.assign(static_cast<size_t>(n) * d, 0.0f);
  labels.assign(klaster ? n : 0, 0);
  const int64_t avail = std::max<int64_t>(0, lines - 2);
  const int64_t rows = std::min(n, avail);
  // pass 2: rows in parallel; the first bad row (lowest index) is reported
  std::vector<int64_t> bad(threads, -1);
  std::vector<size_t> badf(threads, 0);
  {
    std::vector<std::thread> pool;
    const int64_t per = (rows + threads - 1) / std::max(threads, 1);
    for (int t = 0; t < threads; ++t)
      pool.emplace_back([&, t] {
        const int64_t r0 = std::min(rows, t * per), r1 = std::min(rows, r0 + per);
        for (int64_t r = r0; r < r1; ++r) {
          const size_t lb = starts[r + 2];
          size_t le = (r + 3 < lines) ? starts[r + 3] - 1 : total;  // excludes the '\n'
          if (r + 3 >= lines && le > lb && base[le - 1] == '\n') --le;  // the file's final newline
          int32_t lab = 0;
          const size_t nf = parse_line(base + lb, base + le, d, klaster, x.data() + static_cast<size_t>(r) * d, &lab);
          if (nf != static_cast<size_t>(d) + (klaster ? 1 : 0)) {
            bad[t] = r;
            badf[t] = nf;
            return;
          }
          if (klaster) labels[r] = lab;
        }
      });
    for (auto& th : pool) th.join();
  }
  for (int t = 0; t < threads; ++t)
    if (bad[t] >= 0)
      return "data row " + std::to_string(bad[t]) + " has " + std::to_string(badf[t]) + " fields, expected " +
             std::to_string(d + (klaster ? 1 : 0));
  if (rows < n) return "the file holds " + std::to_string(rows) + " data rows, expected " + std::to_string(n);
  return std::string();
}

bool is_fbin(const std::string& path) {
  return path.size() >= 5 && path.compare(path.size() - 5, 5, ".fbin") == 0;
}

bool read_fbin_header(const std::string& path, int* n, int* d) {
  std::ifstream f(path.c_str(), std::ios::binary);
  if (!f.is_open()) return false;
  uint32_t h[2] = {0, 0};
  f.read(reinterpret_cast<char*>(h), sizeof(h));
  if (!f) return false;
  *n = static_cast<int>(std::min<uint32_t>(h[0], 0x7fffffffu));
  *d = static_cast<int>(std::min<uint32_t>(h[1], 0x7fffffffu));
  return true;
}

std::string load_fbin(const std::string& path, int64_t n, int d, std::vector<float>& x) {
  std::ifstream f(path.c_str(), std::ios::binary);
  if (!f.is_open()) return "Error opening the file " + path;
  uint32_t h[2] = {0, 0};
  f.read(reinterpret_cast<char*>(h), sizeof(h));
  if (!f) return "truncated header in " + path;
  if (static_cast<int64_t>(h[0]) < n || static_cast<int>(h[1]) != d)
    return "the file holds " + std::to_string(h[0]) + " x " + std::to_string(h[1]) + " points, expected " +
           std::to_string(n) + " x " + std::to_string(d);
  x.assign(static_cast<size_t>(n) * d, 0.0f);
  f.read(reinterpret_cast<char*>(x.data()), static_cast<std::streamsize>(sizeof(float) * x.size()));
  if (!f) return "the file ends before " + std::to_string(n) + " points";
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
