// Host I/O of the drop-in boundary: the reference's CSV input and binary
// result file.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace aknn {

// loadData (euclideandistancematrixgpu.cu:2712-2757): skips two header lines,
// splits each line on ',' (strtok: empty fields vanish), trims spaces, parses
// with atof and stores float.  With klaster the last field is an atoi label.
// Unlike the reference, rows with the wrong field count or a short file are
// errors (the reference writes out of bounds or leaves rows uninitialised).
// x receives n*d floats, row-major.  Returns an empty string on success.
std::string load_csv(const std::string& path, int64_t n, int d, bool klaster, std::vector<float>& x,
                     std::vector<int32_t>& labels);

// load_csv with the file read at once and parsed by `threads` host threads
// (< 1: up to 16): the same values, labels and errors (the first bad row).
std::string load_csv_parallel(const std::string& path, int64_t n, int d, bool klaster, int threads,
                              std::vector<float>& x, std::vector<int32_t>& labels);

// Binary input (no reference counterpart; for inputs whose CSV text would be
// gigabytes): a file named *.fbin holds uint32 n, uint32 d, then n*d float32,
// little-endian, row-major.  The first n points are read.
bool is_fbin(const std::string& path);
bool read_fbin_header(const std::string& path, int* n, int* d);
std::string load_fbin(const std::string& path, int64_t n, int d, std::vector<float>& x);

// Header lines of the CSV: number of points and dimensions (main.cpp:88-101).
// Returns false if the file cannot be opened.
bool read_csv_header(const std::string& path, int* n, int* d);

// saveResultToResultFile (euclideandistancematrixgpu.cu:2908-2978).
// Returns an empty string on success.
std::string write_result(const std::string& path, bool debug, int metric, int64_t n, int k, const int32_t* ids,
                         const float* vals);

}  // namespace aknn
