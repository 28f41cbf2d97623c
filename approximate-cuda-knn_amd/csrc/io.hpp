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

// Header lines of the CSV: number of points and dimensions (main.cpp:88-101).
// Returns false if the file cannot be opened.
bool read_csv_header(const std::string& path, int* n, int* d);

// saveResultToResultFile (euclideandistancematrixgpu.cu:2908-2978).
// Returns an empty string on success.
std::string write_result(const std::string& path, bool debug, int metric, int64_t n, int k, const int32_t* ids,
                         const float* vals);

}  // namespace aknn
