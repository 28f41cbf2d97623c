// C++ host interface with the reference's class shape
// (EuclideanDistanceMatrixGPU, euclideandistancematrixgpu.h:50-63): the same
// public methods, argument meaning and bool-error returns, implemented over
// the C ABI of libaknn.so.  Messages go to stdout like the reference's.
#pragma once

#include <string>

#include "aknn.h"

class EuclideanDistanceMatrixGPU {
 public:
  EuclideanDistanceMatrixGPU();
  explicit EuclideanDistanceMatrixGPU(bool debugMode);
  ~EuclideanDistanceMatrixGPU();
  EuclideanDistanceMatrixGPU(const EuclideanDistanceMatrixGPU&) = delete;
  EuclideanDistanceMatrixGPU& operator=(const EuclideanDistanceMatrixGPU&) = delete;

  void setDataFile(std::string nameOfFile);
  bool initialize(int numberOfEntities, int dimensionOfEntity, int numberOfNeighbors, int device,
                  int typeOfDistance, bool klaster, int numberOfTrees, int numberOfPropagations, int minSize);
  bool deinitialize();
  bool calculate();
  void setResultsFile(std::string nameOfFile);
  bool saveResultToResultFile();

  // Options beyond the reference's (see aknn_set_option).
  bool setOption(const std::string& name, long long value);
  aknn_handle* handle() { return handle_; }

 private:
  aknn_handle* handle_;
  int device_ = 0;
  int numberOfEntities_ = 0;
};
