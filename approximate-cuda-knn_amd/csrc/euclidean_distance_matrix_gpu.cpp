#include "euclidean_distance_matrix_gpu.hpp"

#include <iostream>

EuclideanDistanceMatrixGPU::EuclideanDistanceMatrixGPU() : handle_(aknn_create(0)) {}

EuclideanDistanceMatrixGPU::EuclideanDistanceMatrixGPU(bool debugMode) : handle_(aknn_create(debugMode ? 1 : 0)) {}

EuclideanDistanceMatrixGPU::~EuclideanDistanceMatrixGPU() { aknn_destroy(handle_); }

void EuclideanDistanceMatrixGPU::setDataFile(std::string nameOfFile) { aknn_set_data_file(handle_, nameOfFile.c_str()); }

void EuclideanDistanceMatrixGPU::setResultsFile(std::string nameOfFile) {
  aknn_set_results_file(handle_, nameOfFile.c_str());
}

bool EuclideanDistanceMatrixGPU::initialize(int numberOfEntities, int dimensionOfEntity, int numberOfNeighbors,
                                            int device, int typeOfDistance, bool klaster, int numberOfTrees,
                                            int numberOfPropagations, int minSize) {
  device_ = device;
  numberOfEntities_ = numberOfEntities;
  const int rc = aknn_initialize(handle_, numberOfEntities, dimensionOfEntity, numberOfNeighbors, device,
                                 typeOfDistance, klaster ? 1 : 0, numberOfTrees, numberOfPropagations, minSize);
  if (rc != AKNN_OK) {
    std::cout << aknn_last_error(handle_) << "\n";
    return true;
  }
  std::cout << "The datafile has been opened\n";
  return false;
}

bool EuclideanDistanceMatrixGPU::calculate() {
  // progress and timing lines of calculate() (.cu:2864, :2893)
  std::cout << "The device " << device_ << " is calculating the neighbours for: 0 - " << numberOfEntities_ - 1 << "\n";
  const int rc = aknn_calculate(handle_);
  if (rc != AKNN_OK) {
    std::cout << aknn_last_error(handle_) << "\n";
    return true;
  }
  double ms[5] = {0, 0, 0, 0, 0};
  aknn_get_stage_times(handle_, ms, 5);
  std::cout << "The device " << device_ << ": has done task in: " << static_cast<float>(ms[4]) << " ms\n";
  return false;
}

bool EuclideanDistanceMatrixGPU::saveResultToResultFile() {
  const int rc = aknn_save_result(handle_);
  if (rc != AKNN_OK) {
    std::cout << aknn_last_error(handle_) << "\n";
    return true;
  }
  return false;
}

bool EuclideanDistanceMatrixGPU::deinitialize() {
  const int rc = aknn_deinitialize(handle_);
  if (rc != AKNN_OK) {
    std::cout << aknn_last_error(handle_) << "\n";
    return true;
  }
  return false;
}

bool EuclideanDistanceMatrixGPU::setOption(const std::string& name, long long value) {
  return aknn_set_option(handle_, name.c_str(), value) != AKNN_OK;
}
