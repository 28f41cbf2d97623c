// Command line drop-in for the reference's `hello` binary (main.cpp:8-124):
//   aknn <input_file> <output_file> <k> <trees> <leaf> <propagations>
//        <EUCLIDEAN|TAXICAB|COSINE> <gpu_ids> <T|F> [--debug]
// Same argument parsing (atoi, exact "T" and "--debug", comma-separated
// device list of which the first is used), messages and exit codes
// (1 argument count, 2 distance, 3 device, 7 input file, 8 initialization,
// 9 deinitialization).  The reference's CUDA compute-capability checks
// (exit 4 and 5) have no counterpart on HIP devices.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "euclidean_distance_matrix_gpu.hpp"

int main(int argc, char** argv) {
  if (argc != 10 && argc != 11) {
    std::cout << "Wrong number of arguments!\n";
    std::cout << "Proper launching: aknn <input_file> <output_file> <number_of_neighbours> <number_of_trees> "
                 "<number_of_elements_in_leaf> <number_of_propagations> <distance_Algorithm> <id_gpu> <cluster: T/F> "
                 "<--debug>\n";
    std::cout << "Example launching: aknn input.csv output.csv 2 10 10 10 EUCLIDEAN 0 T\n";
    std::cout << "Example launching: aknn input.csv output.csv 2 10 10 10 EUCLIDEAN 0 T --debug\n";
    std::cout << "Available distance algorithms: EUCLIDEAN TAXICAB COSINE\n";
    std::cout << "Attention!\n";
    std::cout << "Delimeter -> ,\n";
    std::cout << "If file has cluster (at the end of line) you should run program with cluster set T, otherwise F\n";
    return 1;
  }
  const std::string inputFile(argv[1]);
  const std::string outputFile(argv[2]);
  const int numberOfNeighbours = std::atoi(argv[3]);
  const int numberOfTrees = std::atoi(argv[4]);
  const int numberOfElemsInLeaf = std::atoi(argv[5]);
  const int numberOfPropagation = std::atoi(argv[6]);
  int typeOfDistance;
  const std::string dist(argv[7]);
  if (dist == "EUCLIDEAN") {
    typeOfDistance = AKNN_DISTANCE_EUCLIDEAN;
  } else if (dist == "TAXICAB") {
    typeOfDistance = AKNN_DISTANCE_TAXICAB;
  } else if (dist == "COSINE") {
    typeOfDistance = AKNN_DISTANCE_COSINE;
  } else {
    std::cout << "Wrong distance algorithm!\n";
    std::cout << "Available distance algorithms: EUCLIDEAN TAXICAB COSINE\n";
    return 2;
  }

  std::vector<int> devices;
  for (char* tok = std::strtok(argv[8], ","); tok != nullptr; tok = std::strtok(nullptr, ","))
    devices.push_back(std::atoi(tok));
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) count = 0;
  for (int dev : devices) {
    if (dev < 0 || dev >= count) {
      std::cout << "Device with id " << dev << " is not located in this computer\n";
      return 3;
    }
  }
  if (devices.empty()) {
    std::cout << "Device with id  is not located in this computer\n";
    return 3;
  }
  const bool klaster = std::string(argv[9]) == "T";
  const bool debugMode = argc == 11 && std::string(argv[10]) == "--debug";

  std::ifstream header(inputFile.c_str(), std::ios::binary);
  if (!header.is_open()) {
    std::cout << "Error opening the file\n";
    return 7;
  }
  int numberOfPoints = 0, dimentionOfData = 0;
  const bool fbin = inputFile.size() >= 5 && inputFile.compare(inputFile.size() - 5, 5, ".fbin") == 0;
  if (fbin) {
    // binary input (no reference counterpart): uint32 n, uint32 d, then the points
    uint32_t hd[2] = {0, 0};
    header.read(reinterpret_cast<char*>(hd), sizeof(hd));
    numberOfPoints = static_cast<int>(hd[0] > 0x7fffffffu ? 0x7fffffffu : hd[0]);
    dimentionOfData = static_cast<int>(hd[1] > 0x7fffffffu ? 0x7fffffffu : hd[1]);
  } else {
    std::string line;
    std::getline(header, line);
    numberOfPoints = std::atoi(line.c_str());
    std::getline(header, line);
    dimentionOfData = std::atoi(line.c_str());
  }
  header.close();

  EuclideanDistanceMatrixGPU edm(debugMode);
  edm.setDataFile(inputFile);
  edm.setResultsFile(outputFile);
  if (const char* seed = std::getenv("AKNN_SEED")) edm.setOption("seed", std::atoll(seed));
  bool error = edm.initialize(numberOfPoints, dimentionOfData, numberOfNeighbours, devices[0], typeOfDistance,
                              klaster, numberOfTrees, numberOfPropagation, numberOfElemsInLeaf);
  if (error) {
    std::cout << "Initialization failed.\n";
    return 8;
  }
  std::cout << "I begin analysis.\n";
  edm.calculate();
  std::cout << "I am writing the solution to a file.\n";
  edm.saveResultToResultFile();  // its error does not change the exit code (main.cpp:115)
  error = edm.deinitialize();
  if (error) {
    std::cout << "Deinitialization failed.\n";
    return 9;
  }
  std::cout << "The end\n";
  return 0;
}
