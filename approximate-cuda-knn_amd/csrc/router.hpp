// Learned router of the median-split forest: an MLP
//   u = (x_raw - mu) * inv_sigma  (D, zero-padded to dp = round_up(D, 16))
//   h1 = relu(W1 u + b1), h2 = relu(W2 h1 + b2), z = W3 h2 + b3  (2^bin_bits logits)
// with BF16 weights and activations on the matrix cores, FP32 accumulation,
// FP32 master weights (Adam).  Trained on the k-NN graph the engine built for
// a training set: the labels are a balanced partition of that graph into
// 2^bin_bits parts (few graph edges cut), so the router learns leaf
// assignments that keep neighbours together; its bins then replace the first
// bin_bits levels of routed median-split trees.  The reference has no
// counterpart (its split directions are random pivot differences,
// euclideandistancematrixgpu.cu:2138-2164).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "aknn_common.hpp"

namespace aknn {

class Engine;

struct RouterTrainStats {
  int epochs = 0;
  double loss = 0, accuracy = 0;  // last epoch, over the training points
  double train_ms = 0;
  double cut_before = 0, cut_after = 0;  // fraction of graph edges cut by the labels (init, refined)
};

class Router {
 public:
  Router(int device, int d, int hidden, int bin_bits, uint32_t seed);
  ~Router();
  Router(const Router&) = delete;
  Router& operator=(const Router&) = delete;

  int device() const { return device_; }
  int d() const { return d_; }
  int dp() const { return dp_; }
  int hidden() const { return h_; }
  int bins() const { return nb_; }
  int bin_bits() const { return bin_bits_; }
  // flat FP32 parameters: W1 [h][dp], b1 [h], W2 [h][h], b2 [h], W3 [bins][h], b3 [bins]
  size_t param_count() const { return off_[6]; }
  size_t offset(int i) const { return off_[i]; }

  // Balanced partition of the engine's graph (host): initial parts are equal
  // runs of the leaf order of tree (seed mod trees), refined by `sweeps`
  // capacity-bounded label propagation sweeps over the symmetrised graph.
  void graph_labels(Engine& e, int sweeps, std::vector<int32_t>& labels);
  // Standardisation (mu, inv_sigma) from the engine's raw points, then
  // `epochs` of Adam over shuffled batches of the engine's raw points.
  void fit(Engine& e, const std::vector<int32_t>& labels, int epochs, int batch, float lr);
  // One forward and backward pass over m raw host points with labels
  // (label < 0: row ignored): gradients of the mean cross-entropy (flat
  // layout) and the loss.  Uses the current weights and standardisation.
  double gradients(const float* x_raw, int64_t m, const int32_t* labels, float* grads);
  // Bins (argmax of the logits, lowest index on ties) of n raw device rows
  // (row stride ld floats) into bins[n] on the device, on stream s.
  void assign(const float* x_raw, int64_t n, int ld, int32_t* bins, hipStream_t s);
  // Logits of m raw host points (m x bins), FP32 (testing).
  void logits(const float* x_raw, int64_t m, float* z);
  // Bins of m raw host points through assign() (testing and host callers).
  void bins_of(const float* x_raw, int64_t m, int32_t* bins);
  // assign() through the fused one-kernel path where the shape allows it
  // (hidden 256, dp <= 256, bins <= 256; default on) or the layer-by-layer one
  void set_fused(bool f) { fused_ = f; }

  void get_params(float* p, float* mu, float* inv_sigma) const;
  void set_params(const float* p, const float* mu, float inv_sigma);
  void save(const std::string& path) const;
  void load(const std::string& path);

  const RouterTrainStats& train_stats() const { return stats_; }
  // time of the last assign() (HIP events; valid once its stream has passed it)
  double last_assign_ms() const;

 private:
  template <typename T>
  T* alloc(size_t count);
  void refresh_weights(hipStream_t s);  // BF16 copies (and transposes) of the master weights
  void ensure_batch(int64_t mpad);
  // forward of the staged batch u of mpad rows (train: also the transposed
  // activations the weight gradients need); backward of its first m rows,
  // the loss averaged over `count` labelled rows
  void forward(int64_t mpad, hipStream_t s, bool train);
  void backward(int64_t mpad, int64_t m, int64_t count, hipStream_t s);

  int device_, d_, dp_, h_, nb_, bin_bits_;
  uint32_t seed_;
  size_t off_[7];
  std::vector<void*> allocs_;
  hipStream_t stream_ = nullptr;
  // parameters (FP32 master), gradients, Adam moments
  float *p_ = nullptr, *g_ = nullptr, *m1_ = nullptr, *m2_ = nullptr;
  // BF16 weights: W1 [h][dp], W2 [h][h], W3 [nb][h]; transposes W2T [h][h], W3T [h][nb]
  uint16_t *w1b_ = nullptr, *w2b_ = nullptr, *w3b_ = nullptr, *w2t_ = nullptr, *w3t_ = nullptr;
  // the same weights in MFMA B-fragment order (fused inference)
  uint16_t *w1p_ = nullptr, *w2p_ = nullptr, *w3p_ = nullptr;
  float* mu_ = nullptr;  // [dp], device
  std::vector<float> mu_host_;
  float inv_sigma_ = 1.0f;
  // batch buffers (mpad rows)
  int64_t cap_ = 0;
  uint16_t *u_ = nullptr, *ut_ = nullptr, *h1_ = nullptr, *h1t_ = nullptr, *h2_ = nullptr, *h2t_ = nullptr;
  float *z_ = nullptr, *dz3f_ = nullptr, *dz2f_ = nullptr, *dz1f_ = nullptr;
  uint16_t *dz3_ = nullptr, *dz3t_ = nullptr, *dz2_ = nullptr, *dz2t_ = nullptr, *dz1t_ = nullptr;
  int32_t *lab_ = nullptr, *idx_ = nullptr;
  double* acc_ = nullptr;  // [loss, correct]
  int64_t step_ = 0;
  RouterTrainStats stats_;
  hipEvent_t ev_[2] = {nullptr, nullptr};
  bool assigned_ = false;
  bool fused_ = true;
};

}  // namespace aknn
