// The pivot stream of the forest builder: glibc's rand() (random_r TYPE_3,
// additive feedback x[i] = x[i-3] + x[i-31] over a 31-word table seeded by the
// Park-Miller LCG, 310 warm-up outputs, result = x >> 1).  The reference calls
// the process-global rand() without srand() (euclideandistancematrixgpu.cu:2483-2484,
// :2589), i.e. seed 1; keeping the stream in the handle makes runs reproducible
// without touching the process-global generator.
#pragma once

#include <cstdint>
#include <vector>

namespace aknn {

class GlibcRand {
 public:
  explicit GlibcRand(uint32_t seed = 1) { reseed(seed); }

  void reseed(uint32_t seed) {
    if (seed == 0) seed = 1;
    int64_t w = static_cast<int32_t>(seed);
    tbl_[0] = static_cast<int32_t>(w);
    for (int i = 1; i < 31; ++i) {
      const int64_t hi = w / 127773, lo = w % 127773;
      w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      tbl_[i] = static_cast<int32_t>(w);
    }
    front_ = 3;
    rear_ = 0;
    for (int i = 0; i < 310; ++i) next();
  }

  int32_t next() {
    const uint32_t v = static_cast<uint32_t>(tbl_[front_]) + static_cast<uint32_t>(tbl_[rear_]);
    tbl_[front_] = static_cast<int32_t>(v);
    front_ = front_ == 30 ? 0 : front_ + 1;
    rear_ = rear_ == 30 ? 0 : rear_ + 1;
    return static_cast<int32_t>(v >> 1);
  }

 private:
  int32_t tbl_[31];
  int front_ = 3, rear_ = 0;
};

// A replayable view of the stream: draws are generated on demand and kept, so
// the host can hand a tree more draws than it ends up using and then resume at
// the first unused one.
class DrawStream {
 public:
  explicit DrawStream(uint32_t seed = 1) : gen_(seed) {}
  void reset(uint32_t seed) {
    gen_.reseed(seed);
    draws_.clear();
    pos_ = 0;
  }
  int32_t at(size_t i) {
    while (draws_.size() <= i) draws_.push_back(gen_.next());
    return draws_[i];
  }
  size_t pos() const { return pos_; }
  void advance(size_t n) { pos_ += n; }

 private:
  GlibcRand gen_;
  std::vector<int32_t> draws_;
  size_t pos_ = 0;
};

// int(((double)r/(RAND_MAX))*INT_MAX) % count (euclideandistancematrixgpu.cu:2482-2484, :2322-2323)
inline int32_t pick_index(int32_t r, int32_t count) {
  const double v = (static_cast<double>(r) / 2147483647.0) * 2147483647.0;
  return static_cast<int32_t>(v) % count;
}

}  // namespace aknn
