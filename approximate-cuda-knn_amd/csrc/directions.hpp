// Seeded projection directions of the median-split forest (a rule this
// rebuild adds beside the reference's sign split, which uses pivot
// differences, euclideandistancematrixgpu.cu:2151-2162).
//
// Tree t's directions are Gaussian vectors drawn on the host: a splitmix64
// counter stream seeded by (seed, t) gives uniforms in (0, 1), Box-Muller in
// double turns each pair into two normals (cosine branch first), each rounded
// to float, filled level-major (direction j, then dimension).  The same
// directions are fed to the oracle, so the trees can be compared bit for bit.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace aknn {

class SplitMix64 {
 public:
  explicit SplitMix64(uint64_t s) : s_(s) {}
  uint64_t next() {
    s_ += 0x9E3779B97F4A7C15ull;
    uint64_t z = s_;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  // uniform on the open interval (0, 1): 53 random bits plus one half
  double open01() { return (static_cast<double>(next() >> 11) + 0.5) * 0x1p-53; }

 private:
  uint64_t s_;
};

// `count` directions of length d for tree `tree`, row j at r[j * ld], the
// padding columns d..ld-1 zero.
inline void tree_directions(uint32_t seed, int tree, int count, int d, int ld, std::vector<float>& r) {
  r.assign(static_cast<size_t>(count) * ld, 0.0f);
  SplitMix64 g((static_cast<uint64_t>(seed) << 32) ^
               (static_cast<uint64_t>(static_cast<uint32_t>(tree + 1)) * 0xD1B54A32D192ED03ull));
  const int64_t total = static_cast<int64_t>(count) * d;
  int64_t i = 0;
  auto put = [&](double v) {
    r[static_cast<size_t>(i / d) * ld + static_cast<size_t>(i % d)] = static_cast<float>(v);
    ++i;
  };
  while (i < total) {
    const double u1 = g.open01(), u2 = g.open01();
    const double rad = std::sqrt(-2.0 * std::log(u1)), th = 6.283185307179586 * u2;
    put(rad * std::cos(th));
    if (i < total) put(rad * std::sin(th));
  }
}

// Directions per tree: the levels a median split needs to bring n points
// down to `leaf`, plus 8 for sampling imbalance.  Past them a bucket is split
// by position.
inline int median_depth(int64_t n, int leaf) {
  int lv = 0;
  int64_t c = n;
  const int64_t l = leaf < 1 ? 1 : leaf;
  while (c > l) {
    c = (c + 1) / 2;
    ++lv;
  }
  return lv + 8;
}

}  // namespace aknn
