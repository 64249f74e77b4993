// rng.hpp -- the portable generator of the seeded synthetic inputs:
// splitmix64 seeding xoshiro256**, uniforms from the top 53 bits, normals by
// Box-Muller. The state update of xoshiro256** is linear over GF(2), so the
// state after n steps is T^n s; the device mesh generators use that to jump
// straight to each vertex's draws.
#pragma once

#include <cmath>
#include <cstdint>

namespace dpg {

class Rng {
 public:
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) s_[i] = splitmix(x);
  }
  uint64_t next() {
    uint64_t r = rotl(s_[1] * 5, 7) * 9;
    step(s_);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
  double normal() {
    if (has_spare_) {
      has_spare_ = false;
      return spare_;
    }
    double u1 = 1.0 - uniform();  // (0,1]
    double u2 = uniform();
    double r = std::sqrt(-2.0 * std::log(u1));
    spare_ = r * std::sin(2.0 * M_PI * u2);
    has_spare_ = true;
    return r * std::cos(2.0 * M_PI * u2);
  }
  int integer(int lo, int hi) {  // inclusive
    return lo + (int)(uniform() * (hi - lo + 1));
  }
  const uint64_t* state() const { return s_; }
  // the linear state update alone (no output): s <- T s
  static void step(uint64_t s[4]) {
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
  }

 private:
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  uint64_t s_[4];
  bool has_spare_ = false;
  double spare_ = 0.0;
};

}  // namespace dpg
