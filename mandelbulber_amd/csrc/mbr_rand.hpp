// mbr_rand.hpp — the C library's rand() sequence, reproduced privately.
//
// The reference draws its random numbers with srand/rand (Random,
// src/common_math.cpp:25-32): the DOF splat order (src/image.cpp:96-152)
// and the aux light placement (src/shaders.cpp:1134-1212).  Both results
// depend on the exact sequence, so the library reproduces glibc's generator
// instead of calling rand(), which would also disturb the caller's own
// sequence.
#pragma once

#include <cstdint>

namespace mbr {

// glibc's rand() (the TYPE_3 additive generator of random_r: 31 words
// seeded by the Park-Miller LCG, 310 outputs discarded, then
// r[k] = r[k-31] + r[k-3] and output r[k] >> 1), kept local so the library
// does not disturb the process's own rand() stream.
class GlibcRand {
 public:
  explicit GlibcRand(uint32_t seed) {
    if (seed == 0) seed = 1;
    r_[0] = (int32_t)seed;
    int32_t w = (int32_t)seed;  // Schrage's form of w = 16807 w mod (2^31 - 1), on a signed 32-bit word
    for (int i = 1; i < 31; i++) {
      const int64_t hi = w / 127773, lo = w % 127773;
      w = (int32_t)(16807 * lo - 2836 * hi);
      if (w < 0) w += 2147483647;
      r_[i] = w;
    }
    f_ = 3;
    b_ = 0;
    for (int i = 0; i < 310; i++) next();
  }
  int next() {
    r_[f_] = (int32_t)((uint32_t)r_[f_] + (uint32_t)r_[b_]);
    const int out = (int)(((uint32_t)r_[f_] >> 1) & 0x7fffffff);
    f_ = (f_ + 1) % 31;
    b_ = (b_ + 1) % 31;
    return out;
  }

 private:
  int32_t r_[31];
  int f_, b_;
};

}  // namespace mbr
