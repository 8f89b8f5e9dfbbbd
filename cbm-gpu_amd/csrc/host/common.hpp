// Shared helpers of the host library: error reporting and a seeded PRNG.
#pragma once

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace cbmh {

// Last error message of the calling thread, returned by cbm_host_last_error().
void set_error(const char *fmt, ...);
const char *get_error();

// splitmix64: expands one 64-bit seed into well-mixed state words.
inline uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// xoshiro256** generator (Blackman & Vigna, 2018), seeded through splitmix64.
class Rng {
  public:
    explicit Rng(uint64_t seed) {
        uint64_t s = seed;
        for (auto &w : st_) w = splitmix64(s);
    }
    uint64_t next() {
        const uint64_t r = rotl(st_[1] * 5, 7) * 9;
        const uint64_t t = st_[1] << 17;
        st_[2] ^= st_[0];
        st_[3] ^= st_[1];
        st_[1] ^= st_[2];
        st_[0] ^= st_[3];
        st_[2] ^= t;
        st_[3] = rotl(st_[3], 45);
        return r;
    }
    // Uniform double in [0, 1) from the top 53 bits.
    double uniform() { return double(next() >> 11) * 0x1.0p-53; }
    // Uniform integer in [0, n) by Lemire's multiply-shift (bias < n / 2^64).
    uint64_t below(uint64_t n) {
        return uint64_t((unsigned __int128)next() * n >> 64);
    }
    // Poisson(lam) by Knuth's product-of-uniforms method; for lam > 30 the
    // count is split into independent Poisson(30)-sized pieces.
    int64_t poisson(double lam);

  private:
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t st_[4];
};

}  // namespace cbmh
