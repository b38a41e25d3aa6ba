// Host-side numerics for the small, sequential parts of the fit: the curve
// fit for (a, b), the SmallRng streams used by random init, and dense
// symmetric eigen-solves of at most a few thousand rows.
#pragma once

#include <cstdint>
#include <vector>

namespace umapx {

// rand 0.8.5 SmallRng (xoshiro256++, SplitMix64 seeding) and the
// UniformInt / UniformFloat sampling it uses (published algorithms).
struct SmallRng {
  uint64_t s[4];
  explicit SmallRng(uint64_t seed);
  uint64_t next_u64();
  uint32_t next_u32() { return static_cast<uint32_t>(next_u64() >> 32); }
  uint64_t gen_index(uint64_t n);
  float gen_f32(float low, float high);
  double gen_f64(double low, double high);
  double normal(double sd);  // Box-Muller on this stream
};

// find_ab_params (rust_umap/src/lib.rs:5669-5746): 71x121 grid then an
// 80-step pattern search, all in f32.
void find_ab_params(float spread, float min_dist, float* a, float* b);

// random_init (rust_umap/src/lib.rs:3824-3839): U(low, high) row-major.
std::vector<float> random_init(int64_t n, int dim, uint64_t seed, float low, float high);

// Dense symmetric eigen-decomposition (Householder tridiagonalisation +
// implicit QL).  a is n x n row-major; evecs columns are eigenvectors.
void symmetric_eigen(const std::vector<double>& a, int n, std::vector<double>& evals,
                     std::vector<double>& evecs);

}  // namespace umapx
