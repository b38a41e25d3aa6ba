// Host numerics (compiled with -ffp-contract=off so f32 arithmetic stays
// unfused, as the reference's Rust code is).
#include "hostmath.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace umapx {

SmallRng::SmallRng(uint64_t state) {
  const uint64_t phi = 0x9e3779b97f4a7c15ULL;
  for (int i = 0; i < 4; ++i) {
    state += phi;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    s[i] = z ^ (z >> 31);
  }
  if ((s[0] | s[1] | s[2] | s[3]) == 0) *this = SmallRng(0);
}

uint64_t SmallRng::next_u64() {
  auto rotl = [](uint64_t x, int k) { return (x << k) | (x >> (64 - k)); };
  const uint64_t result = rotl(s[0] + s[3], 23) + s[0];
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl(s[3], 45);
  return result;
}

uint64_t SmallRng::gen_index(uint64_t n) {
  const uint64_t zone = (n << __builtin_clzll(n)) - 1;
  for (;;) {
    const uint64_t v = next_u64();
    const unsigned __int128 m = static_cast<unsigned __int128>(v) * n;
    if (static_cast<uint64_t>(m) <= zone) return static_cast<uint64_t>(m >> 64);
  }
}

float SmallRng::gen_f32(float low, float high) {
  float scale = high - low;
  for (;;) {
    const uint32_t bits = (next_u32() >> 9) | 0x3F800000u;
    float v12;
    std::memcpy(&v12, &bits, 4);
    const float v01 = v12 - 1.0f;
    const float res = v01 * scale + low;
    if (res < high) return res;
    uint32_t sb;
    std::memcpy(&sb, &scale, 4);
    sb -= 1;
    std::memcpy(&scale, &sb, 4);
  }
}

double SmallRng::gen_f64(double low, double high) {
  double scale = high - low;
  for (;;) {
    const uint64_t bits = (next_u64() >> 12) | 0x3FF0000000000000ULL;
    double v12;
    std::memcpy(&v12, &bits, 8);
    const double v01 = v12 - 1.0;
    const double res = v01 * scale + low;
    if (res < high) return res;
    uint64_t sb;
    std::memcpy(&sb, &scale, 8);
    sb -= 1;
    std::memcpy(&scale, &sb, 8);
  }
}

double SmallRng::normal(double sd) {
  double u1 = gen_f64(0.0, 1.0);
  const double u2 = gen_f64(0.0, 1.0);
  if (u1 <= 0.0) u1 = 1e-300;
  return sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
}

namespace {
float curve_loss(float a, float b, const std::vector<float>& xv, const std::vector<float>& yv) {
  float loss = 0.0f;
  const float ex = 2.0f * b;
  for (size_t i = 0; i < xv.size(); ++i) {
    const float model = 1.0f / (1.0f + a * std::pow(xv[i], ex));
    const float e = model - yv[i];
    loss += e * e;
  }
  return loss / static_cast<float>(xv.size());
}
}  // namespace

void find_ab_params(float spread, float min_dist, float* a_out, float* b_out) {
  const int n = 300;
  const float xmax = spread * 3.0f;
  std::vector<float> xv(n), yv(n), xp(n);
  for (int i = 0; i < n; ++i) {
    const float x = xmax * static_cast<float>(i) / (static_cast<float>(n) - 1.0f);
    xv[i] = x;
    yv[i] = x < min_dist ? 1.0f : std::exp(-(x - min_dist) / spread);
  }
  float best_a = 1.0f, best_b = 1.0f, best_loss = std::numeric_limits<float>::infinity();
  for (int bi = 0; bi <= 70; ++bi) {
    const float b = 0.3f + 2.7f * static_cast<float>(bi) / 70.0f;
    const float ex = 2.0f * b;
    for (int i = 0; i < n; ++i) xp[i] = std::pow(xv[i], ex);
    for (int ai = 0; ai <= 120; ++ai) {
      const float la = -3.0f + 6.0f * static_cast<float>(ai) / 120.0f;
      const float a = std::pow(10.0f, la);
      float loss = 0.0f;
      for (int i = 0; i < n; ++i) {
        const float model = 1.0f / (1.0f + a * xp[i]);
        const float e = model - yv[i];
        loss += e * e;
      }
      loss = loss / static_cast<float>(n);
      if (loss < best_loss) {
        best_loss = loss;
        best_a = a;
        best_b = b;
      }
    }
  }
  float step_a = std::max(best_a * 0.35f, 0.05f);
  float step_b = 0.2f;
  for (int it = 0; it < 80; ++it) {
    const float lo_a = std::max(best_a - step_a, 1e-6f);
    const float lo_b = std::max(best_b - step_b, 1e-6f);
    const float cand[8][2] = {{best_a + step_a, best_b}, {lo_a, best_b},
                              {best_a, best_b + step_b}, {best_a, lo_b},
                              {best_a + step_a, best_b + step_b}, {best_a + step_a, lo_b},
                              {lo_a, best_b + step_b}, {lo_a, lo_b}};
    bool improved = false;
    for (const auto& c : cand) {
      const float loss = curve_loss(c[0], c[1], xv, yv);
      if (loss < best_loss) {
        best_loss = loss;
        best_a = c[0];
        best_b = c[1];
        improved = true;
      }
    }
    if (!improved) {
      step_a *= 0.7f;
      step_b *= 0.7f;
      if (step_a < 1e-5f && step_b < 1e-5f) break;
    }
  }
  *a_out = best_a;
  *b_out = best_b;
}

std::vector<float> random_init(int64_t n, int dim, uint64_t seed, float low, float high) {
  SmallRng rng(seed);
  std::vector<float> out(static_cast<size_t>(n) * dim);
  for (auto& v : out) v = rng.gen_f32(low, high);
  return out;
}

void symmetric_eigen(const std::vector<double>& a_in, int n, std::vector<double>& evals,
                     std::vector<double>& evecs) {
  // Householder reduction to tridiagonal form (accumulating Q), then the
  // implicit QL iteration with Wilkinson-style shifts.
  std::vector<double> a = a_in;
  std::vector<double> d(n, 0.0), e(n, 0.0);
  auto A = [&](int i, int j) -> double& { return a[static_cast<size_t>(i) * n + j]; };
  for (int i = n - 1; i > 0; --i) {
    const int l = i - 1;
    double h = 0.0, scale = 0.0;
    if (l > 0) {
      for (int k = 0; k <= l; ++k) scale += std::fabs(A(i, k));
      if (scale == 0.0) {
        e[i] = A(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          A(i, k) /= scale;
          h += A(i, k) * A(i, k);
        }
        double f = A(i, l);
        double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        A(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          A(j, i) = A(i, j) / h;
          g = 0.0;
          for (int k = 0; k <= j; ++k) g += A(j, k) * A(i, k);
          for (int k = j + 1; k <= l; ++k) g += A(k, j) * A(i, k);
          e[j] = g / h;
          f += e[j] * A(i, j);
        }
        const double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          f = A(i, j);
          g = e[j] - hh * f;
          e[j] = g;
          for (int k = 0; k <= j; ++k) A(j, k) -= (f * e[k] + g * A(i, k));
        }
      }
    } else {
      e[i] = A(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    const int l = i - 1;
    if (d[i] != 0.0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0.0;
        for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
        for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
      }
    }
    d[i] = A(i, i);
    A(i, i) = 1.0;
    for (int j = 0; j <= l; ++j) A(j, i) = A(i, j) = 0.0;
  }
  // QL
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  if (n > 0) e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= DBL_EPSILON * dd) break;
      }
      if (m != l) {
        if (iter++ == 60) break;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          const double b = c * e[i];
          r = std::hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          for (int k = 0; k < n; ++k) {
            f = A(k, i + 1);
            A(k, i + 1) = s * A(k, i) + c * f;
            A(k, i) = c * A(k, i) - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  evals = d;
  evecs = a;
}

}  // namespace umapx
