/* CPU restatement of the reference's sparse (CSR) exact kNN -- test
 * infrastructure only (tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it; the product never does).
 *
 * Follows rust_umap/src/sparse.rs:
 *   SparseCsrMatrix::new checks and row summaries     :20-107
 *   sparse_dot (merge over the nonzeros)              :220-239
 *   euclidean_distance_from_dot                       :241-244
 *   row distances (manhattan, chebyshev, minkowski,
 *     cosine, correlation) by merge                   :246-414, 416-603
 *   exact_nearest_neighbors (fit: self at 0.0)        :605-626
 *   exact_nearest_neighbors_sparse_query              :628-657
 *   dense query against CSR rows                      :659-757+
 *   top-k by (f32 total order, index)                 :150-206
 * All f32 arithmetic is in the reference's order without contraction
 * (the Makefile builds with -ffp-contract=off). */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include <omp.h>

#include "umap_oracle.h"

namespace {

using std::vector;

inline int32_t tkey(float v) {
  int32_t b;
  std::memcpy(&b, &v, 4);
  return b ^ static_cast<int32_t>(static_cast<uint32_t>(b >> 31) >> 1);
}

/* sparse.rs:150-206: keep the k best by (dist total order, index) */
struct Top {
  vector<int64_t> idx;
  vector<float> dist;
  size_t k;
  explicit Top(size_t k_) : k(k_) {}
  static bool less(int64_t li, float ld, int64_t ri, float rd) {
    const int32_t a = tkey(ld), b = tkey(rd);
    return a != b ? a < b : li < ri;
  }
  void push(int64_t i, float d) {
    if (k == 0) return;
    if (idx.size() == k && !less(i, d, idx[k - 1], dist[k - 1])) return;
    size_t at = idx.size();
    while (at > 0 && less(i, d, idx[at - 1], dist[at - 1])) --at;
    idx.insert(idx.begin() + at, i);
    dist.insert(dist.begin() + at, d);
    if (idx.size() > k) {
      idx.pop_back();
      dist.pop_back();
    }
  }
  void store(int64_t row, int32_t kk, int64_t* oi, float* od) const {
    for (int32_t c = 0; c < kk; ++c) {
      const bool have = c < static_cast<int32_t>(idx.size());
      oi[row * kk + c] = have ? idx[c] : -1;
      od[row * kk + c] = have ? dist[c] : INFINITY;
    }
  }
};

struct Csr {
  int64_t rows, cols;
  const int64_t* ptr;
  const int64_t* ind;
  const float* val;
  vector<float> sq, l2, sum;
  Csr(int64_t r, int64_t c, const int64_t* p, const int64_t* i, const float* v)
      : rows(r), cols(c), ptr(p), ind(i), val(v), sq(r), l2(r), sum(r) {
    /* sparse.rs:92-100: per-row sums in nonzero order */
    for (int64_t row = 0; row < r; ++row) {
      float s2 = 0.0f, s1 = 0.0f;
      for (int64_t e = p[row]; e < p[row + 1]; ++e) s2 += v[e] * v[e];
      for (int64_t e = p[row]; e < p[row + 1]; ++e) s1 += v[e];
      sq[row] = s2;
      sum[row] = s1;
      l2[row] = std::sqrt(s2);
    }
  }
};

/* sparse.rs:220-239 */
inline float sdot(const int64_t* li, const float* lv, int64_t ln, const int64_t* ri,
                  const float* rv, int64_t rn) {
  int64_t i = 0, j = 0;
  float dot = 0.0f;
  while (i < ln && j < rn) {
    if (li[i] < ri[j]) {
      ++i;
    } else if (li[i] > ri[j]) {
      ++j;
    } else {
      dot += lv[i] * rv[j];
      ++i;
      ++j;
    }
  }
  return dot;
}

/* lib.rs:2678-2698 correlation_distance_from_summaries */
inline float corr_summ(int64_t nfeat, float sx, float qx, float sy, float qy, float dot) {
  const float n = static_cast<float>(nfeat);
  const float cd = dot - (sx * sy / n);
  const float xv = std::fmax(qx - (sx * sx / n), 0.0f);
  const float yv = std::fmax(qy - (sy * sy / n), 0.0f);
  const float xn = std::sqrt(xv), yn = std::sqrt(yv);
  if (xn == 0.0f && yn == 0.0f) return 0.0f;
  if (xn == 0.0f || yn == 0.0f) return 1.0f;
  const float c = std::min(std::max(cd / (xn * yn), -1.0f), 1.0f);
  return 1.0f - c;
}

/* Merged walk over the union of two rows' nonzeros, in column order:
 * op(|a|), op(|b|) or op(|a - b|) per column (sparse.rs:254-366, 416-538). */
template <typename Op>
inline void merge_abs(const int64_t* li, const float* lv, int64_t ln, const int64_t* ri,
                      const float* rv, int64_t rn, Op op) {
  int64_t i = 0, j = 0;
  while (i < ln && j < rn) {
    if (li[i] < ri[j]) {
      op(std::fabs(lv[i]));
      ++i;
    } else if (li[i] > ri[j]) {
      op(std::fabs(rv[j]));
      ++j;
    } else {
      op(std::fabs(lv[i] - rv[j]));
      ++i;
      ++j;
    }
  }
  while (i < ln) op(std::fabs(lv[i++]));
  while (j < rn) op(std::fabs(rv[j++]));
}

/* sparse_row_distance_between (sparse.rs:561-603); both rows from CSR */
float csr_pair(const Csr& a, int64_t ra, const Csr& b, int64_t rb, int32_t metric, float p) {
  const int64_t* li = a.ind + a.ptr[ra];
  const float* lv = a.val + a.ptr[ra];
  const int64_t ln = a.ptr[ra + 1] - a.ptr[ra];
  const int64_t* ri = b.ind + b.ptr[rb];
  const float* rv = b.val + b.ptr[rb];
  const int64_t rn = b.ptr[rb + 1] - b.ptr[rb];
  switch (metric) {
    case 0: { /* euclidean_distance_from_dot */
      const float dot = sdot(li, lv, ln, ri, rv, rn);
      return std::sqrt(std::fmax(a.sq[ra] + b.sq[rb] - 2.0f * dot, 0.0f));
    }
    case 1: { /* sparse_slices_cosine_distance */
      const float nl = a.l2[ra], nr = b.l2[rb];
      if (nl == 0.0f && nr == 0.0f) return 0.0f;
      if (nl == 0.0f || nr == 0.0f) return 1.0f;
      const float dot = sdot(li, lv, ln, ri, rv, rn);
      const float c = std::min(std::max(dot / (nl * nr), -1.0f), 1.0f);
      return 1.0f - c;
    }
    case 2: {
      float s = 0.0f;
      merge_abs(li, lv, ln, ri, rv, rn, [&](float v) { s += v; });
      return s;
    }
    case 3: {
      float m = 0.0f;
      merge_abs(li, lv, ln, ri, rv, rn, [&](float v) { m = std::fmax(m, v); });
      return m;
    }
    case 4: {
      float s = 0.0f;
      merge_abs(li, lv, ln, ri, rv, rn, [&](float v) { s += std::pow(v, p); });
      return std::pow(s, 1.0f / p);
    }
    default: { /* 5 correlation */
      const float dot = sdot(li, lv, ln, ri, rv, rn);
      return corr_summ(a.cols, a.sum[ra], a.sq[ra], b.sum[rb], b.sq[rb], dot);
    }
  }
}

/* dense query row q (d features) against CSR row rb (sparse.rs:659-757+) */
struct DenseQ {
  float sq = 0.0f, l1 = 0.0f, ps = 0.0f, sum = 0.0f, l2 = 0.0f;
};

DenseQ dense_summaries(const float* q, int64_t d, int32_t metric, float p) {
  DenseQ s;
  for (int64_t f = 0; f < d; ++f) s.sq += q[f] * q[f];
  for (int64_t f = 0; f < d; ++f) s.l1 += std::fabs(q[f]);
  if (metric == 4)
    for (int64_t f = 0; f < d; ++f) s.ps += std::pow(std::fabs(q[f]), p);
  for (int64_t f = 0; f < d; ++f) s.sum += q[f];
  s.l2 = std::sqrt(s.sq);
  return s;
}

float dense_pair(const float* q, int64_t d, const DenseQ& qs, const Csr& b, int64_t rb,
                 int32_t metric, float p) {
  const int64_t* ci = b.ind + b.ptr[rb];
  const float* cv = b.val + b.ptr[rb];
  const int64_t cn = b.ptr[rb + 1] - b.ptr[rb];
  switch (metric) {
    case 0: {
      float dot = 0.0f;
      for (int64_t e = 0; e < cn; ++e) dot += q[ci[e]] * cv[e];
      return std::sqrt(std::fmax(qs.sq + b.sq[rb] - 2.0f * dot, 0.0f));
    }
    case 1: {
      if (qs.l2 == 0.0f && b.l2[rb] == 0.0f) return 0.0f;
      if (qs.l2 == 0.0f || b.l2[rb] == 0.0f) return 1.0f;
      float dot = 0.0f;
      for (int64_t e = 0; e < cn; ++e) dot += q[ci[e]] * cv[e];
      const float c = std::min(std::max(dot / (qs.l2 * b.l2[rb]), -1.0f), 1.0f);
      return 1.0f - c;
    }
    case 2: {
      float s = qs.l1;
      for (int64_t e = 0; e < cn; ++e) {
        const float qv = q[ci[e]];
        s += std::fabs(qv - cv[e]) - std::fabs(qv);
      }
      return s;
    }
    case 3: {
      float m = 0.0f;
      int64_t pos = 0;
      for (int64_t f = 0; f < d; ++f) {
        float rv = 0.0f;
        if (pos < cn && ci[pos] == f) rv = cv[pos++];
        m = std::fmax(m, std::fabs(q[f] - rv));
      }
      return m;
    }
    case 4: {
      float s = qs.ps;
      for (int64_t e = 0; e < cn; ++e) {
        const float qv = q[ci[e]];
        s += std::pow(std::fabs(qv - cv[e]), p) - std::pow(std::fabs(qv), p);
      }
      return std::pow(std::fmax(s, 0.0f), 1.0f / p);
    }
    default: {
      float dot = 0.0f;
      for (int64_t e = 0; e < cn; ++e) dot += q[ci[e]] * cv[e];
      return corr_summ(d, qs.sum, qs.sq, b.sum[rb], b.sq[rb], dot);
    }
  }
}

int nthreads(int32_t n) { return n > 0 ? n : omp_get_max_threads(); }

}  // namespace

extern "C" {

/* exact_nearest_neighbors (sparse.rs:605-626): every row against every row,
 * the row itself at distance 0.0. */
void orc_knn_csr(int64_t n, int64_t n_cols, const int64_t* indptr, const int64_t* indices,
                 const float* data, int32_t k, int32_t metric, float p, int32_t n_threads,
                 int64_t* out_idx, float* out_dist) {
  const Csr x(n, n_cols, indptr, indices, data);
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads(n_threads))
  for (int64_t i = 0; i < n; ++i) {
    Top top(static_cast<size_t>(k));
    for (int64_t j = 0; j < n; ++j) top.push(j, i == j ? 0.0f : csr_pair(x, i, x, j, metric, p));
    top.store(i, k, out_idx, out_dist);
  }
}

/* exact_nearest_neighbors_sparse_query (sparse.rs:628-657): no self case. */
void orc_knn_csr_query(int64_t m, const int64_t* q_indptr, const int64_t* q_indices,
                       const float* q_data, int64_t n, const int64_t* r_indptr,
                       const int64_t* r_indices, const float* r_data, int64_t n_cols, int32_t k,
                       int32_t metric, float p, int32_t n_threads, int64_t* out_idx,
                       float* out_dist) {
  const Csr q(m, n_cols, q_indptr, q_indices, q_data);
  const Csr r(n, n_cols, r_indptr, r_indices, r_data);
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads(n_threads))
  for (int64_t i = 0; i < m; ++i) {
    Top top(static_cast<size_t>(k));
    for (int64_t j = 0; j < n; ++j) top.push(j, csr_pair(q, i, r, j, metric, p));
    top.store(i, k, out_idx, out_dist);
  }
}

/* exact_nearest_neighbors_dense_query (sparse.rs:753-900): dense rows q
 * [m x n_cols] against CSR rows. */
void orc_knn_dense_to_csr(const float* q, int64_t m, int64_t n, const int64_t* r_indptr,
                          const int64_t* r_indices, const float* r_data, int64_t n_cols,
                          int32_t k, int32_t metric, float p, int32_t n_threads,
                          int64_t* out_idx, float* out_dist) {
  const Csr r(n, n_cols, r_indptr, r_indices, r_data);
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads(n_threads))
  for (int64_t i = 0; i < m; ++i) {
    const float* qi = q + i * n_cols;
    const DenseQ qs = dense_summaries(qi, n_cols, metric, p);
    Top top(static_cast<size_t>(k));
    for (int64_t j = 0; j < n; ++j) top.push(j, dense_pair(qi, n_cols, qs, r, j, metric, p));
    top.store(i, k, out_idx, out_dist);
  }
}

/* SparseCsrMatrix::new (sparse.rs:20-90): 0 when the CSR is valid, else
 * 1 and the reference's message in msg (at most cap bytes). */
int32_t orc_csr_check(int64_t n_rows, int64_t n_cols, const int64_t* indptr, int64_t indptr_len,
                      const int64_t* indices, int64_t n_indices, const float* data,
                      int64_t n_data, char* msg, int64_t cap) {
  std::string e;
  if (n_cols == 0) {
    e = "csr n_cols must be >= 1";
  } else if (indptr_len != n_rows + 1) {
    e = "csr indptr length must be n_rows + 1 (got " + std::to_string(indptr_len) +
        ", expected " + std::to_string(n_rows + 1) + ")";
  } else if (indptr[0] != 0) {
    e = "csr indptr must start from 0";
  } else if (n_indices != n_data) {
    e = "csr indices/data lengths must match";
  } else if (indptr[n_rows] != n_indices) {
    e = "csr indptr last value (" + std::to_string(indptr[n_rows]) + ") must equal nnz (" +
        std::to_string(n_indices) + ")";
  } else {
    for (int64_t row = 0; row < n_rows && e.empty(); ++row) {
      if (indptr[row] > indptr[row + 1]) {
        e = "csr indptr must be non-decreasing, got indptr[" + std::to_string(row) +
            "]=" + std::to_string(indptr[row]) + " > indptr[" + std::to_string(row + 1) +
            "]=" + std::to_string(indptr[row + 1]);
        break;
      }
      for (int64_t q = indptr[row] + 1; q < indptr[row + 1]; ++q)
        if (indices[q - 1] >= indices[q]) {
          e = "csr row " + std::to_string(row) + " indices must be strictly increasing";
          break;
        }
      if (!e.empty()) break;
      for (int64_t q = indptr[row]; q < indptr[row + 1]; ++q)
        if (indices[q] < 0 || indices[q] >= n_cols) {
          e = "csr row " + std::to_string(row) + " has column index " +
              std::to_string(indices[q]) + " out of bounds for n_cols=" + std::to_string(n_cols);
          break;
        }
    }
    if (e.empty())
      for (int64_t q = 0; q < n_data; ++q)
        if (!std::isfinite(data[q])) {
          e = "csr data must be finite";
          break;
        }
  }
  if (e.empty()) return 0;
  if (msg && cap > 0) {
    const size_t len = std::min<size_t>(e.size(), static_cast<size_t>(cap - 1));
    std::memcpy(msg, e.data(), len);
    msg[len] = '\0';
  }
  return 1;
}

}  // extern "C"
