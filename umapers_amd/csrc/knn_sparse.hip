// Exact kNN over CSR rows in the reference's sparse arithmetic
// (rust_umap/src/sparse.rs): no densifying, any number of columns.
//
// One wave per query row; lane l takes reference rows l, l + 64, ... and
// computes each distance as the reference does, by walking the two rows'
// nonzeros in column order (sparse_dot and the merged walks, sparse.rs:220-414,
// 416-603), or, for dense query rows, by walking the reference row's nonzeros
// against the dense row (sparse.rs:659-757).  The query row's nonzeros sit in
// LDS when they fit (else they are read through the caches); each lane
// streams its own reference row.  The wave keeps its k best (distance total
// order, index) pairs sorted across lanes 0..k-1: a candidate that beats the
// k-th is inserted by one ballot (its rank) and one lane shift, so the list
// needs no sorting and k may be up to 64.  All f32 arithmetic uses the
// reference's operations in its order without contraction, so every distance
// except minkowski's (|x|^p from the hardware log2 / exp2) equals the
// reference's bit for bit, including the expanded-form euclidean
// (|a|^2 + |b|^2 - 2 a.b), whose cancellation the tests pin.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <string>
#include <vector>

#include "knn.hpp"

namespace umapx {
namespace {

constexpr int kWaves = 4;    // query rows per block, one per wave
constexpr int kQCap = 1024;  // query nonzeros staged in LDS per wave

struct CsrArgs {
  const int64_t* ptr;
  const int32_t* ind;
  const float* val;
  const float* sq;
  const float* l2;
  const float* sum;
};

CsrArgs args_of(const CsrDevice& c) {
  return {c.ptr.get(), c.ind.get(), c.val.get(), c.sq.get(), c.l2.get(), c.sum.get()};
}

__device__ __forceinline__ float key_float(int key) {
  return __int_as_float(key ^ static_cast<int>(static_cast<unsigned>(key >> 31) >> 1));
}

// Rust's f32::clamp(-1, 1): NaN passes through.
__device__ __forceinline__ float clamp1(float c) { return c < -1.f ? -1.f : (c > 1.f ? 1.f : c); }

__device__ __forceinline__ float abs_pow(float v, float p) {
  const float a = fabsf(v);
  return a == 0.f ? 0.f : __builtin_amdgcn_exp2f(p * __builtin_amdgcn_logf(a));
}

// correlation_distance_from_summaries (lib.rs:2678-2698)
__device__ __forceinline__ float corr_summ(float nf, float sx, float qx, float sy, float qy,
                                           float dot) {
  const float cd = fsub(dot, fdiv(fmul(sx, sy), nf));
  const float xv = fmaxf(fsub(qx, fdiv(fmul(sx, sx), nf)), 0.f);
  const float yv = fmaxf(fsub(qy, fdiv(fmul(sy, sy), nf)), 0.f);
  const float xn = fsqrt(xv), yn = fsqrt(yv);
  if (xn == 0.f && yn == 0.f) return 0.f;
  if (xn == 0.f || yn == 0.f) return 1.f;
  return fsub(1.f, clamp1(fdiv(cd, fmul(xn, yn))));
}

// sparse_dot (sparse.rs:220-239)
__device__ __forceinline__ float merge_dot(const int32_t* qc, const float* qv, int64_t qn,
                                           const int32_t* rc, const float* rv, int64_t rn) {
  int64_t a = 0, b = 0;
  float dot = 0.f;
  while (a < qn && b < rn) {
    const int32_t ca = qc[a], cb = rc[b];
    if (ca < cb) {
      ++a;
    } else if (ca > cb) {
      ++b;
    } else {
      dot = fadd(dot, fmul(qv[a], rv[b]));
      ++a;
      ++b;
    }
  }
  return dot;
}

// The merged walks of sparse.rs:254-366: op(|a|), op(|b|) or op(|a - b|)
// per column of the union, in column order.
template <typename Op>
__device__ __forceinline__ void merge_abs(const int32_t* qc, const float* qv, int64_t qn,
                                          const int32_t* rc, const float* rv, int64_t rn, Op op) {
  int64_t a = 0, b = 0;
  while (a < qn && b < rn) {
    const int32_t ca = qc[a], cb = rc[b];
    if (ca < cb) {
      op(fabsf(qv[a]));
      ++a;
    } else if (ca > cb) {
      op(fabsf(rv[b]));
      ++b;
    } else {
      op(fabsf(fsub(qv[a], rv[b])));
      ++a;
      ++b;
    }
  }
  for (; a < qn; ++a) op(fabsf(qv[a]));
  for (; b < rn; ++b) op(fabsf(rv[b]));
}

// Row summaries of SparseCsrMatrix::new (sparse.rs:92-100), one row per thread.
__global__ void csr_summaries_kernel(const int64_t* __restrict__ ptr, const float* __restrict__ val,
                                     int64_t rows, float* __restrict__ sq, float* __restrict__ l2,
                                     float* __restrict__ sum) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  float s2 = 0.f, s1 = 0.f;
  for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) {
    const float v = val[e];
    s2 = fadd(s2, fmul(v, v));
    s1 = fadd(s1, v);
  }
  sq[i] = s2;
  sum[i] = s1;
  l2[i] = fsqrt(s2);
}

// Dense query summaries of exact_nearest_neighbors_dense_query
// (sparse.rs:774-830): squared norm, L1 norm, minkowski p-sum, sum, L2 norm.
__global__ void dense_summaries_kernel(const float* __restrict__ q, int64_t m, int64_t d,
                                       int metric, float p, float* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const float* r = q + i * d;
  float s2 = 0.f, l1 = 0.f, ps = 0.f, s1 = 0.f;
  for (int64_t f = 0; f < d; ++f) {
    const float v = r[f];
    s2 = fadd(s2, fmul(v, v));
    l1 = fadd(l1, fabsf(v));
    if (metric == 4) ps = fadd(ps, abs_pow(v, p));
    s1 = fadd(s1, v);
  }
  out[i * 5 + 0] = s2;
  out[i * 5 + 1] = l1;
  out[i * 5 + 2] = ps;
  out[i * 5 + 3] = s1;
  out[i * 5 + 4] = fsqrt(s2);
}

__device__ __forceinline__ int64_t readlane64(int64_t v, int l) {
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(v & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(v >> 32), l);
  return (static_cast<int64_t>(hi) << 32) | static_cast<uint32_t>(lo);
}

// MODE 0: the fit's own rows (row i at 0.0 from itself); 1: CSR queries;
// 2: dense queries (qd [m x cols], qs their 5 summaries per row).
template <int METRIC, int MODE>
__global__ __launch_bounds__(64 * kWaves) void csr_knn_kernel(
    CsrArgs q, const float* __restrict__ qd, const float* __restrict__ qs, int64_t m, CsrArgs r,
    int64_t n, int64_t cols, int k, float p, int32_t* __restrict__ out_idx,
    float* __restrict__ out_dist) {
  __shared__ int32_t s_col[kWaves][kQCap];
  __shared__ float s_val[kWaves][kQCap];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t qi = static_cast<int64_t>(blockIdx.x) * kWaves + wv;
  if (qi >= m) return;  // whole wave; the kernel has no block barriers
  const int32_t* qc = nullptr;
  const float* qv = nullptr;
  int64_t qn = 0;
  float q_sq = 0.f, q_l2 = 0.f, q_sum = 0.f, q_l1 = 0.f, q_ps = 0.f;
  if (MODE < 2) {
    const int64_t q0 = q.ptr[qi];
    qn = q.ptr[qi + 1] - q0;
    if (qn <= kQCap) {
      for (int64_t e = lane; e < qn; e += 64) {
        s_col[wv][e] = q.ind[q0 + e];
        s_val[wv][e] = q.val[q0 + e];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      qc = s_col[wv];
      qv = s_val[wv];
    } else {
      qc = q.ind + q0;
      qv = q.val + q0;
    }
    q_sq = q.sq[qi];
    q_l2 = q.l2[qi];
    q_sum = q.sum[qi];
  } else {
    q_sq = qs[qi * 5 + 0];
    q_l1 = qs[qi * 5 + 1];
    q_ps = qs[qi * 5 + 2];
    q_sum = qs[qi * 5 + 3];
    q_l2 = qs[qi * 5 + 4];
  }
  const float* qrow = MODE == 2 ? qd + qi * cols : nullptr;
  const float nf = static_cast<float>(cols);
  int64_t entry = INT64_MAX;  // lane l < k: the l-th best (key << 32 | index)
  int64_t thr = INT64_MAX;
  for (int64_t j0 = 0; j0 < n; j0 += 64) {
    const int64_t j = j0 + lane;
    int64_t cand = INT64_MAX;
    if (j < n) {
      const int64_t r0 = r.ptr[j];
      const int64_t rn = r.ptr[j + 1] - r0;
      const int32_t* rc = r.ind + r0;
      const float* rv = r.val + r0;
      float dist;
      if (MODE == 0 && j == qi) {
        dist = 0.f;
      } else if (MODE < 2) {
        if (METRIC == 0) {  // euclidean_distance_from_dot
          const float dot = merge_dot(qc, qv, qn, rc, rv, rn);
          dist = fsqrt(fmaxf(fsub(fadd(q_sq, r.sq[j]), fmul(2.f, dot)), 0.f));
        } else if (METRIC == 1) {  // sparse_row_cosine_distance
          const float nr = r.l2[j];
          if (q_l2 == 0.f && nr == 0.f) {
            dist = 0.f;
          } else if (q_l2 == 0.f || nr == 0.f) {
            dist = 1.f;
          } else {
            const float dot = merge_dot(qc, qv, qn, rc, rv, rn);
            dist = fsub(1.f, clamp1(fdiv(dot, fmul(q_l2, nr))));
          }
        } else if (METRIC == 2) {
          float acc = 0.f;
          merge_abs(qc, qv, qn, rc, rv, rn, [&](float v) { acc = fadd(acc, v); });
          dist = acc;
        } else if (METRIC == 3) {
          float acc = 0.f;
          merge_abs(qc, qv, qn, rc, rv, rn, [&](float v) { acc = fmaxf(acc, v); });
          dist = acc;
        } else if (METRIC == 4) {
          float acc = 0.f;
          merge_abs(qc, qv, qn, rc, rv, rn, [&](float v) { acc = fadd(acc, abs_pow(v, p)); });
          dist = powf(acc, fdiv(1.f, p));
        } else {
          const float dot = merge_dot(qc, qv, qn, rc, rv, rn);
          dist = corr_summ(nf, q_sum, q_sq, r.sum[j], r.sq[j], dot);
        }
      } else {  // dense query row against the CSR row (sparse.rs:659-757, 836-900)
        if (METRIC == 0 || METRIC == 1 || METRIC == 5) {
          float dot = 0.f;
          if (METRIC != 1 || (q_l2 != 0.f && r.l2[j] != 0.f))
            for (int64_t e = 0; e < rn; ++e) dot = fadd(dot, fmul(qrow[rc[e]], rv[e]));
          if (METRIC == 0) {
            dist = fsqrt(fmaxf(fsub(fadd(q_sq, r.sq[j]), fmul(2.f, dot)), 0.f));
          } else if (METRIC == 1) {
            const float nr = r.l2[j];
            if (q_l2 == 0.f && nr == 0.f)
              dist = 0.f;
            else if (q_l2 == 0.f || nr == 0.f)
              dist = 1.f;
            else
              dist = fsub(1.f, clamp1(fdiv(dot, fmul(q_l2, nr))));
          } else {
            dist = corr_summ(nf, q_sum, q_sq, r.sum[j], r.sq[j], dot);
          }
        } else if (METRIC == 2) {
          float acc = q_l1;
          for (int64_t e = 0; e < rn; ++e) {
            const float x = qrow[rc[e]];
            acc = fadd(acc, fsub(fabsf(fsub(x, rv[e])), fabsf(x)));
          }
          dist = acc;
        } else if (METRIC == 3) {
          float acc = 0.f;
          int64_t pos = 0;
          for (int64_t f = 0; f < cols; ++f) {
            float y = 0.f;
            if (pos < rn && rc[pos] == f) y = rv[pos++];
            acc = fmaxf(acc, fabsf(fsub(qrow[f], y)));
          }
          dist = acc;
        } else {
          float acc = q_ps;
          for (int64_t e = 0; e < rn; ++e) {
            const float x = qrow[rc[e]];
            acc = fadd(acc, fsub(abs_pow(fsub(x, rv[e]), p), abs_pow(x, p)));
          }
          dist = powf(fmaxf(acc, 0.f), fdiv(1.f, p));
        }
      }
      // (signed key, index) as one signed 64-bit value: key * 2^32 + index
      cand = static_cast<int64_t>(static_cast<uint64_t>(static_cast<uint32_t>(total_key(dist)))
                                  << 32) |
             static_cast<uint32_t>(j);
    }
    // insert the lanes that beat the current k-th, one at a time
    uint64_t pass = __ballot(cand < thr);
    while (pass) {
      const int src = __ffsll(static_cast<unsigned long long>(pass)) - 1;
      pass &= pass - 1;
      const int64_t c = readlane64(cand, src);
      if (c >= thr) continue;  // wave-uniform
      const int pos = __popcll(__ballot(lane < k && entry < c));
      const int64_t up = __shfl_up(entry, 1);
      if (lane < k) {
        if (lane > pos)
          entry = up;
        else if (lane == pos)
          entry = c;
      }
      thr = readlane64(entry, k - 1);
    }
  }
  if (lane < k) {
    const bool have = entry != INT64_MAX;
    out_idx[qi * k + lane] = have ? static_cast<int32_t>(entry & 0xffffffff) : -1;
    out_dist[qi * k + lane] = have ? key_float(static_cast<int>(entry >> 32)) : INFINITY;
  }
}

template <int MODE>
void launch(int metric, CsrArgs q, const float* qd, const float* qs, int64_t m, CsrArgs r,
            int64_t n, int64_t cols, int k, float p, int32_t* idx, float* dist, hipStream_t s) {
  const dim3 grid(static_cast<unsigned>((m + kWaves - 1) / kWaves)), block(64 * kWaves);
#define UMAPX_CSR_CASE(M)                                                                     \
  case M:                                                                                     \
    hipLaunchKernelGGL((csr_knn_kernel<M, MODE>), grid, block, 0, s, q, qd, qs, m, r, n, cols, \
                       k, p, idx, dist);                                                      \
    break;
  switch (metric) {
    UMAPX_CSR_CASE(0)
    UMAPX_CSR_CASE(1)
    UMAPX_CSR_CASE(2)
    UMAPX_CSR_CASE(3)
    UMAPX_CSR_CASE(4)
    UMAPX_CSR_CASE(5)
    default:
      throw_invalid("metric is not available for sparse input");
  }
#undef UMAPX_CSR_CASE
  KERNEL_CHECK();
}

void check_knn_args(int64_t n, int k, int metric) {
  if (metric < 0 || metric > 5) throw_invalid("metric is not available for sparse input");
  if (k < 1 || k > 64) throw_invalid("n_neighbors must be <= 64 for sparse input on the GPU path");
  if (k > n) throw_invalid("n_neighbors must not exceed the number of reference rows");
  if (n >= (int64_t{1} << 31)) throw_invalid("sparse kNN supports fewer than 2^31 rows");
}

}  // namespace

void csr_validate_host(int64_t rows, int64_t cols, const int64_t* indptr, int64_t indptr_len,
                       const int64_t* indices, const float* data, int64_t nnz) {
  if (cols < 1) throw_invalid("csr n_cols must be >= 1");
  if (indptr_len != rows + 1)
    throw_invalid("csr indptr length must be n_rows + 1 (got " + std::to_string(indptr_len) +
                  ", expected " + std::to_string(rows + 1) + ")");
  if (indptr[0] != 0) throw_invalid("csr indptr must start from 0");
  if (indptr[rows] != nnz)
    throw_invalid("csr indptr last value (" + std::to_string(indptr[rows]) +
                  ") must equal nnz (" + std::to_string(nnz) + ")");
  for (int64_t row = 0; row < rows; ++row) {
    if (indptr[row] > indptr[row + 1])
      throw_invalid("csr indptr must be non-decreasing, got indptr[" + std::to_string(row) +
                    "]=" + std::to_string(indptr[row]) + " > indptr[" + std::to_string(row + 1) +
                    "]=" + std::to_string(indptr[row + 1]));
    for (int64_t e = indptr[row] + 1; e < indptr[row + 1]; ++e)
      if (indices[e - 1] >= indices[e])
        throw_invalid("csr row " + std::to_string(row) + " indices must be strictly increasing");
    for (int64_t e = indptr[row]; e < indptr[row + 1]; ++e)
      if (indices[e] < 0 || indices[e] >= cols)
        throw_invalid("csr row " + std::to_string(row) + " has column index " +
                      std::to_string(indices[e]) + " out of bounds for n_cols=" +
                      std::to_string(cols));
  }
  for (int64_t e = 0; e < nnz; ++e)
    if (!std::isfinite(data[e])) throw_invalid("csr data must be finite");
  if (cols > INT32_MAX) throw_invalid("sparse input on the GPU path supports fewer than 2^31 columns");
}

void csr_upload(CsrDevice& c, int64_t rows, int64_t cols, const int64_t* indptr,
                const int64_t* indices, const float* data, hipStream_t s) {
  const int64_t nnz = indptr[rows];
  c.rows = rows;
  c.cols = cols;
  c.nnz = nnz;
  std::vector<int32_t> ind32(static_cast<size_t>(nnz));
  for (int64_t e = 0; e < nnz; ++e) ind32[e] = static_cast<int32_t>(indices[e]);
  c.ptr.alloc(rows + 1);
  c.ind.alloc(std::max<int64_t>(nnz, 1));
  c.val.alloc(std::max<int64_t>(nnz, 1));
  c.sq.alloc(std::max<int64_t>(rows, 1));
  c.l2.alloc(std::max<int64_t>(rows, 1));
  c.sum.alloc(std::max<int64_t>(rows, 1));
  HIP_CHECK(hipMemcpyAsync(c.ptr.get(), indptr, 8 * (rows + 1), hipMemcpyHostToDevice, s));
  if (nnz > 0) {
    HIP_CHECK(hipMemcpyAsync(c.ind.get(), ind32.data(), 4 * nnz, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipMemcpyAsync(c.val.get(), data, 4 * nnz, hipMemcpyHostToDevice, s));
  }
  if (rows > 0) {
    hipLaunchKernelGGL(csr_summaries_kernel, dim3(blocks_for(rows, 256)), dim3(256), 0, s,
                       c.ptr.get(), c.val.get(), rows, c.sq.get(), c.l2.get(), c.sum.get());
    KERNEL_CHECK();
  }
  HIP_CHECK(hipStreamSynchronize(s));  // ind32 is freed on return
}

void csr_knn_device(const CsrDevice& q, const CsrDevice& r, bool self, int k, int metric,
                    float p, int32_t* out_idx, float* out_dist, hipStream_t s) {
  check_knn_args(r.rows, k, metric);
  if (q.cols != r.cols) throw Error(kFeatureMismatch, "feature mismatch: expected " +
                                    std::to_string(r.cols) + ", got " + std::to_string(q.cols),
                                    -1, r.cols, q.cols);
  if (q.rows == 0) return;
  if (self)
    launch<0>(metric, args_of(q), nullptr, nullptr, q.rows, args_of(r), r.rows, r.cols, k, p,
              out_idx, out_dist, s);
  else
    launch<1>(metric, args_of(q), nullptr, nullptr, q.rows, args_of(r), r.rows, r.cols, k, p,
              out_idx, out_dist, s);
}

void dense_to_csr_knn_device(const float* q, int64_t m, const CsrDevice& r, int k, int metric,
                             float p, int32_t* out_idx, float* out_dist, hipStream_t s) {
  check_knn_args(r.rows, k, metric);
  if (m == 0) return;
  DevBuf<float> qs(static_cast<size_t>(m) * 5);
  hipLaunchKernelGGL(dense_summaries_kernel, dim3(blocks_for(m, 256)), dim3(256), 0, s, q, m,
                     r.cols, metric, p, qs.get());
  KERNEL_CHECK();
  const CsrArgs none{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  launch<2>(metric, none, q, qs.get(), m, args_of(r), r.rows, r.cols, k, p, out_idx, out_dist, s);
  HIP_CHECK(hipStreamSynchronize(s));  // qs is released on return
}

}  // namespace umapx
