// Spectral embedding initialisation on the GPU.
#pragma once

#include "common.hpp"
#include "graph.hpp"

namespace umapx {

struct SpectralWorkspace {
  DevBuf<int32_t> label, flag, perm, comp_of, items, slot_of;
  DevBuf<double> deg, inv, gram, rmat, vmat, cent;
  DevBuf<float> y, z;  // the iterated block and its product, [rows][18] f32
  DevBuf<uint64_t> seeds;
  DevBuf<float> maxabs, scale, anchor;
  DevBuf<int32_t> fail;
};

struct SpectralReport {
  int64_t n_components = 0;
  int64_t n_iterative = 0;  // components solved by GPU subspace iteration
  int64_t n_dense = 0;      // components solved by a host dense eigen-solve
  int64_t n_random = 0;     // components laid out at random
  bool used_spectral = false;
};

// connected_components_from_edges (lib.rs:4394-4437): device labels, one per
// node, numbered by each component's smallest member.  Returns the count.
int64_t connected_components_device(const Graph& g, std::vector<int32_t>& labels_host,
                                    SpectralWorkspace& ws, hipStream_t s);

// initialize_embedding with init=spectral (lib.rs:3855-3870, 4355-4376,
// 4624-4764).  Returns false when the reference would fall back to random
// init; y (device, n x dim) is written only on success.
// For a sparse fit x is null and csr holds the rows (the component centroids
// of the multi-component layout are then summed from its nonzeros, as
// component_centroids_sparse does, lib.rs:4487-4516).
struct CsrDevice;
bool spectral_init_device(const float* x, int64_t n, int64_t d, int dim, const Graph& g,
                          uint64_t seed, float* y, SpectralWorkspace& ws, hipStream_t s,
                          SpectralReport* report, const CsrDevice* csr = nullptr);

}  // namespace umapx
