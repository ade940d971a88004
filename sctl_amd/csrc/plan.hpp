// The plan object behind sctl_amd_plan (include/sctl_amd.h), shared by the
// batched block kernels (batched.hip) and the near-field operator kernels
// (near.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "internal.hpp"

namespace sctl_amd {

// One work item of the batched Eval: a chunk of one block's targets.
struct Item {
  int32_t block;  // block index
  int32_t t_off;  // first target of the chunk, relative to the block
  int32_t t_cnt;  // targets in the chunk (<= 64*TPL)
  int32_t lg;     // log2 of the number of lane groups
};

// Wave tiles of the near-field operators, as a flattened list: tile i
// belongs to block tile_blk[i] and is the (i - tile_dsp[block])-th tile of
// that block.
struct TileList {
  int64_t ntiles = 0;
  int32_t* d_tile_blk = nullptr;
  int64_t* d_tile_dsp = nullptr;  // nblocks entries
};

template <class T> int dev_upload(T** dst, const std::vector<T>& v) {
  *dst = nullptr;
  if (v.empty()) return SCTL_AMD_SUCCESS;
  SCTL_HIP_TRY(hipMalloc(dst, sizeof(T) * v.size()));
  SCTL_HIP_TRY(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return SCTL_AMD_SUCCESS;
}

}  // namespace sctl_amd

struct sctl_amd_plan_s {
  int kernel = 0, precision = SCTL_AMD_FP64, newton = 2, K0 = 1, K1 = 1, ND = 0;
  int device = 0;
  int64_t nblocks = 0, nt_total = 0, ns_total = 0, interactions = 0, nonempty = 0;
  std::vector<int64_t> trg_dsp, trg_cnt, src_dsp, src_cnt;  // host copies
  int nitems = 0;
  sctl_amd::PlanShape shape;  // resolved knobs
  int grid = 0;               // persistent grid (workgroups)
  size_t dyn_lds = 0;         // dynamic LDS that enforces shape.occ
  bool has_geometry = false;
  double ctx = 0.0;     // wavenumber of the Helmholtz kernels (sctl_amd_plan_set_ctx)
  bool has_ctx = false;
  // Bumped whenever state a captured graph has baked in changes (work items,
  // launch shape and grid, context value): graphs record it at capture and
  // refuse to launch once it differs.
  uint64_t generation = 0;
  sctl_amd::Item* d_items = nullptr;
  int64_t *d_trg_dsp = nullptr, *d_src_dsp = nullptr;
  int32_t *d_src_cnt = nullptr, *d_trg_cnt = nullptr;
  int* d_counter = nullptr;
  double *d_xt = nullptr, *d_xs = nullptr, *d_xn = nullptr;

  // Near-field matrix layout (sctl_amd_plan_near_layout): per-block offsets
  // of the target-major block matrices in K_near (-1: no matrix), the number
  // of rows per target (K1, or 1 with the target-normal contraction) and the
  // length K_near must have.
  bool has_near = false;
  int near_dof = 1;
  bool near_ndot = false;
  int64_t near_len = 0;
  std::vector<int64_t> near_off;
  int64_t* d_near_off = nullptr;
  sctl_amd::TileList mat_tiles, gemv_tiles;

  void free_near() {
    void* ptrs[] = {d_near_off, mat_tiles.d_tile_blk, mat_tiles.d_tile_dsp, gemv_tiles.d_tile_blk,
                    gemv_tiles.d_tile_dsp};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    d_near_off = nullptr;
    mat_tiles = sctl_amd::TileList();
    gemv_tiles = sctl_amd::TileList();
    has_near = false;
  }

  ~sctl_amd_plan_s() {
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != device) (void)hipSetDevice(device);
    free_near();
    void* ptrs[] = {d_items, d_trg_dsp, d_src_dsp, d_src_cnt, d_trg_cnt, d_counter, d_xt, d_xs, d_xn};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
  }
};
