// Morton ordering of point sets, used by the mixed-precision path so that each
// source tile is spatially compact and its FP32 offsets from the tile centre
// are small.  30-bit keys (10 bits per axis) over each set's bounding cube,
// sorted with rocPRIM's stable radix sort: the permutation, and so the
// summation order, is deterministic.  Two sets (sources and targets) are
// ordered by one sort, the second set's keys carrying bit 30: stability
// makes each set's permutation the one a sort of that set alone would give,
// in 7 launches instead of twice as many.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "internal.hpp"

namespace sctl_amd {

namespace {

constexpr int BB_WG = 256;
constexpr int64_t BB_PER_BLK = 1024;  // points per bounding-box workgroup: 4 per thread

// Per-workgroup min/max of a chunk of points: part[blk*6 + {0..2 min, 3..5
// max}].  Workgroups [0, blk_a) take chunks of set a, the others of set b.
__global__ void __launch_bounds__(BB_WG) k_bbox_partial(const double* __restrict__ pa, int64_t na,
                                                         const double* __restrict__ pb, int64_t nb, int blk_a,
                                                         int64_t per_blk, double* __restrict__ part) {
  __shared__ double red[6][BB_WG];
  const int tid = threadIdx.x;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  const bool in_a = (int)blockIdx.x < blk_a;
  const double* __restrict__ pts = in_a ? pa : pb;
  const int64_t n = in_a ? na : nb;
  const int64_t b = (int64_t)(in_a ? blockIdx.x : blockIdx.x - blk_a) * per_blk;
  const int64_t e = b + per_blk < n ? b + per_blk : n;
  for (int64_t i = b + tid; i < e; i += BB_WG) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const double v = pts[i * 3 + k];
      lo[k] = fmin(lo[k], v);
      hi[k] = fmax(hi[k], v);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; k++) {
    red[k][tid] = lo[k];
    red[3 + k][tid] = hi[k];
  }
  __syncthreads();
  for (int w = BB_WG / 2; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        red[k][tid] = fmin(red[k][tid], red[k][tid + w]);
        red[3 + k][tid] = fmax(red[3 + k][tid], red[3 + k][tid + w]);
      }
    }
    __syncthreads();
  }
  if (tid < 6) part[blockIdx.x * 6 + tid] = red[tid][0];
}

// Workgroup g reduces the partial boxes [first_g, first_g + count_g) into
// out[6g .. 6g+5] (first_0 = 0, first_1 = count_0).
__global__ void __launch_bounds__(BB_WG) k_bbox_final(const double* __restrict__ part, int count0, int count1,
                                                       double* __restrict__ out) {
  __shared__ double red[6][BB_WG];
  const int tid = threadIdx.x;
  const int first = blockIdx.x == 0 ? 0 : count0, nblk = blockIdx.x == 0 ? count0 : count1;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = first + tid; b < first + nblk; b += BB_WG) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      v[k] = fmin(v[k], part[b * 6 + k]);
      v[3 + k] = fmax(v[3 + k], part[b * 6 + 3 + k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; k++) red[k][tid] = v[k];
  __syncthreads();
  for (int w = BB_WG / 2; w > 0; w >>= 1) {
    if (tid < w) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        red[k][tid] = fmin(red[k][tid], red[k][tid + w]);
        red[3 + k][tid] = fmax(red[3 + k][tid], red[3 + k][tid + w]);
      }
    }
    __syncthreads();
  }
  if (tid < 6) out[blockIdx.x * 6 + tid] = red[tid][0];
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {  // 10 bits -> every third bit
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Keys of the points of set a (i < na, over box[0..5]) and of set b (over
// box[6..11], with bit 30 set so that every key of b sorts after those of a);
// idx = i.
__global__ void __launch_bounds__(BB_WG) k_morton_keys(const double* __restrict__ pa, int64_t na,
                                                        const double* __restrict__ pb, int64_t nb,
                                                        const double* __restrict__ boxes, uint32_t* __restrict__ keys,
                                                        uint32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * BB_WG + threadIdx.x;
  if (i >= na + nb) return;
  const bool in_a = i < na;
  const double* __restrict__ box = boxes + (in_a ? 0 : 6);
  const double* __restrict__ p = in_a ? pa + i * 3 : pb + (i - na) * 3;
  const double ext = fmax(fmax(box[3] - box[0], box[4] - box[1]), box[5] - box[2]);
  const double inv = ext > 0 ? 1024.0 / ext : 0.0;
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double u = (p[k] - box[k]) * inv;
    u = fmin(fmax(u, 0.0), 1023.0);  // fmax returns 0 for a NaN u
    q[k] = (uint32_t)u;
  }
  keys[i] = (spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2)) | (in_a ? 0u : 1u << 30);
  idx[i] = (uint32_t)i;
}

// Splits the sorted indices into the two sets' permutations.
__global__ void __launch_bounds__(BB_WG) k_split_perm(const uint32_t* __restrict__ sorted, int64_t na, int64_t nb,
                                                       uint32_t* __restrict__ perm_a, uint32_t* __restrict__ perm_b) {
  const int64_t i = (int64_t)blockIdx.x * BB_WG + threadIdx.x;
  if (i >= na + nb) return;
  const uint32_t v = sorted[i];
  if (i < na) perm_a[i] = v;
  else perm_b[i - na] = v - (uint32_t)na;
}

}  // namespace

int bounding_box(const double* a, int64_t na, const double* b, int64_t nb, double* box6, hipStream_t stream) {
  const int64_t per_blk = BB_PER_BLK;
  const int blk_a = (int)((na + per_blk - 1) / per_blk), blk_b = (int)((nb + per_blk - 1) / per_blk);
  SCTL_CHECK_ARG(blk_a + blk_b > 0, "bounding box of no points");
  Scratch part;
  SCTL_TRY(part.alloc(sizeof(double) * 6 * (blk_a + blk_b), stream));
  hipLaunchKernelGGL(k_bbox_partial, dim3(blk_a + blk_b), dim3(BB_WG), 0, stream, a, na, b, nb, blk_a, per_blk,
                     part.as<double>());
  SCTL_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(BB_WG), 0, stream, part.as<double>(), blk_a + blk_b, 0, box6);
  SCTL_HIP_TRY(hipGetLastError());
  return SCTL_AMD_SUCCESS;
}

namespace {
struct MortonLayout {
  size_t part, boxes, keys_in, keys_out, idx, idx_out, tmp, tmp_bytes, bytes;
};
MortonLayout morton_layout(int64_t na, int64_t nb) {
  MortonLayout m;
  const int64_t n = na + nb, nblk = (na + BB_PER_BLK - 1) / BB_PER_BLK + (nb + BB_PER_BLK - 1) / BB_PER_BLK;
  Carve c;
  m.part = c.take(sizeof(double) * 6 * nblk);
  m.boxes = c.take(sizeof(double) * 12);
  m.keys_in = c.take(sizeof(uint32_t) * n);
  m.keys_out = c.take(sizeof(uint32_t) * n);
  m.idx = c.take(sizeof(uint32_t) * n);
  m.idx_out = c.take(sizeof(uint32_t) * n);
  m.tmp_bytes = 0;
  (void)rocprim::radix_sort_pairs(nullptr, m.tmp_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                  (uint32_t*)nullptr, (size_t)n, 0, 31);
  m.tmp = c.take(m.tmp_bytes);
  m.bytes = c.bytes;
  return m;
}
}  // namespace

size_t morton_order_bytes(int64_t na, int64_t nb) { return na + nb > 0 ? morton_layout(na, nb).bytes : 0; }

int morton_order(const double* a, int64_t na, uint32_t* perm_a, const double* b, int64_t nb, uint32_t* perm_b,
                 void* work, hipStream_t stream) {
  const int64_t n = na + nb;
  if (n <= 0) return SCTL_AMD_SUCCESS;
  SCTL_CHECK_ARG(work, "no workspace for the Morton order");
  SCTL_CHECK_ARG(n < ((int64_t)1 << 32), "too many points for 32-bit indices");
  const int64_t per_blk = BB_PER_BLK;
  const int blk_a = (int)((na + per_blk - 1) / per_blk), blk_b = (int)((nb + per_blk - 1) / per_blk);
  const MortonLayout m = morton_layout(na, nb);
  char* w = static_cast<char*>(work);
  double* part = reinterpret_cast<double*>(w + m.part);
  double* boxes = reinterpret_cast<double*>(w + m.boxes);
  uint32_t* keys_in = reinterpret_cast<uint32_t*>(w + m.keys_in);
  uint32_t* keys_out = reinterpret_cast<uint32_t*>(w + m.keys_out);
  uint32_t* idx = reinterpret_cast<uint32_t*>(w + m.idx);
  uint32_t* idx_out = reinterpret_cast<uint32_t*>(w + m.idx_out);
  hipLaunchKernelGGL(k_bbox_partial, dim3(blk_a + blk_b), dim3(BB_WG), 0, stream, a, na, b, nb, blk_a, per_blk, part);
  SCTL_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bbox_final, dim3(2), dim3(BB_WG), 0, stream, part, blk_a, blk_b, boxes);
  SCTL_HIP_TRY(hipGetLastError());
  const unsigned grid = (unsigned)((n + BB_WG - 1) / BB_WG);
  hipLaunchKernelGGL(k_morton_keys, dim3(grid), dim3(BB_WG), 0, stream, a, na, b, nb, boxes, keys_in, idx);
  SCTL_HIP_TRY(hipGetLastError());
  size_t bytes = m.tmp_bytes;
  SCTL_HIP_TRY(rocprim::radix_sort_pairs(w + m.tmp, bytes, keys_in, keys_out, idx, idx_out, (size_t)n, 0, 31, stream));
  hipLaunchKernelGGL(k_split_perm, dim3(grid), dim3(BB_WG), 0, stream, idx_out, na, nb, perm_a, perm_b);
  SCTL_HIP_TRY(hipGetLastError());
  return SCTL_AMD_SUCCESS;
}

}  // namespace sctl_amd
