// Morton ordering of point sets, used by the mixed-precision path so that each
// source tile is spatially compact and its FP32 offsets from the tile centre
// are small.  30-bit keys (10 bits per axis) over the points' bounding cube,
// sorted with rocPRIM's stable radix sort: the permutation, and so the
// summation order, is deterministic.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "internal.hpp"

namespace sctl_amd {

namespace {

constexpr int BB_WG = 256;

// Per-workgroup min/max of a chunk of points: part[blk*6 + {0..2 min, 3..5 max}].
__global__ void __launch_bounds__(BB_WG) k_bbox_partial(const double* __restrict__ pts, int64_t n, int64_t per_blk,
                                                         double* __restrict__ part) {
  __shared__ double red[6][BB_WG];
  const int tid = threadIdx.x;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  const int64_t b = (int64_t)blockIdx.x * per_blk;
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

// Reduces nblk partial boxes (one workgroup) into part[0..5].
__global__ void __launch_bounds__(BB_WG) k_bbox_final(double* __restrict__ part, int nblk) {
  __shared__ double red[6][BB_WG];
  const int tid = threadIdx.x;
  double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = tid; b < nblk; b += BB_WG) {
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
  if (tid < 6) part[tid] = red[tid][0];
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {  // 10 bits -> every third bit
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ void __launch_bounds__(BB_WG) k_morton_keys(const double* __restrict__ pts, int64_t n,
                                                        const double* __restrict__ box, uint32_t* __restrict__ keys,
                                                        uint32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * BB_WG + threadIdx.x;
  if (i >= n) return;
  const double ext = fmax(fmax(box[3] - box[0], box[4] - box[1]), box[5] - box[2]);
  const double inv = ext > 0 ? 1024.0 / ext : 0.0;
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double u = (pts[i * 3 + k] - box[k]) * inv;
    u = u < 0 ? 0 : (u > 1023 ? 1023 : u);  // also maps NaN to 0
    q[k] = (uint32_t)u;
  }
  keys[i] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
  idx[i] = (uint32_t)i;
}

}  // namespace

int bounding_box(const double* a, int64_t na, const double* b, int64_t nb, double* box6, hipStream_t stream) {
  const int64_t per_blk = 8192;
  const int blk_a = (int)((na + per_blk - 1) / per_blk), blk_b = (int)((nb + per_blk - 1) / per_blk);
  SCTL_CHECK_ARG(blk_a + blk_b > 0, "bounding box of no points");
  Scratch part;
  SCTL_TRY(part.alloc(sizeof(double) * 6 * (blk_a + blk_b), stream));
  if (blk_a) {
    hipLaunchKernelGGL(k_bbox_partial, dim3(blk_a), dim3(BB_WG), 0, stream, a, na, per_blk, part.as<double>());
    SCTL_HIP_TRY(hipGetLastError());
  }
  if (blk_b) {
    hipLaunchKernelGGL(k_bbox_partial, dim3(blk_b), dim3(BB_WG), 0, stream, b, nb, per_blk,
                       part.as<double>() + 6 * blk_a);
    SCTL_HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(BB_WG), 0, stream, part.as<double>(), blk_a + blk_b);
  SCTL_HIP_TRY(hipGetLastError());
  SCTL_HIP_TRY(hipMemcpyAsync(box6, part.as<double>(), sizeof(double) * 6, hipMemcpyDeviceToDevice, stream));
  return SCTL_AMD_SUCCESS;
}

namespace {
struct MortonLayout {
  size_t part, keys_in, keys_out, idx, tmp, tmp_bytes, bytes;
};
MortonLayout morton_layout(int64_t n) {
  MortonLayout m;
  const int64_t nblk = (n + 8191) / 8192;
  Carve c;
  m.part = c.take(sizeof(double) * 6 * nblk);
  m.keys_in = c.take(sizeof(uint32_t) * n);
  m.keys_out = c.take(sizeof(uint32_t) * n);
  m.idx = c.take(sizeof(uint32_t) * n);
  m.tmp_bytes = 0;
  (void)rocprim::radix_sort_pairs(nullptr, m.tmp_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                  (uint32_t*)nullptr, (size_t)n, 0, 30);
  m.tmp = c.take(m.tmp_bytes);
  m.bytes = c.bytes;
  return m;
}
}  // namespace

size_t morton_order_bytes(int64_t n) { return n > 0 ? morton_layout(n).bytes : 0; }

int morton_order(const double* pts, int64_t n, uint32_t* perm, void* work, hipStream_t stream) {
  if (n <= 0) return SCTL_AMD_SUCCESS;
  SCTL_CHECK_ARG(work, "no workspace for the Morton order");
  const int64_t per_blk = 8192;
  const int nblk = (int)((n + per_blk - 1) / per_blk);
  const MortonLayout m = morton_layout(n);
  char* w = static_cast<char*>(work);
  double* part = reinterpret_cast<double*>(w + m.part);
  uint32_t* keys_in = reinterpret_cast<uint32_t*>(w + m.keys_in);
  uint32_t* keys_out = reinterpret_cast<uint32_t*>(w + m.keys_out);
  uint32_t* idx = reinterpret_cast<uint32_t*>(w + m.idx);
  hipLaunchKernelGGL(k_bbox_partial, dim3(nblk), dim3(BB_WG), 0, stream, pts, n, per_blk, part);
  SCTL_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(BB_WG), 0, stream, part, nblk);
  SCTL_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_morton_keys, dim3((unsigned)((n + BB_WG - 1) / BB_WG)), dim3(BB_WG), 0, stream, pts, n, part,
                     keys_in, idx);
  SCTL_HIP_TRY(hipGetLastError());
  size_t bytes = m.tmp_bytes;
  SCTL_HIP_TRY(rocprim::radix_sort_pairs(w + m.tmp, bytes, keys_in, keys_out, idx, perm, (size_t)n, 0, 30, stream));
  return SCTL_AMD_SUCCESS;
}

}  // namespace sctl_amd
