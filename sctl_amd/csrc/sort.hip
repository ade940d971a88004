// Hilbert-curve ordering of point sets, used by the mixed-precision path so
// that each source tile and each wave of targets is spatially compact and its
// FP32 offsets from the tile centre are small.  A point's key is its cell's
// position along a 3-D Hilbert curve over the set's bounding cube.  Two sets
// (sources and targets) are ordered by one sort, the second set's keys
// carrying the top key bit: stability makes each set's permutation the one a
// sort of that set alone would give.  The permutation, and so the summation
// order, is deterministic.
//  - Up to SMALL_N points in all: 15-bit keys (a 32^3 grid), ordered by a
//    two-pass least-significant-digit radix sort written here.  Each pass
//    ranks the 8-bit digit within 1024-point chunks by ballots into per-wave
//    LDS counters, and each chunk's scatter scans the chunk counts itself:
//    five launches in all (bounding boxes, then rank and scatter per digit).
//    A point set this small holds at most two points per cell of the grid on
//    average, so tiles of 64 and 256 points take whole cells and come out as
//    compact as with finer keys.
//  - More points: 30-bit keys (a 1024^3 grid), sorted with rocPRIM's stable
//    radix sort.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "internal.hpp"
#include "wave.hpp"

namespace sctl_amd {

namespace {

constexpr int BB_WG = 256;
constexpr int64_t BB_PER_BLK = 1024;  // points per bounding-box workgroup: 4 per thread

// Per-workgroup min/max of a chunk of BB_PER_BLK points: part[blk*6 + {0..2
// min, 3..5 max}].  Workgroups [0, blk_a) take chunks of set a, the others of
// set b.  Each thread reads its four points at once; waves reduce by
// exchanges, then one wave over the four waves.
__global__ void __launch_bounds__(BB_WG) k_bbox_partial(const double* __restrict__ pa, int64_t na,
                                                         const double* __restrict__ pb, int64_t nb, int blk_a,
                                                         double* __restrict__ part) {
  constexpr int PER = (int)(BB_PER_BLK / BB_WG);
  __shared__ double red[BB_WG / 64][6];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const bool in_a = (int)blockIdx.x < blk_a;
  const double* __restrict__ pts = in_a ? pa : pb;
  const int64_t n = in_a ? na : nb;
  const int64_t b = (int64_t)(in_a ? blockIdx.x : blockIdx.x - blk_a) * BB_PER_BLK;
  double x[PER][3];
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int64_t i = b + j * BB_WG + tid;
#pragma unroll
    for (int k = 0; k < 3; k++) x[j][k] = i < n ? pts[i * 3 + k] : NAN;  // fmin/fmax skip NaN
  }
  double v[6];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    v[k] = INFINITY;
    v[3 + k] = -INFINITY;
#pragma unroll
    for (int j = 0; j < PER; j++) {
      v[k] = fmin(v[k], x[j][k]);
      v[3 + k] = fmax(v[3 + k], x[j][k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 6; k++) {
    v[k] = k < 3 ? wave_min(v[k]) : wave_max(v[k]);
    if (lane == 0) red[wv][k] = v[k];
  }
  __syncthreads();
  if (tid < 6) {
    double r = red[0][tid];
#pragma unroll
    for (int w = 1; w < BB_WG / 64; w++) r = tid < 3 ? fmin(r, red[w][tid]) : fmax(r, red[w][tid]);
    part[blockIdx.x * 6 + tid] = r;
  }
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

__device__ __forceinline__ uint32_t spread5(uint32_t v) {  // 5 bits -> every third bit
  v &= 0x1fu;
  v = (v | (v << 8)) & 0x0000f00fu;
  v = (v | (v << 4)) & 0x000c30c3u;
  v = (v | (v << 2)) & 0x00249249u;
  return v;
}
__device__ __forceinline__ uint32_t spread10(uint32_t v) {  // 10 bits -> every third bit
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000ffu;
  v = (v | (v << 8)) & 0x0300f00fu;
  v = (v | (v << 4)) & 0x030c30c3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// Position of cell (x, y, z) of the 2^BITS grid along a 3-D Hilbert curve
// (3 BITS bits): J. Skilling's transform of the axes to the curve's
// "transposed" index (Programming the Hilbert curve, AIP Conf. Proc. 707,
// 2004), without branches, then the transposed bits interleaved with x's the
// most significant of each triple.  Consecutive positions are face-adjacent
// cells, so runs of the order are more compact than Morton runs (a 64-point
// run of a uniform cube of 20 000 points: mean radius 0.141 against 0.181).
template <int BITS> __device__ __forceinline__ uint32_t hilbert3(uint32_t x, uint32_t y, uint32_t z) {
  uint32_t X[3] = {x, y, z};
#pragma unroll
  for (uint32_t Q = 1u << (BITS - 1); Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const bool hit = (X[i] & Q) != 0;
      const uint32_t t = (X[0] ^ X[i]) & P;  // 0 for i == 0
      X[0] ^= hit ? P : t;
      if (i > 0) X[i] ^= hit ? 0u : t;
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  uint32_t t = 0;
#pragma unroll
  for (uint32_t Q = 1u << (BITS - 1); Q > 1; Q >>= 1) t ^= (X[2] & Q) ? Q - 1 : 0u;
#pragma unroll
  for (int i = 0; i < 3; i++) X[i] ^= t;
  if constexpr (BITS == 10) return (spread10(X[0]) << 2) | (spread10(X[1]) << 1) | spread10(X[2]);
  else return (spread5(X[0]) << 2) | (spread5(X[1]) << 1) | spread5(X[2]);
}

// Keys of the points of set a (i < na, over box[0..5]) and of set b (over
// box[6..11], with bit 30 set so that every key of b sorts after those of a);
// idx = i.
__global__ void __launch_bounds__(BB_WG) k_hilbert_keys(const double* __restrict__ pa, int64_t na,
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
  keys[i] = hilbert3<10>(q[0], q[1], q[2]) | (in_a ? 0u : 1u << 30);
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

// ---- Small sets: two-pass radix sort of 16-bit keys ------------------------
constexpr int64_t SMALL_N = 65536;  // points of both sets together
constexpr int RX_WG = 256;          // four waves
constexpr int RX_L = 1024;          // points per chunk: 256 per wave, 4 ballot steps
constexpr int RX_D = 256;           // buckets of an 8-bit digit
static_assert(SMALL_N / RX_L <= 64, "chunk counts of a digit are read in one go");

// Ranks of one wave's 256 consecutive points of a chunk by their BITS-bit
// digit, in point order: rank[s] = points of the same digit before point s of
// this wave.  h (LDS, this wave's counters) holds the wave's digit counts
// afterwards.  Lanes of equal digits are found with BITS ballots per step.
template <int BITS, class C>
__device__ __forceinline__ void rank_wave(const uint32_t (&dig)[4], const bool (&ok)[4], C* h, uint32_t (&rank)[4]) {
  const int lane = threadIdx.x & 63;
  const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int s = 0; s < 4; s++) {
    uint64_t eq = __builtin_amdgcn_ballot_w64(ok[s]);
#pragma unroll
    for (int b = 0; b < BITS; b++) {
      const uint64_t bal = __builtin_amdgcn_ballot_w64(ok[s] && ((dig[s] >> b) & 1u));
      eq &= ((dig[s] >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t before = ok[s] ? (uint32_t)h[dig[s]] : 0u;
    rank[s] = before + (uint32_t)__popcll(eq & lt);
    if (ok[s] && (eq & lt) == 0) h[dig[s]] = (C)(before + (uint32_t)__popcll(eq));
  }
}

// Chunk `chunk`'s counts of the 256 digits: ranks of the four waves made
// chunk-wide, counts written to hist[chunk * 256 + digit].  h: the four
// waves' counters (LDS).
__device__ __forceinline__ void rank_chunk(const uint32_t (&dig)[4], const bool (&ok)[4], uint32_t (*h)[RX_D],
                                           uint32_t (&rank)[4], uint32_t* __restrict__ hist, int chunk) {
  const int tid = threadIdx.x, w = tid >> 6;
#pragma unroll
  for (int k = 0; k < 4; k++) h[k][tid] = 0u;
  __syncthreads();
  rank_wave<8>(dig, ok, h[w], rank);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 4; s++) {
    uint32_t add = 0;
    for (int k = 0; k < w; k++) add += h[k][dig[s]];
    rank[s] += add;
  }
  hist[(int64_t)chunk * RX_D + tid] = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
}

// Start of chunk `chunk`'s points of each digit in the pass's output, for a
// workgroup of 256 threads (thread d = digit d): offs[d] = (points of smaller
// digits) + (points of digit d in earlier chunks).  Thread d reads the counts
// of all chunks at once (at most 64); the digit totals are scanned by
// exchanges within each wave and across the four waves.
__device__ __forceinline__ void chunk_offsets(const uint32_t* __restrict__ hist, int nchunk, int chunk,
                                              uint32_t* offs) {
  constexpr int MAXC = (int)(SMALL_N / RX_L);
  __shared__ uint32_t wsum[RX_D / 64];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint32_t v[MAXC];
#pragma unroll
  for (int c = 0; c < MAXC; c++) v[c] = c < nchunk ? hist[c * RX_D + tid] : 0u;
  uint32_t tot = 0, pre = 0;
#pragma unroll
  for (int c = 0; c < MAXC; c++) {
    tot += v[c];
    pre += c < chunk ? v[c] : 0u;
  }
  uint32_t inc = tot;  // inclusive scan of the totals within the wave
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    const uint32_t o = __shfl_up(inc, m, 64);
    if (lane >= m) inc += o;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  uint32_t base = 0;
#pragma unroll
  for (int k = 0; k < RX_D / 64; k++) base += k < w ? wsum[k] : 0u;
  offs[tid] = base + inc - tot + pre;
  __syncthreads();
}

// Hilbert keys of the four points per thread of a chunk (256 threads), 5 bits
// per axis: set a over box a, set b over box b with bit 15 set.  The boxes are
// reduced from the bounding-box partials [0, blk_a) of set a and
// [blk_a, blk_a + blk_b) of set b (at most 64 each) into box (LDS; box[s][6] =
// 32 / extent).  Point s of this thread: c0 + 64 s + lane.
__device__ __forceinline__ void chunk_keys(const double* __restrict__ pa, int64_t na, const double* __restrict__ pb,
                                           int64_t nb, const double* __restrict__ part, int blk_a, int blk_b,
                                           double (*box)[7], int64_t c0, uint32_t (&key)[4], bool (&ok)[4]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t n = na + nb;
  double x[4][3];  // the points, read before the boxes are known
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int64_t i = c0 + s * 64 + lane;
    ok[s] = i < n;
    const double* __restrict__ p = i < na ? pa + i * 3 : pb + (i - na) * 3;
#pragma unroll
    for (int k = 0; k < 3; k++) x[s][k] = ok[s] ? p[k] : 0.0;
  }
  if (w < 2) {  // wave 0: set a's box, wave 1: set b's
    const int first = w == 0 ? 0 : blk_a, cnt = w == 0 ? blk_a : blk_b;
    double v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = lane < cnt ? part[(first + lane) * 6 + k] : (k < 3 ? INFINITY : -INFINITY);
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = k < 3 ? wave_min(v[k]) : wave_max(v[k]);
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < 6; k++) box[w][k] = v[k];
      const double ext = fmax(fmax(v[3] - v[0], v[4] - v[1]), v[5] - v[2]);
      box[w][6] = ext > 0 ? 32.0 / ext : 0.0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int64_t i = c0 + s * 64 + lane;
    key[s] = 0;
    if (ok[s]) {
      const bool in_a = i < na;
      const double* __restrict__ bx = box[in_a ? 0 : 1];
      uint32_t q[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double u = (x[s][k] - bx[k]) * bx[6];
        u = fmin(fmax(u, 0.0), 31.0);  // fmax returns 0 for a NaN u
        q[k] = (uint32_t)u;
      }
      key[s] = hilbert3<5>(q[0], q[1], q[2]) | (in_a ? 0u : 1u << 15);
    }
  }
}

// Pass 1: keys (5 bits per axis) and their ranks by the low digit.
// kr[i] = key << 16 | rank.
__global__ void __launch_bounds__(RX_WG) k_rx_rank_low(const double* __restrict__ pa, int64_t na,
                                                        const double* __restrict__ pb, int64_t nb,
                                                        const double* __restrict__ part, int blk_a, int blk_b,
                                                        uint32_t* __restrict__ kr, uint32_t* __restrict__ hist) {
  __shared__ double box[2][7];
  __shared__ uint32_t h[4][RX_D];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * RX_L + w * (RX_L / 4);
  uint32_t key[4], dig[4], rank[4];
  bool ok[4];
  chunk_keys(pa, na, pb, nb, part, blk_a, blk_b, box, c0, key, ok);
#pragma unroll
  for (int s = 0; s < 4; s++) dig[s] = key[s] & 0xffu;
  rank_chunk(dig, ok, h, rank, hist, blockIdx.x);
#pragma unroll
  for (int s = 0; s < 4; s++)
    if (ok[s]) kr[c0 + s * 64 + lane] = key[s] << 16 | rank[s];
}

// Pass 1 scatter: point i goes to the low-digit order, tkey/tidx = key, i.  The chunk's four points per thread are read before the
// chunk counts.
__global__ void __launch_bounds__(RX_WG) k_rx_scatter_low(const uint32_t* __restrict__ kr, int64_t n,
                                                           const uint32_t* __restrict__ hist, int nchunk,
                                                           uint32_t* __restrict__ tkey, uint32_t* __restrict__ tidx) {
  __shared__ uint32_t offs[RX_D];
  constexpr int PER = RX_L / RX_WG;
  const int64_t i0 = (int64_t)blockIdx.x * RX_L + threadIdx.x;
  uint32_t v[PER];
#pragma unroll
  for (int j = 0; j < PER; j++) v[j] = i0 + j * RX_WG < n ? kr[i0 + j * RX_WG] : 0u;
  chunk_offsets(hist, nchunk, blockIdx.x, offs);
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int64_t i = i0 + j * RX_WG;
    if (i < n) {
      const uint32_t key = v[j] >> 16, pos = offs[key & 0xffu] + (v[j] & 0xffffu);
      tkey[pos] = key;
      tidx[pos] = (uint32_t)i;
    }
  }
}

// Pass 2: ranks of the points in low-digit order by their high digit.
__global__ void __launch_bounds__(RX_WG) k_rx_rank_high(const uint32_t* __restrict__ tkey, int64_t n,
                                                         uint32_t* __restrict__ rk, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[4][RX_D];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t c0 = (int64_t)blockIdx.x * RX_L + w * (RX_L / 4);
  uint32_t dig[4], rank[4];
  bool ok[4];
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const int64_t p = c0 + s * 64 + lane;
    ok[s] = p < n;
    dig[s] = ok[s] ? tkey[p] >> 8 : 0u;
  }
  rank_chunk(dig, ok, h, rank, hist, blockIdx.x);
#pragma unroll
  for (int s = 0; s < 4; s++)
    if (ok[s]) rk[c0 + s * 64 + lane] = rank[s];
}

// Pass 2 scatter into the two permutations: positions [0, na) hold set a
// (high digits below 128), the rest set b.
__global__ void __launch_bounds__(RX_WG) k_rx_scatter_high(const uint32_t* __restrict__ tkey,
                                                            const uint32_t* __restrict__ tidx,
                                                            const uint32_t* __restrict__ rk, int64_t na, int64_t nb,
                                                            const uint32_t* __restrict__ hist, int nchunk,
                                                            uint32_t* __restrict__ perm_a,
                                                            uint32_t* __restrict__ perm_b) {
  __shared__ uint32_t offs[RX_D];
  constexpr int PER = RX_L / RX_WG;
  const int64_t n = na + nb, p0 = (int64_t)blockIdx.x * RX_L + threadIdx.x;
  uint32_t key[PER], idx[PER], rank[PER];
#pragma unroll
  for (int j = 0; j < PER; j++) {
    const int64_t p = p0 + j * RX_WG;
    key[j] = p < n ? tkey[p] : 0u;
    idx[j] = p < n ? tidx[p] : 0u;
    rank[j] = p < n ? rk[p] : 0u;
  }
  chunk_offsets(hist, nchunk, blockIdx.x, offs);
#pragma unroll
  for (int j = 0; j < PER; j++) {
    if (p0 + j * RX_WG < n) {
      const int64_t pos = offs[key[j] >> 8] + rank[j];
      if (pos < na) perm_a[pos] = idx[j];
      else perm_b[pos - na] = idx[j] - (uint32_t)na;
    }
  }
}

}  // namespace

int bounding_box(const double* a, int64_t na, const double* b, int64_t nb, double* box6, hipStream_t stream) {
  const int64_t per_blk = BB_PER_BLK;
  const int blk_a = (int)((na + per_blk - 1) / per_blk), blk_b = (int)((nb + per_blk - 1) / per_blk);
  SCTL_CHECK_ARG(blk_a + blk_b > 0, "bounding box of no points");
  Scratch part;
  SCTL_TRY(part.alloc(sizeof(double) * 6 * (blk_a + blk_b), stream));
  hipLaunchKernelGGL(k_bbox_partial, dim3(blk_a + blk_b), dim3(BB_WG), 0, stream, a, na, b, nb, blk_a,
                     part.as<double>());
  SCTL_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(BB_WG), 0, stream, part.as<double>(), blk_a + blk_b, 0, box6);
  SCTL_HIP_TRY(hipGetLastError());
  return SCTL_AMD_SUCCESS;
}

namespace {
struct OrderLayout {
  bool small;
  size_t part, boxes, keys_in, keys_out, idx, idx_out, tmp, tmp_bytes, bytes;
  // small sets: keys and low-digit ranks (later the high-digit ranks), chunk
  // counts of the two digits, keys and indices in low-digit order
  size_t kr, hist_lo, hist_hi, tkey, tidx;
  int nchunk;
};
OrderLayout order_layout(int64_t na, int64_t nb) {
  OrderLayout m{};
  const int64_t n = na + nb, nblk = (na + BB_PER_BLK - 1) / BB_PER_BLK + (nb + BB_PER_BLK - 1) / BB_PER_BLK;
  Carve c;
  m.part = c.take(sizeof(double) * 6 * nblk);
  m.small = n <= SMALL_N;
  if (m.small) {
    m.nchunk = (int)((n + RX_L - 1) / RX_L);
    m.kr = c.take(sizeof(uint32_t) * n);
    m.hist_lo = c.take(sizeof(uint32_t) * RX_D * m.nchunk);
    m.hist_hi = c.take(sizeof(uint32_t) * RX_D * m.nchunk);
    m.tkey = c.take(sizeof(uint32_t) * n);
    m.tidx = c.take(sizeof(uint32_t) * n);
    m.bytes = c.bytes;
    return m;
  }
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

size_t hilbert_order_bytes(int64_t na, int64_t nb) { return na + nb > 0 ? order_layout(na, nb).bytes : 0; }

int hilbert_order(const double* a, int64_t na, uint32_t* perm_a, const double* b, int64_t nb, uint32_t* perm_b,
                 void* work, hipStream_t stream) {
  const int64_t n = na + nb;
  if (n <= 0) return SCTL_AMD_SUCCESS;
  SCTL_CHECK_ARG(work, "no workspace for the Hilbert order");
  SCTL_CHECK_ARG(na >= 0 && nb >= 0 && n < ((int64_t)1 << 32), "too many points for 32-bit indices");
  const int64_t per_blk = BB_PER_BLK;
  const int blk_a = (int)((na + per_blk - 1) / per_blk), blk_b = (int)((nb + per_blk - 1) / per_blk);
  const OrderLayout m = order_layout(na, nb);
  char* w = static_cast<char*>(work);
  double* part = reinterpret_cast<double*>(w + m.part);
  hipLaunchKernelGGL(k_bbox_partial, dim3(blk_a + blk_b), dim3(BB_WG), 0, stream, a, na, b, nb, blk_a, part);
  SCTL_HIP_TRY(hipGetLastError());
  if (m.small) {
    // The rank kernels reduce at most 64 partial boxes per set, one per lane.
    SCTL_CHECK_ARG(blk_a <= 64 && blk_b <= 64 && m.nchunk * (int64_t)RX_L >= n && m.nchunk <= SMALL_N / RX_L,
                   "small Hilbert order layout");
    uint32_t* kr = reinterpret_cast<uint32_t*>(w + m.kr);
    uint32_t* hist_lo = reinterpret_cast<uint32_t*>(w + m.hist_lo);
    uint32_t* hist_hi = reinterpret_cast<uint32_t*>(w + m.hist_hi);
    uint32_t* tkey = reinterpret_cast<uint32_t*>(w + m.tkey);
    uint32_t* tidx = reinterpret_cast<uint32_t*>(w + m.tidx);
    const dim3 grid(m.nchunk), blk(RX_WG);
    hipLaunchKernelGGL(k_rx_rank_low, grid, blk, 0, stream, a, na, b, nb, part, blk_a, blk_b, kr, hist_lo);
    SCTL_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_rx_scatter_low, grid, blk, 0, stream, kr, n, hist_lo, m.nchunk, tkey, tidx);
    SCTL_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_rx_rank_high, grid, blk, 0, stream, tkey, n, kr, hist_hi);
    SCTL_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_rx_scatter_high, grid, blk, 0, stream, tkey, tidx, kr, na, nb, hist_hi, m.nchunk, perm_a,
                       perm_b);
    SCTL_HIP_TRY(hipGetLastError());
    return SCTL_AMD_SUCCESS;
  }
  double* boxes = reinterpret_cast<double*>(w + m.boxes);
  uint32_t* keys_in = reinterpret_cast<uint32_t*>(w + m.keys_in);
  uint32_t* keys_out = reinterpret_cast<uint32_t*>(w + m.keys_out);
  uint32_t* idx = reinterpret_cast<uint32_t*>(w + m.idx);
  uint32_t* idx_out = reinterpret_cast<uint32_t*>(w + m.idx_out);
  hipLaunchKernelGGL(k_bbox_final, dim3(2), dim3(BB_WG), 0, stream, part, blk_a, blk_b, boxes);
  SCTL_HIP_TRY(hipGetLastError());
  const unsigned grid = (unsigned)((n + BB_WG - 1) / BB_WG);
  hipLaunchKernelGGL(k_hilbert_keys, dim3(grid), dim3(BB_WG), 0, stream, a, na, b, nb, boxes, keys_in, idx);
  SCTL_HIP_TRY(hipGetLastError());
  size_t bytes = m.tmp_bytes;
  SCTL_HIP_TRY(rocprim::radix_sort_pairs(w + m.tmp, bytes, keys_in, keys_out, idx, idx_out, (size_t)n, 0, 31, stream));
  hipLaunchKernelGGL(k_split_perm, dim3(grid), dim3(BB_WG), 0, stream, idx_out, na, nb, perm_a, perm_b);
  SCTL_HIP_TRY(hipGetLastError());
  return SCTL_AMD_SUCCESS;
}

}  // namespace sctl_amd

extern "C" int sctl_amd_hilbert_order(const double* r_a, int64_t na, uint32_t* perm_a, const double* r_b, int64_t nb,
                                      uint32_t* perm_b, void* stream) {
  using namespace sctl_amd;
  SCTL_CHECK_ARG(na >= 0 && nb >= 0, "negative size");
  SCTL_CHECK_ARG((na == 0 || (r_a && perm_a)) && (nb == 0 || (r_b && perm_b)), "NULL coordinates or permutation");
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  const hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  Scratch work;
  SCTL_TRY(work.alloc(hilbert_order_bytes(na, nb), s));
  return hilbert_order(r_a, na, perm_a, r_b, nb, perm_b, work.as<void>(), s);
}
