// bit_reverse.hip -- per-column bit-reversal permutation.
//
// Reference: boojum_cuda::ops_complex::bit_reverse[_in_place] (crates/boojum-cuda/src/
// ops_complex.rs:262-331, kernels native/ops_complex.cu:88-156).  For log_rows >= 10 an index
// is split as (A: top 5 bits, M: middle bits, B: low 5 bits); bitrev maps it to
// (rev5(B), rev(M), rev5(A)), so the 32x32 tile with middle bits M lands, transposed, on the
// tile with middle bits rev(M).  A workgroup owns the pair {M, rev(M)}, so the same kernel
// serves in-place and out-of-place calls; loads and stores are 256-byte row segments.
#include <algorithm>

#include "api_internal.h"

namespace {

using gl::u32;
using gl::u64;

__device__ __forceinline__ u32 brev(u32 x, u32 bits) { return bits ? (__brev(x) >> (32 - bits)) : 0; }

__global__ __launch_bounds__(256) void bit_reverse_tiles(const u64 *src, u64 src_stride, u64 *dst,
                                                         u64 dst_stride, u32 log_rows) {
  __shared__ u64 tile[2][32][33];
  const u32 mid_bits = log_rows - 10;
  const u32 M = blockIdx.x;
  const u32 Mr = brev(M, mid_bits);
  if (M > Mr) return;
  const u64 *s = src + (u64)blockIdx.y * src_stride;
  u64 *d = dst + (u64)blockIdx.y * dst_stride;
  const u32 tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const u32 hi_shift = log_rows - 5;
#pragma unroll
  for (u32 j = 0; j < 4; j++) {
    const u32 A = ty + 8 * j;
    tile[0][A][tx] = s[((u64)A << hi_shift) | (M << 5) | tx];
    if (Mr != M) tile[1][A][tx] = s[((u64)A << hi_shift) | (Mr << 5) | tx];
  }
  __syncthreads();
#pragma unroll
  for (u32 j = 0; j < 4; j++) {
    const u32 R = ty + 8 * j;
    d[((u64)R << hi_shift) | (Mr << 5) | tx] = tile[0][brev(tx, 5)][brev(R, 5)];
    if (Mr != M) d[((u64)R << hi_shift) | (M << 5) | tx] = tile[1][brev(tx, 5)][brev(R, 5)];
  }
}

__global__ __launch_bounds__(256) void bit_reverse_pairs(const u64 *src, u64 src_stride, u64 *dst,
                                                         u64 dst_stride, u32 log_rows) {
  const u32 i = blockIdx.x * 256 + threadIdx.x;
  if (i >> log_rows) return;
  const u32 j = brev(i, log_rows);
  if (i > j) return;
  const u64 *s = src + (u64)blockIdx.y * src_stride;
  u64 *d = dst + (u64)blockIdx.y * dst_stride;
  const u64 a = s[i], b = s[j];
  d[i] = b;
  d[j] = a;
}

}  // namespace

extern "C" int ba_bit_reverse(const uint64_t *src, uint64_t src_stride, uint64_t *dst,
                              uint64_t dst_stride, uint32_t log_rows, uint32_t cols,
                              ba_stream_t stream) {
  if (log_rows > 31 || !src || !dst) return BA_ERROR_INVALID_VALUE;
  const uint64_t rows = 1ull << log_rows;
  if (cols > 1 && (src_stride < rows || dst_stride < rows)) return BA_ERROR_INVALID_VALUE;
  if (src != dst && cols > 0) {
    // distinct buffers must not overlap (an in-place call passes the same pointer)
    const uint64_t *s_end = src + (uint64_t)(cols - 1) * src_stride + rows;
    const uint64_t *d_end = dst + (uint64_t)(cols - 1) * dst_stride + rows;
    if (src < d_end && dst < s_end) return BA_ERROR_INVALID_VALUE;
  }
  hipStream_t s = (hipStream_t)stream;
  for (uint32_t c0 = 0; c0 < cols; c0 += 65535) {
    const uint32_t nc = std::min<uint32_t>(65535, cols - c0);
    const uint64_t *sp = src + (uint64_t)c0 * src_stride;
    uint64_t *dp = dst + (uint64_t)c0 * dst_stride;
    if (log_rows >= 10) {
      dim3 grid(1u << (log_rows - 10), nc);
      hipLaunchKernelGGL(bit_reverse_tiles, grid, dim3(256), 0, s, sp, src_stride, dp, dst_stride,
                         log_rows);
    } else {
      dim3 grid((uint32_t)((rows + 255) / 256), nc);
      hipLaunchKernelGGL(bit_reverse_pairs, grid, dim3(256), 0, s, sp, src_stride, dp, dst_stride,
                         log_rows);
    }
    const hipError_t err = hipGetLastError();
    if (err != hipSuccess) return (int)err;
  }
  return BA_SUCCESS;
}

extern "C" int ba_bit_reverse_in_place(uint64_t *values, uint64_t stride, uint32_t log_rows,
                                       uint32_t cols, ba_stream_t stream) {
  return ba_bit_reverse(values, stride, values, stride, log_rows, cols, stream);
}
