// helpers.hip -- seeded device input generation and an element-wise field-op kernel that
// exposes the NTT kernels' arithmetic to the parity tests.
#include <algorithm>

#include "api_internal.h"

namespace {

using gl::u32;
using gl::u64;

__device__ __forceinline__ u64 splitmix(u64 z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

constexpr u64 kGamma = 0x9E3779B97F4A7C15ull;
constexpr u64 kColumnStep = 0xD1B54A32D192ED03ull;

__global__ __launch_bounds__(256) void fill_uniform_kernel(u64 *dst, u64 stride, u64 rows, u32 cols,
                                                           u64 seed) {
  const u64 total = rows * cols;
  for (u64 k = blockIdx.x * 256ull + threadIdx.x; k < total; k += (u64)gridDim.x * 256) {
    const u64 col = k / rows, i = k - col * rows;
    const u64 s = seed + col * kColumnStep;
    u64 x = splitmix(s + (i + 1) * kGamma);
    while (x >= gl::P) x = splitmix(x ^ kGamma);
    dst[col * stride + i] = x;
  }
}

// Exponents op 5 applies (x 2^E for E = kPow2Exps[b mod 24]): every shift the NTT blocks use
// (multiples of 12 below 192) and the boundaries of the shift forms (E < 32, 32 <= E < 64,
// 64 <= E < 96 as a right shift, and their negated copies above 96).
constexpr int kPow2Exps[24] = {12, 24, 36, 48, 60, 72, 84, 96, 108, 120, 132, 144,
                               156, 168, 180, 0, 1, 31, 32, 33, 63, 64, 65, 191};
template <int I>
__device__ __forceinline__ u64 pow2_case(u64 x, u32 sel) {
  if constexpr (I == 24) {
    return 0;
  } else {
    return sel == I ? gl::mul_pow2<kPow2Exps[I]>(x) : pow2_case<I + 1>(x, sel);
  }
}

__global__ __launch_bounds__(256) void field_op_kernel(int op, const u64 *a, const u64 *b, u64 *out,
                                                       size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const u64 x = gl::canon(a[i]);
    const u64 y = b ? gl::canon(b[i]) : 0;
    u64 r;
    switch (op) {
      case 0: r = gl::add(x, y); break;
      case 1: r = gl::sub(x, y); break;
      case 2: r = gl::mul(a[i], b[i]); break;  // mul accepts unreduced operands
      case 3: r = gl::reduce1(a[i]); break;    // canonical form of any u64
      case 5: r = pow2_case<0>(a[i], (u32)(b[i] % 24)); break;  // x 2^E, x any u64
      // the reductions themselves on chosen words: lo = a, hl = b[0:32), hh = b[32:63), bi = b[63]
      case 6: {
        const u64 bi = (b[i] >> 63) ? ~0ull : 0ull;  // a whole-wave mask built per lane below
        const u32 hl = (u32)b[i], hh = (u32)(b[i] >> 32) & 0x7FFFFFFFu;
        u64 m;
        asm("v_cmp_ne_u64_e64 %0, %1, 0" : "=s"(m) : "v"(bi));
        r = gl::reduce3b(a[i], hl, hh, m);
        break;
      }
      case 7: r = gl::reduce2(a[i], (u32)b[i]); break;  // lo = a, hl = b[0:32)
      case 8: r = gl::reduce3(a[i], (u32)b[i], (u32)(b[i] >> 32)); break;  // hh = b[32:64)
      default: r = gl::inv(x); break;
    }
    out[i] = r;
  }
}

// Register-only loop over the NTT passes' butterfly arithmetic (one general product, one
// canonical add and one canonical subtract per butterfly, 8 independent values per thread): the
// VALU issue rate this instruction mix reaches with no memory, LDS or barrier in the way.
constexpr int kBflyValues = 8;
__global__ __launch_bounds__(256) void butterfly_rate_kernel(const u64 *in, u64 *out, u32 iters) {
  const u64 i = blockIdx.x * 256ull + threadIdx.x;
  const u64 w = gl::canon(in[i]) | 1;
  u64 v[kBflyValues];
#pragma unroll
  for (int c = 0; c < kBflyValues; c++) v[c] = gl::canon(in[i] + c);
  for (u32 it = 0; it < iters; it++) {
#pragma unroll
    for (int c = 0; c < kBflyValues; c++) {
      const u64 a = v[c], t = gl::mul(v[(c + 1) % kBflyValues], w);
      v[c] = gl::add(a, t);
      v[(c + 1) % kBflyValues] = gl::sub(a, t);
    }
  }
  u64 acc = 0;
#pragma unroll
  for (int c = 0; c < kBflyValues; c++) acc ^= v[c];
  out[i] = acc;
}

}  // namespace

extern "C" int ba_diag_butterfly_rate(const uint64_t *in, uint64_t *out, uint32_t blocks, uint32_t iters,
                                      ba_stream_t stream) {
  if (!in || !out || blocks == 0) return BA_ERROR_INVALID_VALUE;
  hipLaunchKernelGGL(butterfly_rate_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, in, out, iters);
  return (int)hipGetLastError();
}

extern "C" int ba_fill_uniform(uint64_t *dst, uint64_t stride, uint64_t rows, uint32_t cols,
                               uint64_t seed, ba_stream_t stream) {
  if (!dst || (cols > 1 && stride < rows)) return BA_ERROR_INVALID_VALUE;
  if (rows == 0 || cols == 0) return BA_SUCCESS;
  const u64 total = rows * cols;
  const u32 grid = (u32)std::min<u64>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, dst, stride,
                     rows, cols, seed);
  return (int)hipGetLastError();
}

extern "C" int ba_field_op(int op, const uint64_t *a, const uint64_t *b, uint64_t *out, size_t n,
                           ba_stream_t stream) {
  if (op < 0 || op > 8 || !a || !out || ((op <= 2 || op >= 5) && !b)) return BA_ERROR_INVALID_VALUE;
  if (n == 0) return BA_SUCCESS;
  const u32 grid = (u32)std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(field_op_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, op, a, b, out, n);
  return (int)hipGetLastError();
}
