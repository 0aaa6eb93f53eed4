// powers.hip -- vectors of powers and omega shifts that read the context's power tables:
// boojum_cuda::ops_complex::{get_powers_of_w, get_powers_of_g, get_powers_by_val,
// get_powers_by_ref, omega_shift, omega_shift_in_place} (crates/boojum-cuda/src/ops_complex.rs:31-260).
//
// Semantics kept from the reference (its kernels: crates/boojum-cuda/native/ops_complex.cu:8-86),
// with u32 index arithmetic wrapping as there:
//   powers_of_w: r[i] = omega_{2^30}^(+-(j << (30 - log_degree))), j = bitrev_{log_degree}(i + offset)
//                or i + offset; powers_of_g: 7^(that exponent mod 2^30) (7^-1 when inverse);
//   powers_by_val/ref: r[i] = base^((bit_reverse ? brev32(i) : i) + offset), base field or the
//                quadratic extension u^2 = 7 (its two coordinates stored as two planes of count);
//   omega_shift: dst[i] = src[i] * omega_{2^30}^(+-((shift * (i + offset)) << (30 - log_degree))).
// Every kernel is a grid-stride loop, one element per lane per step, coalesced 8-byte accesses;
// where the exponent is linear in the index a lane steps its power by one multiply per element.
#include <algorithm>

#include "api_internal.h"

namespace {

using gl::u32;
using gl::u64;

constexpr u32 kThreads = 256;

u32 grid_for(u64 count) { return (u32)std::max<u64>(1, std::min<u64>((count + kThreads - 1) / kThreads, 65536)); }

__device__ __forceinline__ u32 raw_exponent(u32 i, u32 log_degree, bool bit_reverse) {
  return bit_reverse ? gl::ntt::bitrev32(i, log_degree) : i;
}

// omega_{2^30} has order 2^30 and every u32 wrap below is a multiple of 2^30, so omega^e depends on e
// mod 2^30 only: an exponent linear in the index (natural order) gives powers a lane can step by
// one constant factor per grid stride, instead of two dependent table reads per element.
__global__ __launch_bounds__(kThreads) void powers_of_w_kernel(gl::Tables tab, u32 log_degree, u32 offset,
                                                               bool inverse, bool bit_reverse, u64 *result,
                                                               u32 count) {
  const u32 shift = gl::OMEGA_LOG_ORDER - log_degree;
  const u32 stride = gridDim.x * kThreads;
  u32 i = blockIdx.x * kThreads + threadIdx.x;
  if (bit_reverse) {
    for (u64 k = i; k < count; k += stride)
      result[k] = gl::pow_w(tab, raw_exponent((u32)k + offset, log_degree, true) << shift, inverse);
    return;
  }
  if (i >= count) return;
  u64 w = gl::pow_w(tab, (i + offset) << shift, inverse);
  const u64 step = gl::pow_w(tab, stride << shift, inverse);
  for (;;) {
    result[i] = w;
    if (count - i <= stride) break;
    i += stride;
    w = gl::mul(w, step);
  }
}

__global__ __launch_bounds__(kThreads) void powers_of_g_kernel(gl::Tables tab, u32 log_degree, u32 offset,
                                                               bool inverse, bool bit_reverse, u64 *result,
                                                               u32 count) {
  const u32 shift = gl::OMEGA_LOG_ORDER - log_degree;
  for (u64 i = blockIdx.x * kThreads + threadIdx.x; i < count; i += gridDim.x * kThreads) {
    const u32 e = raw_exponent((u32)i + offset, log_degree, bit_reverse) << shift;
    result[i] = gl::table_pow(inverse ? tab.g_inv : tab.g, e);
  }
}

// quadratic extension F_p[u] / (u^2 - 7)
struct Ext {
  u64 a, b;
};
__device__ __forceinline__ Ext ext_mul(const Ext &x, const Ext &y) {
  const u64 ac = gl::mul(x.a, y.a), bd = gl::mul(x.b, y.b);
  const u64 ad = gl::mul(x.a, y.b), bc = gl::mul(x.b, y.a);
  return {gl::add(ac, gl::mul(bd, 7)), gl::add(ad, bc)};
}
__device__ __forceinline__ Ext ext_pow(Ext x, u32 e) {
  Ext r{1, 0};
  while (e) {
    if (e & 1) r = ext_mul(r, x);
    e >>= 1;
    if (e) x = ext_mul(x, x);
  }
  return r;
}

__device__ __forceinline__ u32 by_val_exponent(u32 i, u32 offset, bool bit_reverse) {
  return (bit_reverse ? __brev(i) : i) + offset;
}

// natural order: lane i starts at base^(i + offset) and steps by base^(grid stride), one
// multiply per element; bit-reversed order, or an exponent that wraps past 2^32 (which the
// reference's u32 arithmetic does), one exponentiation per element
__global__ __launch_bounds__(kThreads) void powers_bf_kernel(const u64 *base_ptr, u64 base_val, u32 offset,
                                                             bool bit_reverse, u64 *result, u32 count) {
  const u64 base = gl::canon(base_ptr ? *base_ptr : base_val);
  const u32 stride = gridDim.x * kThreads;
  u32 i = blockIdx.x * kThreads + threadIdx.x;
  if (bit_reverse || (u64)offset + count > 0x100000000ull) {
    for (u64 k = i; k < count; k += stride) result[k] = gl::pow_u64(base, by_val_exponent((u32)k, offset, bit_reverse));
    return;
  }
  if (i >= count) return;
  u64 x = gl::pow_u64(base, by_val_exponent(i, offset, false));
  const u64 step = gl::pow_u64(base, stride);
  for (;;) {
    result[i] = x;
    if (count - i <= stride) break;
    i += stride;
    x = gl::mul(x, step);
  }
}

__global__ __launch_bounds__(kThreads) void powers_ef_kernel(const u64 *base_ptr, u64 base_a, u64 base_b, u32 offset,
                                                             bool bit_reverse, u64 *result, u32 count) {
  const Ext base{gl::canon(base_ptr ? base_ptr[0] : base_a), gl::canon(base_ptr ? base_ptr[1] : base_b)};
  const u32 stride = gridDim.x * kThreads;
  u32 i = blockIdx.x * kThreads + threadIdx.x;
  if (bit_reverse || (u64)offset + count > 0x100000000ull) {
    for (u64 k = i; k < count; k += stride) {
      const Ext v = ext_pow(base, by_val_exponent((u32)k, offset, bit_reverse));
      result[k] = v.a;
      result[k + count] = v.b;
    }
    return;
  }
  if (i >= count) return;
  Ext x = ext_pow(base, by_val_exponent(i, offset, false));
  const Ext step = ext_pow(base, stride);
  for (;;) {
    result[i] = x.a;
    result[(u64)i + count] = x.b;
    if (count - i <= stride) break;
    i += stride;
    x = ext_mul(x, step);
  }
}

// the exponent (shift (i + offset)) << degree_shift is linear in i modulo 2^30: each lane steps its
// factor by omega^(+-((shift stride) << degree_shift)) per grid stride
__global__ __launch_bounds__(kThreads) void omega_shift_kernel(gl::Tables tab, const u64 *src, u32 log_degree,
                                                               u32 offset, bool inverse, u32 shift, u64 *dst,
                                                               u32 count) {
  const u32 degree_shift = gl::OMEGA_LOG_ORDER - log_degree;
  const u32 stride = gridDim.x * kThreads;
  u32 i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= count) return;
  u64 w = gl::pow_w(tab, (shift * (i + offset)) << degree_shift, inverse);
  const u64 step = gl::pow_w(tab, (shift * stride) << degree_shift, inverse);
  for (;;) {
    dst[i] = gl::mul(src[i], w);
    if (count - i <= stride) break;
    i += stride;
    w = gl::mul(w, step);
  }
}

int tables(const gl::Tables **t) {
  const ba_context *ctx = ba_context_for_current_device();
  if (!ctx) return BA_ERROR_NOT_INITIALIZED;
  *t = &ctx->tab;
  return BA_SUCCESS;
}

}  // namespace

extern "C" int ba_get_powers_of_w(uint32_t log_degree, uint32_t offset, int inverse, int bit_reverse,
                                  uint64_t *result, uint32_t count, ba_stream_t stream) {
  if (log_degree > gl::OMEGA_LOG_ORDER || (count && !result)) return BA_ERROR_INVALID_VALUE;
  const gl::Tables *t;
  if (int rc = tables(&t)) return rc;
  if (!count) return BA_SUCCESS;
  hipLaunchKernelGGL(powers_of_w_kernel, dim3(grid_for(count)), dim3(kThreads), 0, (hipStream_t)stream, *t,
                     log_degree, offset, inverse != 0, bit_reverse != 0, result, count);
  return (int)hipGetLastError();
}

extern "C" int ba_get_powers_of_g(uint32_t log_degree, uint32_t offset, int inverse, int bit_reverse,
                                  uint64_t *result, uint32_t count, ba_stream_t stream) {
  if (log_degree > gl::OMEGA_LOG_ORDER || (count && !result)) return BA_ERROR_INVALID_VALUE;
  const gl::Tables *t;
  if (int rc = tables(&t)) return rc;
  if (!count) return BA_SUCCESS;
  hipLaunchKernelGGL(powers_of_g_kernel, dim3(grid_for(count)), dim3(kThreads), 0, (hipStream_t)stream, *t,
                     log_degree, offset, inverse != 0, bit_reverse != 0, result, count);
  return (int)hipGetLastError();
}

extern "C" int ba_get_powers_by_val_bf(uint64_t base, uint32_t offset, int bit_reverse, uint64_t *result,
                                       uint32_t count, ba_stream_t stream) {
  if (count && !result) return BA_ERROR_INVALID_VALUE;
  if (!count) return BA_SUCCESS;
  hipLaunchKernelGGL(powers_bf_kernel, dim3(grid_for(count)), dim3(kThreads), 0, (hipStream_t)stream,
                     (const u64 *)nullptr, base, offset, bit_reverse != 0, result, count);
  return (int)hipGetLastError();
}

extern "C" int ba_get_powers_by_ref_bf(const uint64_t *base, uint32_t offset, int bit_reverse, uint64_t *result,
                                       uint32_t count, ba_stream_t stream) {
  if (!base || (count && !result)) return BA_ERROR_INVALID_VALUE;
  if (!count) return BA_SUCCESS;
  hipLaunchKernelGGL(powers_bf_kernel, dim3(grid_for(count)), dim3(kThreads), 0, (hipStream_t)stream, base,
                     (u64)0, offset, bit_reverse != 0, result, count);
  return (int)hipGetLastError();
}

extern "C" int ba_get_powers_by_val_ef(uint64_t base_c0, uint64_t base_c1, uint32_t offset, int bit_reverse,
                                       uint64_t *result, uint32_t count, ba_stream_t stream) {
  if (count && !result) return BA_ERROR_INVALID_VALUE;
  if (!count) return BA_SUCCESS;
  hipLaunchKernelGGL(powers_ef_kernel, dim3(grid_for(count)), dim3(kThreads), 0, (hipStream_t)stream,
                     (const u64 *)nullptr, base_c0, base_c1, offset, bit_reverse != 0, result, count);
  return (int)hipGetLastError();
}

extern "C" int ba_get_powers_by_ref_ef(const uint64_t *base, uint32_t offset, int bit_reverse, uint64_t *result,
                                       uint32_t count, ba_stream_t stream) {
  if (!base || (count && !result)) return BA_ERROR_INVALID_VALUE;
  if (!count) return BA_SUCCESS;
  hipLaunchKernelGGL(powers_ef_kernel, dim3(grid_for(count)), dim3(kThreads), 0, (hipStream_t)stream, base,
                     (u64)0, (u64)0, offset, bit_reverse != 0, result, count);
  return (int)hipGetLastError();
}

extern "C" int ba_omega_shift(const uint64_t *src, uint32_t log_degree, uint32_t offset, int inverse, uint32_t shift,
                              uint64_t *dst, uint32_t count, ba_stream_t stream) {
  if (log_degree > gl::OMEGA_LOG_ORDER || (count && (!src || !dst))) return BA_ERROR_INVALID_VALUE;
  // element i is read and written by the same lane: in place is the only overlap that is safe
  if (src != dst && (const uint64_t *)dst < src + count && src < (const uint64_t *)dst + count)
    return BA_ERROR_INVALID_VALUE;
  const gl::Tables *t;
  if (int rc = tables(&t)) return rc;
  if (!count) return BA_SUCCESS;
  hipLaunchKernelGGL(omega_shift_kernel, dim3(grid_for(count)), dim3(kThreads), 0, (hipStream_t)stream, *t, src,
                     log_degree, offset, inverse != 0, shift, dst, count);
  return (int)hipGetLastError();
}

extern "C" int ba_omega_shift_in_place(uint64_t *values, uint32_t log_degree, uint32_t offset, int inverse,
                                       uint32_t shift, uint32_t count, ba_stream_t stream) {
  return ba_omega_shift(values, log_degree, offset, inverse, shift, values, count, stream);
}
