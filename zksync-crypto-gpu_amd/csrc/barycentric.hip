// barycentric.hip -- evaluation of polynomials given by their values on a coset domain at one
// point X, without interpolating: boojum_cuda::barycentric (crates/boojum-cuda/src/barycentric.rs:83-300).
//
// For the domain {c w^i : i < N} (w = omega_N, c the coset shift) and evaluations y_i,
//   f(X) = sum_i L_i y_i,   L_i = common / (X w^-i - c),   common = c (X^N - c^N) / (N c^N),
// the same factorisation the reference uses (native/barycentric.cu:11-93).  X is a base-field
// or extension-field element (F_p[u]/(u^2 - 7)); evaluations are base- or extension-field
// columns.  Layouts follow the reference's vectorized extension type: an extension vector of
// count elements is two planes (c0[count], c1[count]); an extension matrix column k starts at
// 2 k count; extension results are interleaved (c0, c1) per polynomial.
//
// Kernels:
//  * coefficients: each lane takes 8 elements (grid-strided), forms X w^-i - c from the inverse
//    power table and inverts the 8 at once (one field inversion and 3 multiplies per element,
//    Montgomery's trick), as the reference batches its inversions per thread;
//  * batch evaluation: a grid over (row chunks of 4096, column groups), each lane 16 rows of a
//    column, coefficients kept in registers across the columns of its group, a wave-level
//    reduction through LDS, one partial sum per (chunk, column); a second kernel sums the chunks.
// Both are HBM-bound for base-field data (8 B per element read once) and VALU-bound for
// extension-field products.
#include <algorithm>
#include <mutex>
#include <vector>

#include "api_internal.h"

namespace {

using gl::u32;
using gl::u64;

struct Ext {
  u64 a, b;
};
__device__ __forceinline__ Ext ext_add(const Ext &x, const Ext &y) { return {gl::add(x.a, y.a), gl::add(x.b, y.b)}; }
__device__ __forceinline__ Ext ext_sub(const Ext &x, const Ext &y) { return {gl::sub(x.a, y.a), gl::sub(x.b, y.b)}; }
__device__ __forceinline__ Ext ext_mul(const Ext &x, const Ext &y) {
  const u64 ac = gl::mul(x.a, y.a), bd = gl::mul(x.b, y.b);
  const u64 ad = gl::mul(x.a, y.b), bc = gl::mul(x.b, y.a);
  return {gl::add(ac, gl::mul(bd, 7)), gl::add(ad, bc)};
}
__device__ __forceinline__ Ext ext_scale(const Ext &x, u64 s) { return {gl::mul(x.a, s), gl::mul(x.b, s)}; }
__device__ __forceinline__ Ext ext_pow(Ext x, u64 e) {
  Ext r{1, 0};
  while (e) {
    if (e & 1) r = ext_mul(r, x);
    e >>= 1;
    if (e) x = ext_mul(x, x);
  }
  return r;
}
// (a + b u)^-1 = (a - b u) / (a^2 - 7 b^2); 0 maps to 0
__device__ __forceinline__ Ext ext_inv(const Ext &x) {
  const u64 norm = gl::sub(gl::sqr(x.a), gl::mul(gl::sqr(x.b), 7));
  const u64 ni = gl::inv(norm);
  return {gl::mul(x.a, ni), gl::neg(gl::mul(x.b, ni))};
}

// Field traits: base field (u64) or extension (Ext) element operations
template <bool EXT>
struct F;
template <>
struct F<false> {
  typedef u64 T;
  static __device__ __forceinline__ T one() { return 1; }
  static __device__ __forceinline__ T zero() { return 0; }
  static __device__ __forceinline__ T mul(const T &x, const T &y) { return gl::mul(x, y); }
  static __device__ __forceinline__ T add(const T &x, const T &y) { return gl::add(x, y); }
  static __device__ __forceinline__ T sub(const T &x, const T &y) { return gl::sub(x, y); }
  static __device__ __forceinline__ T scale(const T &x, u64 s) { return gl::mul(x, s); }
  static __device__ __forceinline__ T inv(const T &x) { return gl::inv(x); }
  static __device__ __forceinline__ T pow(const T &x, u64 e) { return gl::pow_u64(x, e); }
  static __device__ __forceinline__ T from_base(u64 c) { return c; }
  static __device__ __forceinline__ T load(const u64 *p, u64 i, u64) { return gl::canon(p[i]); }
  static __device__ __forceinline__ void store(u64 *p, u64 i, u64, const T &v) { p[i] = v; }
};
template <>
struct F<true> {
  typedef Ext T;
  static __device__ __forceinline__ T one() { return {1, 0}; }
  static __device__ __forceinline__ T zero() { return {0, 0}; }
  static __device__ __forceinline__ T mul(const T &x, const T &y) { return ext_mul(x, y); }
  static __device__ __forceinline__ T add(const T &x, const T &y) { return ext_add(x, y); }
  static __device__ __forceinline__ T sub(const T &x, const T &y) { return ext_sub(x, y); }
  static __device__ __forceinline__ T scale(const T &x, u64 s) { return ext_scale(x, s); }
  static __device__ __forceinline__ T inv(const T &x) { return ext_inv(x); }
  static __device__ __forceinline__ T pow(const T &x, u64 e) { return ext_pow(x, e); }
  static __device__ __forceinline__ T from_base(u64 c) { return {c, 0}; }
  // element i of a two-plane vector whose planes are `plane` apart
  static __device__ __forceinline__ T load(const u64 *p, u64 i, u64 plane) {
    return {gl::canon(p[i]), gl::canon(p[i + plane])};
  }
  static __device__ __forceinline__ void store(u64 *p, u64 i, u64 plane, const T &v) {
    p[i] = v.a;
    p[i + plane] = v.b;
  }
};

// common = c (X^N - c^N) / (N c^N)   (one lane)
template <bool EXT>
__global__ void common_factor_kernel(const u64 *x_ptr, u64 *common, u64 coset, u32 count) {
  typedef F<EXT> Fx;
  if (blockIdx.x | threadIdx.x) return;
  const typename Fx::T x = Fx::load(x_ptr, 0, 1);
  const u64 c = gl::canon(coset);
  const u64 cn = gl::pow_u64(c, count);
  const typename Fx::T num = Fx::scale(Fx::sub(Fx::pow(x, count), Fx::from_base(cn)), c);
  const u64 den_inv = gl::inv(gl::mul(count % gl::P, cn));
  Fx::store(common, 0, 1, Fx::scale(num, den_inv));
}

constexpr u32 kThreads = 256;
constexpr int kInvBatch = 8;

// L_i = common / (X w^-i - c): each lane takes kInvBatch elements i = lane + k * stride and
// inverts their product once
template <bool EXT>
__global__ __launch_bounds__(kThreads) void coeffs_kernel(gl::Tables tab, const u64 *x_ptr, const u64 *common_ptr,
                                                          u64 coset, u32 log_count, u64 *coeffs) {
  typedef F<EXT> Fx;
  typedef typename Fx::T T;
  const u64 count = 1ull << log_count;
  const u64 stride = (u64)gridDim.x * kThreads;
  const u64 i0 = (u64)blockIdx.x * kThreads + threadIdx.x;
  if (i0 >= count) return;
  const T x = Fx::load(x_ptr, 0, 1);
  const T common = Fx::load(common_ptr, 0, 1);
  const u64 c = gl::canon(coset);
  const u32 shift = gl::OMEGA_LOG_ORDER - log_count;
  T den[kInvBatch], pre[kInvBatch];
  T run = Fx::one();
  int n = 0;
  for (int k = 0; k < kInvBatch; k++) {
    const u64 i = i0 + k * stride;
    if (i >= count) break;
    const u64 w_inv = gl::pow_w(tab, (u32)i << shift, true);
    den[k] = Fx::sub(Fx::scale(x, w_inv), Fx::from_base(c));
    pre[k] = run;  // product of den[0..k)
    run = Fx::mul(run, den[k]);
    n = k + 1;
  }
  T inv = Fx::mul(Fx::inv(run), common);  // common / prod
  for (int k = n - 1; k >= 0; k--) {
    Fx::store(coeffs, i0 + k * stride, count, Fx::mul(inv, pre[k]));  // common / den[k]
    inv = Fx::mul(inv, den[k]);
  }
}

constexpr int kRows = 16;                        // rows per lane per column
constexpr u32 kChunk = kThreads * kRows;         // rows per workgroup
constexpr u32 kColsPerGroup = 8;                 // columns a workgroup visits

// partial[chunk, col] = sum over the chunk's rows of L_row * y[row, col]
template <bool XE, bool YE>
__global__ __launch_bounds__(kThreads) void partial_kernel(const u64 *ys, const u64 *coeffs, u32 log_count,
                                                           u32 num_polys, u64 *partial, u32 chunks) {
  typedef F<XE> Fx;
  typedef typename Fx::T T;
  const u64 count = 1ull << log_count;
  const u32 chunk = blockIdx.x;
  const u64 r0 = (u64)chunk * kChunk;
  const u32 t = threadIdx.x;
  T l[kRows];
  for (int k = 0; k < kRows; k++) {
    const u64 r = r0 + t + (u64)k * kThreads;
    l[k] = r < count ? Fx::load(coeffs, r, count) : Fx::zero();
  }
  __shared__ T red[kThreads / 64];
  const u32 col_end = min(num_polys, (blockIdx.y + 1) * kColsPerGroup);
  // base-field values: the next column's are loaded while the current column is accumulated
  // (rows past the end of a short column read as zero and contribute nothing); extension values
  // are read in the loop (two registers each would leave one wave per SIMD)
  u64 nxt[YE ? 1 : kRows];
  auto fetch = [&](u32 col) {
    if constexpr (!YE) {
#pragma unroll
      for (int k = 0; k < kRows; k++) {
        const u64 r = r0 + t + (u64)k * kThreads;
        nxt[k] = r < count ? ys[(u64)col * count + r] : 0;
      }
    }
  };
  u32 col = blockIdx.y * kColsPerGroup;
  if (col < col_end) fetch(col);
  for (; col < col_end; col++) {
    T acc = Fx::zero();
    if constexpr (YE) {
      for (int k = 0; k < kRows; k++) {
        const u64 r = r0 + t + (u64)k * kThreads;
        if (r >= count) break;
        const Ext y = F<true>::load(ys + (u64)col * 2 * count, r, count);
        acc = Fx::add(acc, Fx::mul(l[k], T(y)));
      }
    } else {
      u64 cur[kRows];
#pragma unroll
      for (int k = 0; k < kRows; k++) cur[k] = nxt[k];
      if (col + 1 < col_end) fetch(col + 1);
#pragma unroll
      for (int k = 0; k < kRows; k++) acc = Fx::add(acc, Fx::scale(l[k], gl::canon(cur[k])));
    }
    // wave reduction through shuffles, then across the 4 waves through LDS
    for (int off = 32; off > 0; off >>= 1) {
      T o;
      if constexpr (XE) {
        o.a = __shfl_xor(acc.a, off);
        o.b = __shfl_xor(acc.b, off);
      } else {
        o = __shfl_xor(acc, off);
      }
      acc = Fx::add(acc, o);
    }
    if ((t & 63) == 0) red[t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
      T s = red[0];
      for (u32 w = 1; w < kThreads / 64; w++) s = Fx::add(s, red[w]);
      // partial sums: column `col` holds `chunks` values (two planes for the extension)
      Fx::store(partial + (u64)col * chunks * (XE ? 2 : 1), chunk, chunks, s);
    }
    __syncthreads();
  }
}

// evals[col] = sum over chunks of partial[chunk, col]; extension results interleaved (c0, c1)
template <bool XE>
__global__ __launch_bounds__(kThreads) void final_kernel(const u64 *partial, u32 chunks, u32 num_polys, u64 *evals) {
  typedef F<XE> Fx;
  typedef typename Fx::T T;
  const u32 col = blockIdx.x * kThreads + threadIdx.x;
  if (col >= num_polys) return;
  const u64 *p = partial + (u64)col * chunks * (XE ? 2 : 1);
  T s = Fx::zero();
  for (u32 k = 0; k < chunks; k++) s = Fx::add(s, Fx::load(p, k, chunks));
  if constexpr (XE) {
    evals[2 * (u64)col] = s.a;
    evals[2 * (u64)col + 1] = s.b;
  } else {
    evals[col] = s;
  }
}

bool pow2(u64 x) { return x && !(x & (x - 1)); }
u32 log2u(u64 x) {
  u32 l = 0;
  while ((1ull << l) < x) l++;
  return l;
}

}  // namespace

extern "C" int ba_barycentric_precompute_lagrange_coeffs(int ext, const uint64_t *x, uint64_t *common_factor,
                                                         uint64_t coset, uint64_t *lagrange_coeffs, uint32_t count,
                                                         int bit_reverse, ba_stream_t stream) {
  if (!x || !common_factor || !lagrange_coeffs || !pow2(count) || (ext != 0 && ext != 1)) return BA_ERROR_INVALID_VALUE;
  const u32 log_count = log2u(count);
  if (log_count > gl::OMEGA_LOG_ORDER) return BA_ERROR_INVALID_VALUE;
  const ba_context *ctx = ba_context_for_current_device();
  if (!ctx) return BA_ERROR_NOT_INITIALIZED;
  hipStream_t s = (hipStream_t)stream;
  const u32 grid = (u32)std::max<u64>(1, ((u64)count + (u64)kThreads * kInvBatch - 1) / ((u64)kThreads * kInvBatch));
  if (ext) {
    hipLaunchKernelGGL(common_factor_kernel<true>, dim3(1), dim3(64), 0, s, x, common_factor, coset, count);
    hipLaunchKernelGGL(coeffs_kernel<true>, dim3(grid), dim3(kThreads), 0, s, ctx->tab, x,
                       (const u64 *)common_factor, coset, log_count, lagrange_coeffs);
  } else {
    hipLaunchKernelGGL(common_factor_kernel<false>, dim3(1), dim3(64), 0, s, x, common_factor, coset, count);
    hipLaunchKernelGGL(coeffs_kernel<false>, dim3(grid), dim3(kThreads), 0, s, ctx->tab, x,
                       (const u64 *)common_factor, coset, log_count, lagrange_coeffs);
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return (int)err;
  // bit-reversed order, as the reference's bit_reverse_in_place epilogue (barycentric.rs:118-124):
  // the extension planes are two columns of count
  if (bit_reverse) return ba_bit_reverse_in_place(lagrange_coeffs, count, log_count, ext ? 2 : 1, stream);
  return BA_SUCCESS;
}

extern "C" int ba_barycentric_batch_eval_temp_elems(int kind, uint32_t count, uint32_t num_polys, uint64_t *elems) {
  if (!elems || !pow2(count) || kind < 0 || kind > 2) return BA_ERROR_INVALID_VALUE;
  const u64 chunks = ((u64)count + kChunk - 1) / kChunk;
  *elems = chunks * num_polys * (kind == 0 ? 1 : 2);
  return BA_SUCCESS;
}

extern "C" int ba_barycentric_batch_eval(int kind, const uint64_t *ys, uint32_t count, uint32_t num_polys,
                                         const uint64_t *lagrange_coeffs, uint64_t *temp, uint64_t temp_elems,
                                         uint64_t *evals, ba_stream_t stream) {
  uint64_t need = 0;
  if (int rc = ba_barycentric_batch_eval_temp_elems(kind, count, num_polys, &need)) return rc;
  if (num_polys == 0) return BA_SUCCESS;
  if (!ys || !lagrange_coeffs || !evals || !temp || temp_elems < need) return BA_ERROR_INVALID_VALUE;
  const u32 log_count = log2u(count);
  const u32 chunks = (u32)(((u64)count + kChunk - 1) / kChunk);
  hipStream_t s = (hipStream_t)stream;
  const dim3 grid(chunks, (num_polys + kColsPerGroup - 1) / kColsPerGroup);
  const dim3 fgrid((num_polys + kThreads - 1) / kThreads);
  if (grid.y > 65535) return BA_ERROR_INVALID_VALUE;
  switch (kind) {
    case 0:  // base-field evaluations at a base-field point
      hipLaunchKernelGGL((partial_kernel<false, false>), grid, dim3(kThreads), 0, s, ys, lagrange_coeffs, log_count,
                         num_polys, temp, chunks);
      hipLaunchKernelGGL(final_kernel<false>, fgrid, dim3(kThreads), 0, s, (const u64 *)temp, chunks, num_polys, evals);
      break;
    case 1:  // base-field evaluations at an extension point
      hipLaunchKernelGGL((partial_kernel<true, false>), grid, dim3(kThreads), 0, s, ys, lagrange_coeffs, log_count,
                         num_polys, temp, chunks);
      hipLaunchKernelGGL(final_kernel<true>, fgrid, dim3(kThreads), 0, s, (const u64 *)temp, chunks, num_polys, evals);
      break;
    default:  // extension evaluations at an extension point
      hipLaunchKernelGGL((partial_kernel<true, true>), grid, dim3(kThreads), 0, s, ys, lagrange_coeffs, log_count,
                         num_polys, temp, chunks);
      hipLaunchKernelGGL(final_kernel<true>, fgrid, dim3(kThreads), 0, s, (const u64 *)temp, chunks, num_polys, evals);
      break;
  }
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------- batch evaluation, no allocation
// For callers that cannot allocate device memory (code under stream capture, tight serving loops):
// the partial sums live in this file-scope array instead of a caller workspace.  Every GPU the
// library runs on holds its own copy, at its own address, until the process ends.  The partial
// kernel writes every element the final kernel reads, so nothing has to start at zero; calls that
// overlap on one GPU share the array and race.
constexpr uint64_t kStaticTempElems = 1ull << 19;  // 4 MiB
static __device__ __align__(256) uint64_t scratch[kStaticTempElems];

namespace {
std::mutex g_scratch_mutex;
std::vector<uint64_t *> g_scratch_ptr;  // per device: the device address of `scratch`

// The device address of this GPU's copy of `scratch`, looked up once per device (with that device
// current) and kept; the array's name is a host-side stand-in and is never passed to a kernel.
int scratch_for_current_device(uint64_t **ptr) {
  int device = -1;
  hipError_t err = hipGetDevice(&device);
  if (err != hipSuccess) return (int)err;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (err != hipSuccess) return (int)err;
    if (!p) return BA_ERROR_NOT_INITIALIZED;
    g_scratch_ptr[device] = (uint64_t *)p;
  }
  *ptr = g_scratch_ptr[device];
  return BA_SUCCESS;
}
}  // namespace

extern "C" int ba_barycentric_batch_eval_noalloc(int kind, const uint64_t *ys, uint32_t count, uint32_t num_polys,
                                                 const uint64_t *lagrange_coeffs, uint64_t *evals,
                                                 ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    uint64_t need = 0;
    if (int rc = ba_barycentric_batch_eval_temp_elems(kind, count, num_polys, &need)) return rc;
    if (num_polys == 0) return BA_SUCCESS;
    if (need > kStaticTempElems) return BA_ERROR_INVALID_VALUE;  // beyond the array's 2^19 elements
    if (!ys || !lagrange_coeffs || !evals) return BA_ERROR_INVALID_VALUE;
    uint64_t *temp = nullptr;
    if (int rc = scratch_for_current_device(&temp)) return rc;
    return ba_barycentric_batch_eval(kind, ys, count, num_polys, lagrange_coeffs, temp, kStaticTempElems, evals,
                                     stream);
  });
}
