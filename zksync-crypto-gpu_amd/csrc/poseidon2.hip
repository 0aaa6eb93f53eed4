// poseidon2.hip -- Poseidon2 over Goldilocks, width 12 (rate 8, capacity 4): the Merkle leaf
// and node hashes that consume LDE cosets.
//
// Entry points mirror boojum_cuda::poseidon2::GpuTreeHasher for the Goldilocks hasher
// (crates/boojum-cuda/src/poseidon2.rs:165-316: compute_leaf_hashes, compute_node_hashes,
// build_merkle_tree_leaves, build_merkle_tree_nodes, build_merkle_tree) with the same slice
// lengths, layouts and preconditions (violations return BA_ERROR_INVALID_VALUE instead of
// panicking).  The permutation is the Poseidon2 paper's (eprint 2023/323, Fig. 1 and Appendix B)
// for the instance the reference runs (crates/boojum-cuda/native/poseidon2/gl/poseidon2_gl_st.cu:9-22,
// poseidon2_gl_constants.cuh.template): external matrix M_E = circ(2 M4, M4, M4) first; 4 full
// rounds; 22 partial rounds with M_I = J + diag(2^mu); 4 full rounds; S-box x^7.  Its 30 x 12
// round constants come from boojum (poseidon_goldilocks_params), which the reference tree does
// not hold: a caller installs them once per context with ba_poseidon2_gl_set_round_constants.
//
// One thread per digest holds the 12-lane state in registers (canonical between layers; the
// linear layers sum exactly and reduce once per output); the round constants are read with
// wave-uniform indices from the context's copy.
#include <vector>

#include "api_internal.h"
#include "gl_field.h"

namespace {

using gl::u32;
using gl::u64;

constexpr int WIDTH = 12, RATE = 8, CAPACITY = 4;
constexpr int HALF_FULL_ROUNDS = 4, PARTIAL_ROUNDS = 22, ROUNDS = 2 * HALF_FULL_ROUNDS + PARTIAL_ROUNDS;
constexpr int THREADS = 256;

// The linear layers are evaluated on exact integers held as (lo, hi) words and reduced once per
// output (gl::reduce2 takes any lo + hi 2^64 with hi < 2^32): with canonical inputs M4's outputs
// are < 16 p, M_E's < 64 p and M_I's < (12 + 2^14) p, far inside that range, and each output is
// the same residue a reduction after every addition would give.
struct Wide {
  u64 lo;
  u32 hi;
};
__device__ __forceinline__ Wide wide(u64 x) { return {x, 0}; }
__device__ __forceinline__ Wide operator+(Wide a, Wide b) {
  u64 lo;
  const bool c = __builtin_add_overflow(a.lo, b.lo, &lo);
  return {lo, a.hi + b.hi + (u32)c};
}
template <int K>
__device__ __forceinline__ Wide shifted(Wide a) {  // a 2^K, K < 32
  return {a.lo << K, (a.hi << K) | (u32)(a.lo >> (64 - K))};
}
__device__ __forceinline__ u64 reduced(Wide a) { return gl::reduce2(a.lo, a.hi); }

// M4 of Appendix B, [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]], as an addition chain
__device__ __forceinline__ void m4(Wide &x0, Wide &x1, Wide &x2, Wide &x3) {
  const Wide a = x0 + x1, b = x2 + x3;
  const Wide c = shifted<1>(x1) + b;  // 2 x1 + x2 + x3
  const Wide d = shifted<1>(x3) + a;  // x0 + x1 + 2 x3
  const Wide e = shifted<2>(b) + d;   // x0 + x1 + 4 x2 + 6 x3
  const Wide f = shifted<2>(a) + c;   // 4 x0 + 6 x1 + x2 + x3
  x0 = d + f;
  x1 = f;
  x2 = c + e;
  x3 = e;
}

// M_E: M4 on each 4-lane tile, then every lane plus the sum of its position over the tiles
__device__ __forceinline__ void external_layer(u64 (&s)[WIDTH]) {
  Wide w[WIDTH];
#pragma unroll
  for (int i = 0; i < WIDTH; i++) w[i] = wide(s[i]);
#pragma unroll
  for (int t = 0; t < WIDTH; t += 4) m4(w[t], w[t + 1], w[t + 2], w[t + 3]);
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const Wide sum = w[j] + w[4 + j] + w[8 + j];
    s[j] = reduced(w[j] + sum);
    s[4 + j] = reduced(w[4 + j] + sum);
    s[8 + j] = reduced(w[8 + j] + sum);
  }
}

template <int MU>
__device__ __forceinline__ u64 sum_plus(Wide sum, u64 x) {  // sum + x 2^MU, reduced
  if constexpr (MU == 0)
    return reduced(sum + wide(x));
  else
    return reduced(sum + Wide{x << MU, (u32)(x >> (64 - MU))});
}

// M_I = J + diag(2^mu), mu = {4, 14, 11, 8, 0, 5, 2, 9, 13, 6, 3, 12}
__device__ __forceinline__ void internal_layer(u64 (&s)[WIDTH]) {
  Wide sum = wide(s[0]);
#pragma unroll
  for (int i = 1; i < WIDTH; i++) sum = sum + wide(s[i]);
  s[0] = sum_plus<4>(sum, s[0]);
  s[1] = sum_plus<14>(sum, s[1]);
  s[2] = sum_plus<11>(sum, s[2]);
  s[3] = sum_plus<8>(sum, s[3]);
  s[4] = sum_plus<0>(sum, s[4]);
  s[5] = sum_plus<5>(sum, s[5]);
  s[6] = sum_plus<2>(sum, s[6]);
  s[7] = sum_plus<9>(sum, s[7]);
  s[8] = sum_plus<13>(sum, s[8]);
  s[9] = sum_plus<6>(sum, s[9]);
  s[10] = sum_plus<3>(sum, s[10]);
  s[11] = sum_plus<12>(sum, s[11]);
}

__device__ __forceinline__ u64 sbox(u64 x) {
  const u64 x2 = gl::sqr(x), x3 = gl::mul(x, x2), x4 = gl::sqr(x2);
  return gl::mul(x3, x4);
}

__device__ __forceinline__ void permute(u64 (&s)[WIDTH], const u64 *__restrict__ rc) {
  external_layer(s);
#pragma unroll
  for (int r = 0; r < ROUNDS; r++) {
    if (r < HALF_FULL_ROUNDS || r >= HALF_FULL_ROUNDS + PARTIAL_ROUNDS) {
#pragma unroll
      for (int i = 0; i < WIDTH; i++) s[i] = sbox(gl::add(s[i], rc[WIDTH * r + i]));
      external_layer(s);
    } else {
      s[0] = sbox(gl::add(s[0], rc[WIDTH * r]));
      internal_layer(s);
    }
  }
}

// Leaf gid absorbs values[gid * rows + r + c * rows * count] (c < cols, r < rows, column by
// column) RATE at a time into lanes 0..7 (overwrite mode, zero-filled past the last value) and
// permutes after each chunk; digest word k goes to results[gid + k * count].
__global__ __launch_bounds__(THREADS) void leaves_kernel(const u64 *__restrict__ values, u64 *__restrict__ results,
                                                         u32 rows, u32 cols, u32 count, int load_intermediate,
                                                         int store_intermediate, const u64 *__restrict__ rc) {
  const u32 gid = blockIdx.x * THREADS + threadIdx.x;
  if (gid >= count) return;
  u64 s[WIDTH];
#pragma unroll
  for (int i = 0; i < WIDTH; i++) s[i] = 0;
  if (load_intermediate) {
#pragma unroll
    for (int k = 0; k < CAPACITY; k++) s[RATE + k] = results[gid + (u64)k * count];
  }
  const u64 *leaf = values + (u64)gid * rows;
  const u64 col_stride = (u64)rows * count;
  const u64 total = (u64)rows * cols;
  u32 r = 0;  // (row, column) of the next value, the same for every lane
  u64 c = 0;
  for (u64 off = 0; off < total; off += RATE) {
#pragma unroll
    for (int i = 0; i < RATE; i++) {
      s[i] = c < cols ? gl::canon(leaf[r + c * col_stride]) : 0;
      if (++r == rows) r = 0, c++;
    }
    permute(s, rc);
  }
  const int from = store_intermediate ? RATE : 0;
#pragma unroll
  for (int k = 0; k < CAPACITY; k++) results[gid + (u64)k * count] = s[from + k];
}

// Node gid hashes its children's digests: word k of the left child at values[2 (gid + k count)],
// of the right child at values[2 (gid + k count) + 1] (the layer below, read as word pairs).
__global__ __launch_bounds__(THREADS) void nodes_kernel(const u64 *__restrict__ values, u64 *__restrict__ results,
                                                        u32 count, const u64 *__restrict__ rc) {
  const u32 gid = blockIdx.x * THREADS + threadIdx.x;
  if (gid >= count) return;
  u64 s[WIDTH];
#pragma unroll
  for (int k = 0; k < CAPACITY; k++) {
    const u64 at = 2 * (gid + (u64)k * count);
    s[k] = gl::canon(values[at]);
    s[CAPACITY + k] = gl::canon(values[at + 1]);
  }
#pragma unroll
  for (int i = 2 * CAPACITY; i < WIDTH; i++) s[i] = 0;
  permute(s, rc);
#pragma unroll
  for (int k = 0; k < CAPACITY; k++) results[gid + (u64)k * count] = s[k];
}

// The current device's context with installed round constants, or an error code.
int hasher_context(ba_context *&ctx) {
  ctx = ba_context_for_current_device();
  if (!ctx || !ctx->p2_round_constants) return BA_ERROR_NOT_INITIALIZED;
  return BA_SUCCESS;
}

int launch_nodes(const ba_context *ctx, const u64 *values, u64 *results, u32 count, hipStream_t stream) {
  if (count == 0) return BA_SUCCESS;
  nodes_kernel<<<(count + THREADS - 1) / THREADS, THREADS, 0, stream>>>(values, results, count,
                                                                        ctx->p2_round_constants);
  return (int)hipGetLastError();
}

// compute_node_hashes' preconditions (poseidon2.rs:231-239)
bool nodes_shape_ok(size_t values_len, size_t results_len) {
  return values_len % RATE == 0 && results_len % CAPACITY == 0 && values_len == 2 * results_len &&
         results_len / CAPACITY <= 0xFFFFFFFFull;
}

// build_merkle_tree_nodes' preconditions for every layer it will hash (poseidon2.rs:276-295)
bool node_layers_ok(size_t values_len, size_t results_len, uint32_t layers_count) {
  for (; layers_count; layers_count--) {
    if (values_len % RATE != 0 || values_len != results_len) return false;
    const size_t units = values_len / RATE;
    if (units == 0 || (units & (units - 1)) != 0) return false;  // values_len = RATE << layer
    if (!nodes_shape_ok(values_len, results_len >> 1)) return false;
    values_len = results_len >> 1;
    results_len -= results_len >> 1;
  }
  return true;
}

int run_node_layers(const ba_context *ctx, const u64 *values, size_t values_len, u64 *results, size_t results_len,
                    uint32_t layers_count, hipStream_t stream) {
  for (; layers_count; layers_count--) {
    const size_t half = results_len >> 1;
    const int rc = launch_nodes(ctx, values, results, (u32)(half / CAPACITY), stream);
    if (rc) return rc;
    values = results;
    values_len = half;
    results += half;
    results_len -= half;
  }
  (void)values_len;
  return BA_SUCCESS;
}

}  // namespace

extern "C" int ba_poseidon2_gl_set_round_constants(ba_context_t ctx, const uint64_t *round_constants) {
  return ba_guarded([&]() -> int {
    if (!ctx) return BA_ERROR_NOT_INITIALIZED;
    if (!round_constants) return BA_ERROR_INVALID_VALUE;
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device != ctx->device) return BA_ERROR_INVALID_VALUE;
    std::vector<u64> host(round_constants, round_constants + ROUNDS * WIDTH);
    for (u64 &x : host) x = gl::canon(x);
    if (!ctx->p2_round_constants) {
      void *p = nullptr;
      const hipError_t err = hipMalloc(&p, host.size() * sizeof(u64));
      if (err != hipSuccess) return err == hipErrorOutOfMemory ? BA_ERROR_OUT_OF_MEMORY : (int)err;
      ctx->p2_round_constants = (u64 *)p;
    }
    return (int)hipMemcpy(ctx->p2_round_constants, host.data(), host.size() * sizeof(u64), hipMemcpyHostToDevice);
  });
}

extern "C" int ba_poseidon2_gl_compute_leaf_hashes(const uint64_t *values, size_t values_len, uint64_t *results,
                                                   size_t results_len, uint32_t log_rows_per_hash,
                                                   int load_intermediate, int store_intermediate, ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    ba_context *ctx = nullptr;
    if (const int rc = hasher_context(ctx)) return rc;
    // poseidon2.rs:186-207
    if (results_len % CAPACITY != 0 || log_rows_per_hash >= 32) return BA_ERROR_INVALID_VALUE;
    const size_t count = results_len / CAPACITY;
    if (count == 0) return BA_ERROR_INVALID_VALUE;
    const size_t rows = (size_t)1 << log_rows_per_hash;
    if (values_len % (count * rows) != 0) return BA_ERROR_INVALID_VALUE;
    const size_t cols = values_len / (count * rows);
    if (store_intermediate && (rows * cols) % RATE != 0) return BA_ERROR_INVALID_VALUE;
    if (cols > 0xFFFFFFFFull || count > 0xFFFFFFFFull || (values_len && !values) || !results)
      return BA_ERROR_INVALID_VALUE;
    leaves_kernel<<<(u32)((count + THREADS - 1) / THREADS), THREADS, 0, (hipStream_t)stream>>>(
        values, results, (u32)rows, (u32)cols, (u32)count, load_intermediate, store_intermediate,
        ctx->p2_round_constants);
    return (int)hipGetLastError();
  });
}

extern "C" int ba_poseidon2_gl_compute_node_hashes(const uint64_t *values, size_t values_len, uint64_t *results,
                                                   size_t results_len, ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    ba_context *ctx = nullptr;
    if (const int rc = hasher_context(ctx)) return rc;
    if (!nodes_shape_ok(values_len, results_len) || (results_len && (!values || !results)))
      return BA_ERROR_INVALID_VALUE;
    return launch_nodes(ctx, values, results, (u32)(results_len / CAPACITY), (hipStream_t)stream);
  });
}

extern "C" int ba_poseidon2_gl_build_merkle_tree_leaves(const uint64_t *values, size_t values_len, uint64_t *results,
                                                        size_t results_len, uint32_t log_rows_per_hash,
                                                        int load_intermediate, int store_intermediate,
                                                        ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    // poseidon2.rs:253-274: the leaves count divides the values
    if (results_len % CAPACITY != 0 || results_len == 0 || values_len % (results_len / CAPACITY) != 0)
      return BA_ERROR_INVALID_VALUE;
    return ba_poseidon2_gl_compute_leaf_hashes(values, values_len, results, results_len, log_rows_per_hash,
                                               load_intermediate, store_intermediate, stream);
  });
}

extern "C" int ba_poseidon2_gl_build_merkle_tree_nodes(const uint64_t *values, size_t values_len, uint64_t *results,
                                                       size_t results_len, uint32_t layers_count, ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    ba_context *ctx = nullptr;
    if (const int rc = hasher_context(ctx)) return rc;
    if (layers_count == 0) return BA_SUCCESS;
    if (!node_layers_ok(values_len, results_len, layers_count) || !values || !results) return BA_ERROR_INVALID_VALUE;
    return run_node_layers(ctx, values, values_len, results, results_len, layers_count, (hipStream_t)stream);
  });
}

extern "C" int ba_poseidon2_gl_build_merkle_tree(const uint64_t *values, size_t values_len, uint64_t *results,
                                                 size_t results_len, uint32_t log_rows_per_hash, uint32_t layers_count,
                                                 ba_stream_t stream) {
  return ba_guarded([&]() -> int {
    ba_context *ctx = nullptr;
    if (const int rc = hasher_context(ctx)) return rc;
    // poseidon2.rs:297-316
    if (layers_count == 0 || layers_count > 40 || results_len % (2 * CAPACITY) != 0) return BA_ERROR_INVALID_VALUE;
    const size_t leaves_count = results_len / (2 * CAPACITY);
    if (leaves_count == 0 || ((size_t)1 << (layers_count - 1)) > leaves_count || values_len % leaves_count != 0)
      return BA_ERROR_INVALID_VALUE;
    const size_t half = results_len >> 1;
    // every node layer is checked before the leaves launch, so a bad shape launches nothing
    if (!node_layers_ok(half, results_len - half, layers_count - 1)) return BA_ERROR_INVALID_VALUE;
    const int rc = ba_poseidon2_gl_compute_leaf_hashes(values, values_len, results, half, log_rows_per_hash, 0, 0, stream);
    if (rc) return rc;
    return run_node_layers(ctx, results, half, results + half, results_len - half, layers_count - 1, (hipStream_t)stream);
  });
}
