/*
 * gl_oracle.h -- CPU oracle for the Goldilocks NTT / coset-LDE path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the CPU NTT that
 * the reference checks its GPU kernels against (boojum's fft_natural_to_bitreversed /
 * ifft_natural_to_natural, crates.io `boojum = "=0.32.10"`, pinned at Cargo.toml:32,
 * not vendored in the reference tree).  It is linked only by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as the checker
 * or as the timed CPU baseline, never as the product path.
 *
 * Conventions restated (reference call sites):
 *   - field p = 2^64 - 2^32 + 1 (crates/boojum-cuda/native/goldilocks.cuh:23)
 *   - omega_{2^k} = domain_generator_for_size(2^k): the 2^32-th root 7^((p-1)/2^32)
 *     squared down (crates/boojum-cuda/src/context.rs:179-198)
 *   - twiddles = precompute_twiddles_for_fft: powers of omega_n, first n/2, bit-reversed
 *     (crates/boojum-cuda/src/ntt.rs:404-416)
 *   - batch semantics of batch_ntt_internal as pinned by its test
 *     (crates/boojum-cuda/src/ntt.rs:391-611)
 *   - LDE composition of shivini (crates/shivini/src/data_structures/storage.rs:154-260,
 *     crates/shivini/src/primitives/ntt.rs:10-42, crates/shivini/src/utils.rs:5-16)
 *
 * All outputs are canonical (< p).  Inputs may be any u64 (values >= p are read mod p,
 * as the reference test feeds raw rng u64 at ntt.rs:424).
 */
#ifndef GL_ORACLE_H
#define GL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GL_P 0xFFFFFFFF00000001ull
#define GL_OMEGA_LOG_ORDER 30u

/* field arithmetic (canonical in, canonical out; gl_reduce accepts any u64) */
uint64_t gl_reduce(uint64_t x);
uint64_t gl_add(uint64_t a, uint64_t b);
uint64_t gl_sub(uint64_t a, uint64_t b);
uint64_t gl_mul(uint64_t a, uint64_t b);
uint64_t gl_pow(uint64_t a, uint64_t e);
uint64_t gl_inv(uint64_t a);

/* domain_generator_for_size(2^log_size), log_size <= 32 */
uint64_t gl_omega(unsigned log_size);
/* multiplicative generator 7 (GoldilocksField::multiplicative_generator) */
uint64_t gl_generator(void);

size_t gl_bitrev(size_t i, unsigned bits);
void gl_bitreverse_enumeration(uint64_t *v, unsigned log_n);
void gl_distribute_powers(uint64_t *v, size_t n, uint64_t base);
/* n/2 twiddles: tw[k] = omega_n^{+-bitrev_{log_n-1}(k)} */
void gl_precompute_twiddles(uint64_t *tw, unsigned log_n, int inverse);
/* radix-2 Cooley-Tukey, natural in, bit-reversed out, bit-reversed twiddles */
void gl_serial_ct_n2b(uint64_t *a, unsigned log_n, const uint64_t *tw);
void gl_fft_natural_to_bitreversed(uint64_t *v, unsigned log_n, uint64_t coset, const uint64_t *tw);
void gl_ifft_natural_to_natural(uint64_t *v, unsigned log_n, uint64_t coset, const uint64_t *tw);

/* coset shift s used by batch_ntt_internal: 1 if log_ext == 0 else 7 * omega_{n*2^e}^c */
uint64_t gl_coset_shift(unsigned log_n, unsigned log_ext, unsigned coset_idx);

/*
 * batch_ntt_internal semantics (crates/boojum-cuda/src/ntt.rs:179-191), column k read at
 * in + k*stride_in, written at out + k*stride_out (in == out allowed: in place).
 *   !bitrev_inputs && !inverse: out[bitrev(k)] = sum_j s^j x_j w^jk
 *    bitrev_inputs && !inverse: in is x bit-reversed, out natural
 *   !bitrev_inputs &&  inverse: out[bitrev(j)] = n^-1 s^-j sum_k X_k w^-jk
 *    bitrev_inputs &&  inverse: in is X bit-reversed, out natural
 * Returns 0, or -1 for arguments the reference rejects with assert!.
 * nthreads > 1 splits the columns over pthreads.
 */
int gl_oracle_batch_ntt(const uint64_t *in, uint64_t *out, unsigned log_n, unsigned num_ntts,
                        uint64_t stride_in, uint64_t stride_out, int bitrev_inputs, int inverse,
                        unsigned log_ext, unsigned coset_idx, int nthreads);

/*
 * Vectors of powers and omega shifts, boojum_cuda::ops_complex (crates/boojum-cuda/src/
 * ops_complex.rs:31-260; kernels crates/boojum-cuda/native/ops_complex.cu:8-86), u32 index
 * arithmetic wrapping as there.  Computed with gl_pow from the definitions, no tables.
 *   powers_of_w: r[i] = omega_{2^30}^(+-(j << (30 - log_degree)) mod 2^30),
 *                j = bit_reverse ? bitrev_{log_degree}(low log_degree bits of i + offset) : i + offset
 *   powers_of_g: r[i] = 7^(that exponent mod 2^30), or (7^-1)^(...) when inverse
 *   powers_by_val_bf: r[i] = base^((bit_reverse ? brev32(i) : i) + offset mod 2^32)
 *   powers_by_val_ef: the same in F_p[u]/(u^2 - 7); r[i] = c0, r[count + i] = c1
 *   omega_shift: dst[i] = src[i] * omega_{2^30}^(+-((shift * (i + offset)) << (30 - log_degree)))
 */
void gl_oracle_powers_of_w(unsigned log_degree, uint32_t offset, int inverse, int bit_reverse,
                           uint64_t *result, uint32_t count);
void gl_oracle_powers_of_g(unsigned log_degree, uint32_t offset, int inverse, int bit_reverse,
                           uint64_t *result, uint32_t count);
void gl_oracle_powers_by_val_bf(uint64_t base, uint32_t offset, int bit_reverse, uint64_t *result,
                                uint32_t count);
void gl_oracle_powers_by_val_ef(uint64_t base_c0, uint64_t base_c1, uint32_t offset, int bit_reverse,
                                uint64_t *result, uint32_t count);
void gl_oracle_omega_shift(const uint64_t *src, unsigned log_degree, uint32_t offset, int inverse,
                           uint32_t shift, uint64_t *dst, uint32_t count);

/*
 * Barycentric evaluation (crates/boojum-cuda/src/barycentric.rs:83-300, kernels
 * crates/boojum-cuda/native/barycentric.cu:11-158), from the formula with one inversion per
 * element: L_i = coset (X^N - coset^N) / (N coset^N) / (X w^-i - coset), w = omega_N; ext selects
 * X = x0 + x1 u in F_p[u]/(u^2 - 7) and two-plane coefficients; bit_reverse permutes them.
 * eval kind 0: base ys at base X; 1: base ys at extension X; 2: extension ys (column k at
 * 2 k count, planes c0 and c1) at extension X.  evals: kind 0 one element per column, else
 * (c0, c1) interleaved.
 */
void gl_oracle_barycentric_coeffs(int ext, uint64_t x0, uint64_t x1, uint64_t coset, uint32_t count,
                                  int bit_reverse, uint64_t *coeffs);
void gl_oracle_barycentric_eval(int kind, const uint64_t *ys, uint32_t count, uint32_t num_polys,
                                const uint64_t *coeffs, uint64_t *evals);

/* per-column bit reversal (crates/boojum-cuda/src/ops_complex.rs:262-331) */
void gl_oracle_bit_reverse(const uint64_t *src, uint64_t src_stride, uint64_t *dst,
                           uint64_t dst_stride, unsigned log_rows, unsigned cols);

/*
 * shivini's trace-commitment LDE: evaluations (natural) -> monomials (iNTT + bit reverse)
 * -> for each coset c < 2^log_blowup: forward n2b NTT with coset index bitrev(c) and
 * log_extension_degree = log_blowup.  Coset c of column k is written bit-reversed at
 * out + k*out_col_stride + c*out_coset_stride.  With out_col_stride = n<<log_blowup and
 * out_coset_stride = n each column holds its whole LDE domain in bit-reversed order.
 */
int gl_oracle_lde(const uint64_t *evals, uint64_t in_stride, uint64_t *out,
                  uint64_t out_col_stride, uint64_t out_coset_stride, unsigned log_n,
                  unsigned num_cols, unsigned log_blowup, int nthreads);

/* SplitMix64 with rejection of x >= p: the seeded input generator of the benchmarks */
void gl_fill_uniform(uint64_t *v, size_t n, uint64_t seed);

/*
 * Poseidon2 over Goldilocks, width 12 (rate 8, capacity 4): Merkle leaf and node hashes
 * (crates/boojum-cuda/src/poseidon2.rs:165-300, kernels
 * crates/boojum-cuda/native/poseidon2/gl/poseidon2_gl_st.cu:24-88).  The permutation follows
 * the Poseidon2 paper (eprint 2023/323, Fig. 1 and Appendix B) written as matrix products:
 * M_E = circ(2 M4, M4, M4) with M4 = [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]], applied once
 * first; 4 full rounds (constants on all 12 lanes, x^7 on all, M_E), 22 partial rounds
 * (constant and x^7 on lane 0, M_I = J + diag(2^mu) with mu = {4, 14, 11, 8, 0, 5, 2, 9, 13, 6,
 * 3, 12}, poseidon2_gl_constants.cuh.template:13), 4 full rounds.  The 30 x 12 round constants
 * are an argument: boojum's table (poseidon_goldilocks_params) is not in the reference tree.
 */
void gl_oracle_poseidon2_permutation(uint64_t state[12], const uint64_t *round_constants);
/* the linear layers alone: M_E x (internal = 0) or M_I x (internal = 1), in place */
void gl_oracle_poseidon2_matrix(int internal, uint64_t x[12]);
/* leaf gid absorbs, 8 at a time in overwrite mode (zero-filled past the end), the values
 * values[gid * rows + r + c * rows * count] for c < cols, r < rows, column by column; digest
 * word k goes to results[gid + k * count] (state[k], or the capacity state[8 + k] when
 * store_intermediate; load_intermediate starts from a stored capacity) */
void gl_oracle_poseidon2_leaves(const uint64_t *values, uint64_t *results, uint32_t rows_count,
                                uint32_t cols_count, uint32_t count, int load_intermediate,
                                int store_intermediate, const uint64_t *round_constants);
/* node gid hashes the two child digests of the layer below: word k of the left child at
 * values[2 (gid + k * count)], of the right child at values[2 (gid + k * count) + 1] */
void gl_oracle_poseidon2_nodes(const uint64_t *values, uint64_t *results, uint32_t count,
                               const uint64_t *round_constants);

#ifdef __cplusplus
}
#endif
#endif
