/*
 * gl_oracle.c -- CPU oracle for the Goldilocks NTT / coset-LDE path.
 *
 * TEST INFRASTRUCTURE ONLY: linked by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg, as the checker or the timed CPU baseline.  Never on the product path.
 *
 * Restates, from the field's mathematics, boojum's CPU FFT (crates.io boojum =0.32.10,
 * Cargo.toml:32; absent from the reference tree) exactly as the reference's own test
 * calls it (crates/boojum-cuda/src/ntt.rs:391-611).  Parity of each function is pinned by
 * tests/test_oracle.py (closed-form roots, an independent big-integer DFT, transform
 * invariants and direct polynomial evaluation on coset points).
 */
#include "gl_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;


uint64_t gl_reduce(uint64_t x) { return x >= GL_P ? x - GL_P : x; }

/* 2^64 = 2^32 - 1 and 2^96 = -1 (mod p) */
#define GL_EPS 0xFFFFFFFFull

uint64_t gl_add(uint64_t a, uint64_t b) {
  a = gl_reduce(a);
  b = gl_reduce(b);
  uint64_t s = a + b;
  if (s < a) return s + GL_EPS; /* a + b - 2^64 + (2^64 - p) < p */
  return gl_reduce(s);
}

uint64_t gl_sub(uint64_t a, uint64_t b) {
  a = gl_reduce(a);
  b = gl_reduce(b);
  return a >= b ? a - b : a + (GL_P - b);
}

/* lo + hi 2^64 = lo - hh + hl (2^32 - 1) with hi = hh 2^32 + hl */
static inline uint64_t reduce128(u128 x) {
  uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  uint64_t hh = hi >> 32, hl = hi & GL_EPS;
  uint64_t t0 = lo - hh;
  if (lo < hh) t0 -= GL_EPS;
  uint64_t t1 = hl * GL_EPS;
  uint64_t r = t0 + t1;
  if (r < t1) r += GL_EPS;
  return gl_reduce(r);
}

uint64_t gl_mul(uint64_t a, uint64_t b) { return reduce128((u128)a * b); }

uint64_t gl_pow(uint64_t a, uint64_t e) {
  uint64_t r = 1, x = gl_reduce(a);
  while (e) {
    if (e & 1) r = gl_mul(r, x);
    x = gl_mul(x, x);
    e >>= 1;
  }
  return r;
}

uint64_t gl_inv(uint64_t a) { return gl_pow(a, GL_P - 2); }

uint64_t gl_generator(void) { return 7; }

/* domain_generator_for_size: the 2^32-th root of unity 7^((p-1)/2^32) squared
 * (32 - log_size) times (crates/boojum-cuda/src/context.rs:179-198 call site). */
uint64_t gl_omega(unsigned log_size) {
  uint64_t w = gl_pow(7, (GL_P - 1) >> 32);
  for (unsigned i = log_size; i < 32; i++) w = gl_mul(w, w);
  return w;
}

size_t gl_bitrev(size_t i, unsigned bits) {
  size_t r = 0;
  for (unsigned b = 0; b < bits; b++) r |= ((i >> b) & 1) << (bits - 1 - b);
  return r;
}

void gl_bitreverse_enumeration(uint64_t *v, unsigned log_n) {
  size_t n = (size_t)1 << log_n;
  for (size_t i = 0; i < n; i++) {
    size_t j = gl_bitrev(i, log_n);
    if (i < j) {
      uint64_t t = v[i];
      v[i] = v[j];
      v[j] = t;
    }
  }
}

void gl_distribute_powers(uint64_t *v, size_t n, uint64_t base) {
  uint64_t acc = 1;
  for (size_t i = 0; i < n; i++) {
    v[i] = gl_mul(gl_reduce(v[i]), acc);
    acc = gl_mul(acc, base);
  }
}

/* precompute_twiddles_for_fft (ntt.rs:404-416 call site): the first n/2 powers of
 * omega_n (or its inverse), in bit-reversed order over log_n - 1 bits. */
void gl_precompute_twiddles(uint64_t *tw, unsigned log_n, int inverse) {
  if (log_n == 0) return;
  size_t half = (size_t)1 << (log_n - 1);
  uint64_t w = gl_omega(log_n);
  if (inverse) w = gl_inv(w);
  uint64_t acc = 1;
  for (size_t i = 0; i < half; i++) {
    tw[i] = acc;
    acc = gl_mul(acc, w);
  }
  gl_bitreverse_enumeration(tw, log_n - 1);
}

/* Radix-2 Cooley-Tukey, natural order in, bit-reversed order out: at stage s the 2^s
 * groups of size n/2^s use twiddle tw[group] (bit-reversed table). */
void gl_serial_ct_n2b(uint64_t *a, unsigned log_n, const uint64_t *tw) {
  size_t n = (size_t)1 << log_n;
  for (size_t i = 0; i < n; i++) a[i] = gl_reduce(a[i]);
  size_t groups = 1, dist = n >> 1;
  while (groups < n) {
    for (size_t k = 0; k < groups; k++) {
      size_t start = k * dist * 2;
      uint64_t w = tw[k];
      for (size_t j = start; j < start + dist; j++) {
        uint64_t u = a[j];
        uint64_t v = gl_mul(a[j + dist], w);
        a[j] = gl_add(u, v);
        a[j + dist] = gl_sub(u, v);
      }
    }
    groups <<= 1;
    dist >>= 1;
  }
}

void gl_fft_natural_to_bitreversed(uint64_t *v, unsigned log_n, uint64_t coset, const uint64_t *tw) {
  size_t n = (size_t)1 << log_n;
  if (coset != 1) gl_distribute_powers(v, n, coset);
  gl_serial_ct_n2b(v, log_n, tw);
}

void gl_ifft_natural_to_natural(uint64_t *v, unsigned log_n, uint64_t coset, const uint64_t *tw) {
  size_t n = (size_t)1 << log_n;
  gl_serial_ct_n2b(v, log_n, tw);
  gl_bitreverse_enumeration(v, log_n);
  if (coset != 1) gl_distribute_powers(v, n, gl_inv(coset));
  if (n > 1) {
    uint64_t n_inv = gl_inv(n % GL_P);
    for (size_t i = 0; i < n; i++) v[i] = gl_mul(v[i], n_inv);
  }
}

/* coset = g * omega_{n*2^e}^c when log_ext > 0 (ntt.rs:527-537) */
uint64_t gl_coset_shift(unsigned log_n, unsigned log_ext, unsigned coset_idx) {
  if (log_ext == 0) return 1;
  uint64_t w = gl_omega(log_n + log_ext);
  return gl_mul(gl_generator(), gl_pow(w, coset_idx));
}

/* ---------------------------------------------------------------- batch driver */

typedef struct {
  const uint64_t *in;
  uint64_t *out;
  unsigned log_n;
  uint64_t stride_in, stride_out;
  int bitrev_inputs, inverse;
  uint64_t coset;
  const uint64_t *tw;
  unsigned col_begin, col_end;
} ntt_job;

static void *ntt_worker(void *arg) {
  ntt_job *j = (ntt_job *)arg;
  size_t n = (size_t)1 << j->log_n;
  uint64_t *buf = (uint64_t *)malloc(n * sizeof(uint64_t));
  for (unsigned c = j->col_begin; c < j->col_end; c++) {
    const uint64_t *src = j->in + (size_t)c * j->stride_in;
    memcpy(buf, src, n * sizeof(uint64_t));
    /* b2n takes the transform input bit-reversed (ntt.rs:441-447, 577-583) */
    if (j->bitrev_inputs) gl_bitreverse_enumeration(buf, j->log_n);
    if (j->inverse) {
      gl_ifft_natural_to_natural(buf, j->log_n, j->coset, j->tw);
      /* n2b inverse returns bit-reversed order (ntt.rs:577-578) */
      if (!j->bitrev_inputs) gl_bitreverse_enumeration(buf, j->log_n);
    } else {
      gl_fft_natural_to_bitreversed(buf, j->log_n, j->coset, j->tw);
      /* b2n forward returns natural order (ntt.rs:581-582) */
      if (j->bitrev_inputs) gl_bitreverse_enumeration(buf, j->log_n);
    }
    memcpy(j->out + (size_t)c * j->stride_out, buf, n * sizeof(uint64_t));
  }
  free(buf);
  return NULL;
}

static void run_jobs(void *(*fn)(void *), void *jobs, size_t job_size, int nthreads) {
  if (nthreads <= 1) {
    fn(jobs);
    return;
  }
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, fn, (char *)jobs + (size_t)t * job_size);
  for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  free(th);
}

int gl_oracle_batch_ntt(const uint64_t *in, uint64_t *out, unsigned log_n, unsigned num_ntts,
                        uint64_t stride_in, uint64_t stride_out, int bitrev_inputs, int inverse,
                        unsigned log_ext, unsigned coset_idx, int nthreads) {
  /* argument checks of batch_ntt_internal (ntt.rs:192-200) */
  if (log_n < 1 || log_n + log_ext > GL_OMEGA_LOG_ORDER) return -1;
  uint64_t n = (uint64_t)1 << log_n;
  if (n > stride_in || n > stride_out) return -1;
  if (coset_idx > 0 && (log_ext == 0 || coset_idx >= (1u << log_ext))) return -1;
  if (num_ntts == 0) return 0;
  uint64_t *tw = (uint64_t *)malloc(sizeof(uint64_t) * (n / 2));
  gl_precompute_twiddles(tw, log_n, inverse);
  uint64_t coset = gl_coset_shift(log_n, log_ext, coset_idx);
  if (nthreads < 1) nthreads = 1;
  if ((unsigned)nthreads > num_ntts) nthreads = (int)num_ntts;
  ntt_job *jobs = (ntt_job *)calloc((size_t)nthreads, sizeof(ntt_job));
  for (int t = 0; t < nthreads; t++) {
    ntt_job *j = &jobs[t];
    j->in = in;
    j->out = out;
    j->log_n = log_n;
    j->stride_in = stride_in;
    j->stride_out = stride_out;
    j->bitrev_inputs = bitrev_inputs;
    j->inverse = inverse;
    j->coset = coset;
    j->tw = tw;
    j->col_begin = (unsigned)((uint64_t)num_ntts * t / nthreads);
    j->col_end = (unsigned)((uint64_t)num_ntts * (t + 1) / nthreads);
  }
  run_jobs(ntt_worker, jobs, sizeof(ntt_job), nthreads);
  free(jobs);
  free(tw);
  return 0;
}

/* ---- powers and omega shifts (crates/boojum-cuda/native/ops_complex.cu:8-86 semantics) ---- */

/* j = bitrev over the low `bits` bits of x (x >> 32-bits of the 32-bit reversal; 0 for bits 0) */
static uint32_t rev_low(uint32_t x, unsigned bits) { return bits ? (uint32_t)gl_bitrev(x & (uint32_t)((1ull << bits) - 1), bits) : 0; }

static uint32_t brev32(uint32_t x) { return (uint32_t)gl_bitrev(x, 32); }

void gl_oracle_powers_of_w(unsigned log_degree, uint32_t offset, int inverse, int bit_reverse,
                           uint64_t *result, uint32_t count) {
  const uint64_t w = gl_omega(GL_OMEGA_LOG_ORDER);
  for (uint32_t i = 0; i < count; i++) {
    const uint32_t raw = i + offset;
    const uint32_t j = bit_reverse ? rev_low(raw, log_degree) : raw;
    uint32_t e = (j << (GL_OMEGA_LOG_ORDER - log_degree)) & ((1u << GL_OMEGA_LOG_ORDER) - 1);
    if (inverse) e = (0u - e) & ((1u << GL_OMEGA_LOG_ORDER) - 1);
    result[i] = gl_pow(w, e);
  }
}

void gl_oracle_powers_of_g(unsigned log_degree, uint32_t offset, int inverse, int bit_reverse,
                           uint64_t *result, uint32_t count) {
  const uint64_t g = inverse ? gl_inv(gl_generator()) : gl_generator();
  for (uint32_t i = 0; i < count; i++) {
    const uint32_t raw = i + offset;
    const uint32_t j = bit_reverse ? rev_low(raw, log_degree) : raw;
    const uint32_t e = (j << (GL_OMEGA_LOG_ORDER - log_degree)) & ((1u << GL_OMEGA_LOG_ORDER) - 1);
    result[i] = gl_pow(g, e);
  }
}

void gl_oracle_powers_by_val_bf(uint64_t base, uint32_t offset, int bit_reverse, uint64_t *result,
                                uint32_t count) {
  for (uint32_t i = 0; i < count; i++) result[i] = gl_pow(base, (uint32_t)((bit_reverse ? brev32(i) : i) + offset));
}

/* F_p[u] / (u^2 - 7) */
static void ext_mul(const uint64_t x[2], const uint64_t y[2], uint64_t r[2]) {
  const uint64_t a = gl_add(gl_mul(x[0], y[0]), gl_mul(7, gl_mul(x[1], y[1])));
  const uint64_t b = gl_add(gl_mul(x[0], y[1]), gl_mul(x[1], y[0]));
  r[0] = a;
  r[1] = b;
}

void gl_oracle_powers_by_val_ef(uint64_t base_c0, uint64_t base_c1, uint32_t offset, int bit_reverse,
                                uint64_t *result, uint32_t count) {
  for (uint32_t i = 0; i < count; i++) {
    uint32_t e = (bit_reverse ? brev32(i) : i) + offset;
    uint64_t r[2] = {1, 0}, x[2] = {gl_reduce(base_c0), gl_reduce(base_c1)};
    while (e) {
      if (e & 1) ext_mul(r, x, r);
      ext_mul(x, x, x);
      e >>= 1;
    }
    result[i] = r[0];
    result[(size_t)count + i] = r[1];
  }
}

void gl_oracle_omega_shift(const uint64_t *src, unsigned log_degree, uint32_t offset, int inverse,
                           uint32_t shift, uint64_t *dst, uint32_t count) {
  const uint64_t w = gl_omega(GL_OMEGA_LOG_ORDER);
  for (uint32_t i = 0; i < count; i++) {
    uint32_t e = ((shift * (i + offset)) << (GL_OMEGA_LOG_ORDER - log_degree)) & ((1u << GL_OMEGA_LOG_ORDER) - 1);
    if (inverse) e = (0u - e) & ((1u << GL_OMEGA_LOG_ORDER) - 1);
    dst[i] = gl_mul(gl_reduce(src[i]), gl_pow(w, e));
  }
}

/* ---- barycentric evaluation (crates/boojum-cuda/native/barycentric.cu:11-158 semantics) ---- */

static void ext_inv(const uint64_t x[2], uint64_t r[2]) {
  const uint64_t norm = gl_sub(gl_mul(x[0], x[0]), gl_mul(7, gl_mul(x[1], x[1])));
  const uint64_t ni = gl_inv(norm);
  r[0] = gl_mul(x[0], ni);
  r[1] = gl_sub(0, gl_mul(x[1], ni));
}

static void ext_pow_u64(const uint64_t x[2], uint64_t e, uint64_t r[2]) {
  uint64_t a[2] = {1, 0}, b[2] = {x[0], x[1]};
  while (e) {
    if (e & 1) ext_mul(a, b, a);
    ext_mul(b, b, b);
    e >>= 1;
  }
  r[0] = a[0];
  r[1] = a[1];
}

void gl_oracle_barycentric_coeffs(int ext, uint64_t x0, uint64_t x1, uint64_t coset, uint32_t count,
                                  int bit_reverse, uint64_t *coeffs) {
  unsigned log_count = 0;
  while ((1ull << log_count) < count) log_count++;
  const uint64_t c = gl_reduce(coset);
  const uint64_t x[2] = {gl_reduce(x0), ext ? gl_reduce(x1) : 0};
  /* common = c (X^N - c^N) / (N c^N) */
  const uint64_t cn = gl_pow(c, count);
  uint64_t xn[2];
  ext_pow_u64(x, count, xn);
  const uint64_t scale = gl_mul(c, gl_inv(gl_mul(count, cn)));
  const uint64_t common[2] = {gl_mul(gl_sub(xn[0], cn), scale), gl_mul(xn[1], scale)};
  const uint64_t w_inv = gl_inv(gl_omega(log_count));
  uint64_t wi = 1; /* w^-i */
  for (uint32_t i = 0; i < count; i++) {
    const uint64_t den[2] = {gl_sub(gl_mul(x[0], wi), c), gl_mul(x[1], wi)};
    uint64_t inv[2], l[2];
    ext_inv(den, inv);
    ext_mul(common, inv, l);
    const uint32_t at = bit_reverse ? (uint32_t)gl_bitrev(i, log_count) : i;
    coeffs[at] = l[0];
    if (ext) coeffs[(size_t)count + at] = l[1];
    wi = gl_mul(wi, w_inv);
  }
}

void gl_oracle_barycentric_eval(int kind, const uint64_t *ys, uint32_t count, uint32_t num_polys,
                                const uint64_t *coeffs, uint64_t *evals) {
  for (uint32_t k = 0; k < num_polys; k++) {
    uint64_t s[2] = {0, 0};
    for (uint32_t i = 0; i < count; i++) {
      const uint64_t l[2] = {coeffs[i], kind ? coeffs[(size_t)count + i] : 0};
      uint64_t y[2];
      if (kind == 2) {
        y[0] = gl_reduce(ys[(size_t)k * 2 * count + i]);
        y[1] = gl_reduce(ys[(size_t)k * 2 * count + count + i]);
      } else {
        y[0] = gl_reduce(ys[(size_t)k * count + i]);
        y[1] = 0;
      }
      uint64_t t[2];
      ext_mul(l, y, t);
      s[0] = gl_add(s[0], t[0]);
      s[1] = gl_add(s[1], t[1]);
    }
    if (kind == 0) {
      evals[k] = s[0];
    } else {
      evals[2 * (size_t)k] = s[0];
      evals[2 * (size_t)k + 1] = s[1];
    }
  }
}

void gl_oracle_bit_reverse(const uint64_t *src, uint64_t src_stride, uint64_t *dst,
                           uint64_t dst_stride, unsigned log_rows, unsigned cols) {
  size_t n = (size_t)1 << log_rows;
  uint64_t *buf = (uint64_t *)malloc(n * sizeof(uint64_t));
  for (unsigned c = 0; c < cols; c++) {
    const uint64_t *s = src + (size_t)c * src_stride;
    for (size_t i = 0; i < n; i++) buf[gl_bitrev(i, log_rows)] = s[i];
    memcpy(dst + (size_t)c * dst_stride, buf, n * sizeof(uint64_t));
  }
  free(buf);
}

/* ---------------------------------------------------------------- LDE */

typedef struct {
  const uint64_t *evals;
  uint64_t in_stride;
  uint64_t *out;
  uint64_t out_col_stride, out_coset_stride;
  unsigned log_n, log_blowup;
  const uint64_t *tw_inv, *tw_fwd;
  unsigned col_begin, col_end;
} lde_job;

static void *lde_worker(void *arg) {
  lde_job *j = (lde_job *)arg;
  size_t n = (size_t)1 << j->log_n;
  unsigned blowup = 1u << j->log_blowup;
  uint64_t *mono = (uint64_t *)malloc(n * sizeof(uint64_t));
  uint64_t *buf = (uint64_t *)malloc(n * sizeof(uint64_t));
  for (unsigned k = j->col_begin; k < j->col_end; k++) {
    /* into_monomials: n2b inverse NTT then bit reverse (storage.rs:155-169) */
    memcpy(mono, j->evals + (size_t)k * j->in_stride, n * sizeof(uint64_t));
    gl_ifft_natural_to_natural(mono, j->log_n, 1, j->tw_inv);
    for (unsigned c = 0; c < blowup; c++) {
      /* raw_batch_coset_ntt bit-reverses the coset index (primitives/ntt.rs:25) */
      unsigned cidx = (unsigned)gl_bitrev(c, j->log_blowup);
      uint64_t s = gl_coset_shift(j->log_n, j->log_blowup, cidx);
      memcpy(buf, mono, n * sizeof(uint64_t));
      gl_fft_natural_to_bitreversed(buf, j->log_n, s, j->tw_fwd);
      memcpy(j->out + (size_t)k * j->out_col_stride + (size_t)c * j->out_coset_stride, buf,
             n * sizeof(uint64_t));
    }
  }
  free(buf);
  free(mono);
  return NULL;
}

int gl_oracle_lde(const uint64_t *evals, uint64_t in_stride, uint64_t *out,
                  uint64_t out_col_stride, uint64_t out_coset_stride, unsigned log_n,
                  unsigned num_cols, unsigned log_blowup, int nthreads) {
  if (log_n < 1 || log_blowup < 1 || log_n + log_blowup > GL_OMEGA_LOG_ORDER) return -1;
  if (num_cols == 0) return 0;
  size_t n = (size_t)1 << log_n;
  uint64_t *tw_inv = (uint64_t *)malloc(sizeof(uint64_t) * (n / 2));
  uint64_t *tw_fwd = (uint64_t *)malloc(sizeof(uint64_t) * (n / 2));
  gl_precompute_twiddles(tw_inv, log_n, 1);
  gl_precompute_twiddles(tw_fwd, log_n, 0);
  if (nthreads < 1) nthreads = 1;
  if ((unsigned)nthreads > num_cols) nthreads = (int)num_cols;
  lde_job *jobs = (lde_job *)calloc((size_t)nthreads, sizeof(lde_job));
  for (int t = 0; t < nthreads; t++) {
    lde_job *j = &jobs[t];
    j->evals = evals;
    j->in_stride = in_stride;
    j->out = out;
    j->out_col_stride = out_col_stride;
    j->out_coset_stride = out_coset_stride;
    j->log_n = log_n;
    j->log_blowup = log_blowup;
    j->tw_inv = tw_inv;
    j->tw_fwd = tw_fwd;
    j->col_begin = (unsigned)((uint64_t)num_cols * t / nthreads);
    j->col_end = (unsigned)((uint64_t)num_cols * (t + 1) / nthreads);
  }
  run_jobs(lde_worker, jobs, sizeof(lde_job), nthreads);
  free(jobs);
  free(tw_fwd);
  free(tw_inv);
  return 0;
}

/* ---------------------------------------------------------------- inputs */

/* Counter-based SplitMix64: element i is the SplitMix64 finaliser of seed + (i+1)*gamma;
 * a value >= p is re-mixed until it is below p. */
static uint64_t splitmix_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void gl_fill_uniform(uint64_t *v, size_t n, uint64_t seed) {
  const uint64_t gamma = 0x9E3779B97F4A7C15ull;
  for (size_t i = 0; i < n; i++) {
    uint64_t x = splitmix_mix(seed + (uint64_t)(i + 1) * gamma);
    while (x >= GL_P) x = splitmix_mix(x ^ gamma);
    v[i] = x;
  }
}

/* ---------------------------------------------------------------- Poseidon2 */

static const uint64_t P2_M4[4][4] = {{5, 7, 1, 3}, {4, 6, 1, 1}, {1, 3, 5, 7}, {1, 1, 4, 6}};
static const unsigned P2_MU[12] = {4, 14, 11, 8, 0, 5, 2, 9, 13, 6, 3, 12};

/* y = M_E x with M_E's 4 x 4 blocks 2 M4 on the diagonal and M4 elsewhere */
static void p2_external(uint64_t x[12]) {
  uint64_t y[12];
  for (int r = 0; r < 12; r++) {
    uint64_t acc = 0;
    for (int c = 0; c < 12; c++) {
      uint64_t m = P2_M4[r % 4][c % 4];
      if (r / 4 == c / 4) m *= 2;
      acc = gl_add(acc, gl_mul(m, x[c]));
    }
    y[r] = acc;
  }
  memcpy(x, y, sizeof(y));
}

/* y = (J + diag(2^mu)) x */
static void p2_internal(uint64_t x[12]) {
  uint64_t sum = 0;
  for (int i = 0; i < 12; i++) sum = gl_add(sum, x[i]);
  for (int i = 0; i < 12; i++) x[i] = gl_add(sum, gl_mul(x[i], 1ull << P2_MU[i]));
}

static uint64_t p2_sbox(uint64_t x) { return gl_pow(x, 7); }

void gl_oracle_poseidon2_matrix(int internal, uint64_t x[12]) {
  for (int i = 0; i < 12; i++) x[i] = gl_reduce(x[i]);
  if (internal)
    p2_internal(x);
  else
    p2_external(x);
}

void gl_oracle_poseidon2_permutation(uint64_t state[12], const uint64_t *rc) {
  for (int i = 0; i < 12; i++) state[i] = gl_reduce(state[i]);
  p2_external(state);
  for (int r = 0; r < 30; r++) {
    const int full = r < 4 || r >= 26;
    if (full) {
      for (int i = 0; i < 12; i++) state[i] = p2_sbox(gl_add(state[i], gl_reduce(rc[12 * r + i])));
      p2_external(state);
    } else {
      state[0] = p2_sbox(gl_add(state[0], gl_reduce(rc[12 * r])));
      p2_internal(state);
    }
  }
}

void gl_oracle_poseidon2_leaves(const uint64_t *values, uint64_t *results, uint32_t rows_count,
                                uint32_t cols_count, uint32_t count, int load_intermediate,
                                int store_intermediate, const uint64_t *rc) {
  const uint64_t total = (uint64_t)rows_count * cols_count;
  for (uint64_t gid = 0; gid < count; gid++) {
    uint64_t state[12] = {0};
    if (load_intermediate)
      for (int k = 0; k < 4; k++) state[8 + k] = results[gid + (uint64_t)k * count];
    for (uint64_t off = 0; off < total; off += 8) {
      for (int i = 0; i < 8; i++) {
        const uint64_t o = off + i, r = o % rows_count, c = o / rows_count;
        state[i] = c < cols_count ? values[gid * rows_count + r + c * rows_count * (uint64_t)count] : 0;
      }
      gl_oracle_poseidon2_permutation(state, rc);
    }
    for (int k = 0; k < 4; k++) results[gid + (uint64_t)k * count] = state[(store_intermediate ? 8 : 0) + k];
  }
}

void gl_oracle_poseidon2_nodes(const uint64_t *values, uint64_t *results, uint32_t count,
                               const uint64_t *rc) {
  for (uint64_t gid = 0; gid < count; gid++) {
    uint64_t state[12] = {0};
    for (int k = 0; k < 4; k++) {
      state[k] = values[2 * (gid + (uint64_t)k * count)];
      state[4 + k] = values[2 * (gid + (uint64_t)k * count) + 1];
    }
    gl_oracle_poseidon2_permutation(state, rc);
    for (int k = 0; k < 4; k++) results[gid + (uint64_t)k * count] = state[k];
  }
}
