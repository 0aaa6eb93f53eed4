"""Pins the CPU oracle (oracle/gl_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible slip (wrong root, sign, twiddle index, order,
dropped scaling) fails it:
  * closed-form roots of unity and the generator's order
  * field operations against Python big integers on edge and random values
  * every batch mode against the independent O(n^2) big-integer DFT (tests/bigint_ref.py)
  * the committed golden fixtures (tests/golden/ntt_golden.npz)
  * transform invariants at sizes the brute force cannot reach
  * LDE output equals direct polynomial evaluation at the coset points
"""
import os

import numpy as np
import pytest

import bigint_ref as br
import gl_oracle
from gpu_util import raw_u64

P = br.P
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ntt_golden.npz")


def test_roots_closed_form():
    w32 = pow(7, (P - 1) >> 32, P)
    assert w32 == 0x185629DCDA58878C
    assert gl_oracle.omega(32) == w32
    # order exactly 2^32
    assert pow(w32, 1 << 31, P) == P - 1
    assert gl_oracle.omega(30) == 0x7E9BD009B86A0845
    # omega_64 is the shift 2^39, omega_2 = -1
    assert gl_oracle.omega(6) == pow(2, 39, P)
    assert gl_oracle.omega(1) == P - 1
    # 7 generates F_p^*: p - 1 = 2^32 * 3 * 5 * 17 * 257 * 65537
    for q in (2, 3, 5, 17, 257, 65537):
        assert pow(7, (P - 1) // q, P) != 1
    assert gl_oracle.field_op("gl_inv", 7) == 0x249249246DB6DB6E
    assert gl_oracle.field_op("gl_pow", 2, P - 1 - 20) == 0xFFFFEFFF00001001  # 2^-20


@pytest.mark.parametrize("seed", [0, 1])
def test_field_ops(seed):
    rng = np.random.default_rng(seed)
    edge = [0, 1, 2, P - 2, P - 1, 2**32 - 1, 2**32, 2**63, 2**64 - 2**32]
    vals = edge + [int(v) % P for v in rng.integers(0, 2**64, 200, dtype=np.uint64)]
    for a in vals:
        for b in vals[:20]:
            assert gl_oracle.field_op("gl_add", a, b) == (a + b) % P
            assert gl_oracle.field_op("gl_sub", a, b) == (a - b) % P
            assert gl_oracle.field_op("gl_mul", a, b) == (a * b) % P
    for x in [P, P + 1, 2**64 - 1]:
        assert gl_oracle.field_op("gl_reduce", x) == x % P
        assert gl_oracle.field_op("gl_mul", x, 1) == x % P


def test_twiddles_bitreversed_prefix():
    big = gl_oracle.twiddles(10, False)
    for log_n in range(1, 10):
        tw = gl_oracle.twiddles(log_n, False)
        assert np.array_equal(tw, big[: len(tw)])
    w = br.omega(10)
    for k in range(512):
        assert int(big[k]) == pow(w, br.bitrev(k, 9), P)
    inv = gl_oracle.twiddles(10, True)
    for k in range(512):
        assert int(inv[k]) * int(big[k]) % P == 1


@pytest.mark.parametrize("bitrev_inputs,inverse", [(False, False), (True, False), (False, True), (True, True)])
@pytest.mark.parametrize("log_ext,coset_idx", [(0, 0), (3, 3), (1, 0), (2, 1)])
def test_batch_modes_vs_bruteforce(bitrev_inputs, inverse, log_ext, coset_idx):
    for log_n in range(1, 7):
        n = 1 << log_n
        cols, stride = 3, 2 * n
        rng = np.random.default_rng(log_n * 31 + coset_idx)
        x = rng.integers(0, 2**64, cols * stride, dtype=np.uint64)
        y = gl_oracle.batch_ntt(x, log_n, cols, stride, stride, bitrev_inputs, inverse, log_ext, coset_idx)
        for c in range(cols):
            ref = br.batch_mode([int(t) for t in x[c * stride: c * stride + n]], log_n, bitrev_inputs,
                                inverse, log_ext, coset_idx)
            assert [int(t) for t in y[c * stride: c * stride + n]] == ref


def test_golden_fixtures():
    g = np.load(GOLDEN)
    for log_n in range(1, 10):
        x = g[f"in_{log_n}"]
        n = 1 << log_n
        for b in (0, 1):
            for i in (0, 1):
                for log_ext, coset_idx in [(0, 0), (3, 3), (1, 0)]:
                    y = gl_oracle.batch_ntt(x, log_n, 1, n, n, b, i, log_ext, coset_idx)
                    assert np.array_equal(y, g[f"ntt_{log_n}_{b}{i}_{log_ext}_{coset_idx}"])
    for log_n in range(1, 8):
        ev = g[f"lde_in_{log_n}"]
        for lb in (1, 2, 3):
            assert np.array_equal(gl_oracle.lde(ev, log_n, 1, lb), g[f"lde_{log_n}_{lb}"])


@pytest.mark.parametrize("log_n", [10, 13])
def test_invariants(log_n):
    n = 1 << log_n
    rng = np.random.default_rng(log_n)
    x = gl_oracle.fill_uniform(n, 77 + log_n)
    # inverse(forward) == identity, both orderings
    y = gl_oracle.batch_ntt(x, log_n, 1, n, n, False, False)
    z = gl_oracle.batch_ntt(y, log_n, 1, n, n, True, True)
    assert np.array_equal(z, x)
    # coset forward then coset inverse is identity too
    y = gl_oracle.batch_ntt(x, log_n, 1, n, n, False, False, 2, 3)
    z = gl_oracle.batch_ntt(y, log_n, 1, n, n, True, True, 2, 3)
    assert np.array_equal(z, x)
    # delta_0 -> all ones; e_1 -> omega^k in bit-reversed order
    d = np.zeros(n, dtype=np.uint64)
    d[0] = 1
    assert np.all(gl_oracle.batch_ntt(d, log_n, 1, n, n, False, False) == 1)
    d[0], d[1] = 0, 1
    y = gl_oracle.batch_ntt(d, log_n, 1, n, n, False, False)
    w = br.omega(log_n)
    for pos in rng.integers(0, n, 64):
        assert int(y[pos]) == pow(w, br.bitrev(int(pos), log_n), P)
    # linearity: NTT(a*x + y) == a*NTT(x) + NTT(y)
    a = 0x123456789ABCDEF
    x2 = gl_oracle.fill_uniform(n, 5)
    lhs_in = np.array([(a * int(u) + int(v)) % P for u, v in zip(x, x2)], dtype=np.uint64)
    lhs = gl_oracle.batch_ntt(lhs_in, log_n, 1, n, n, False, False, 1, 1)
    fx = gl_oracle.batch_ntt(x, log_n, 1, n, n, False, False, 1, 1)
    fy = gl_oracle.batch_ntt(x2, log_n, 1, n, n, False, False, 1, 1)
    assert all(int(l) == (a * int(u) + int(v)) % P for l, u, v in zip(lhs[:256], fx[:256], fy[:256]))


@pytest.mark.parametrize("log_n,log_blowup", [(8, 1), (9, 2), (6, 3)])
def test_lde_equals_direct_evaluation(log_n, log_blowup):
    n = 1 << log_n
    ev = gl_oracle.fill_uniform(2 * n, 9)
    out = gl_oracle.lde(ev, log_n, 2, log_blowup)
    big = n << log_blowup
    for c in range(2):
        mono_br = gl_oracle.batch_ntt(ev[c * n:(c + 1) * n], log_n, 1, n, n, False, True)
        coeffs = [int(t) for t in gl_oracle.bit_reverse(mono_br, log_n, 1)]
        # monomials via n2b inverse + bit reverse: they must interpolate the evaluations
        w = br.omega(log_n)
        for j in (0, 1, n - 1):
            assert br.horner(coeffs, pow(w, j, P)) == int(ev[c * n + j])
        for q in list(range(8)) + [big - 1, big // 2 + 3]:
            assert int(out[c * big + q]) == br.horner(coeffs, br.lde_point(log_n, log_blowup, q))


def test_rejects_reference_asserts():
    x = np.zeros(16, dtype=np.uint64)
    with pytest.raises(ValueError):
        gl_oracle.batch_ntt(x, 0, 1, 16, 16, False, False)
    with pytest.raises(ValueError):
        gl_oracle.batch_ntt(x, 4, 1, 8, 16, False, False)
    with pytest.raises(ValueError):
        gl_oracle.batch_ntt(x, 4, 1, 16, 16, False, False, 0, 1)
    with pytest.raises(ValueError):
        gl_oracle.batch_ntt(x, 4, 1, 16, 16, False, False, 2, 4)


def test_bit_reverse():
    log_rows, cols = 6, 3
    n = 1 << log_rows
    src = np.arange(cols * n, dtype=np.uint64)
    dst = gl_oracle.bit_reverse(src, log_rows, cols)
    for c in range(cols):
        for i in range(n):
            assert dst[c * n + br.bitrev(i, log_rows)] == src[c * n + i]


def test_fill_uniform_matches_splitmix():
    def mix(z):
        M = 2**64 - 1
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    g = 0x9E3779B97F4A7C15
    v = gl_oracle.fill_uniform(100, 42)
    for i in range(100):
        x = mix((42 + (i + 1) * g) & (2**64 - 1))
        while x >= P:
            x = mix(x ^ g)
        assert int(v[i]) == x


# ---- powers and omega shifts (crates/boojum-cuda/native/ops_complex.cu:8-86 semantics) ----
W30 = 0x7E9BD009B86A0845  # omega_{2^30}, closed form pinned in test_roots_closed_form


def test_powers_of_w_closed_form_and_group():
    for ld in (0, 1, 4, 10, 30):
        count = min(1 << ld, 64) if ld else 3
        r = gl_oracle.powers_of_w(ld, 0, False, False, count)
        for i in range(count):
            assert int(r[i]) == pow(W30, (i << (30 - ld)) % 2**30, P), (ld, i)
    # the full domain of 2^6: distinct, r[i] r[j] = r[i+j], r[32] = -1, omega_64 = 2^39
    r = [int(x) for x in gl_oracle.powers_of_w(6, 0, False, False, 64)]
    assert len(set(r)) == 64 and r[1] == pow(2, 39, P) and r[32] == P - 1
    assert all(r[i] * r[j] % P == r[(i + j) % 64] for i in range(0, 64, 7) for j in range(0, 64, 5))
    inv = gl_oracle.powers_of_w(6, 0, True, False, 64)
    assert all(int(a) * int(b) % P == 1 for a, b in zip(r, inv))
    br_ = gl_oracle.powers_of_w(6, 0, False, True, 64)
    assert [int(br_[i]) for i in range(64)] == [r[br.bitrev(i, 6)] for i in range(64)]


def test_powers_of_w_offset_and_wrap():
    # i + offset wraps at 2^32 and, bit-reversed, only its low log_degree bits count
    off = 2**32 - 5
    r = gl_oracle.powers_of_w(8, off, False, True, 10)
    for i in range(10):
        raw = (i + off) % 2**32
        j = br.bitrev(raw & 255, 8)
        assert int(r[i]) == pow(W30, (j << 22) % 2**30, P)
    r = gl_oracle.powers_of_w(3, 5, True, False, 20)
    for i in range(20):
        e = (-(((i + 5) << 27) % 2**32)) % 2**30
        assert int(r[i]) == pow(W30, e, P)


def test_powers_of_g():
    g_inv = pow(7, P - 2, P)
    for ld, off, inv, bitrev in [(5, 0, False, False), (5, 3, True, True), (30, 7, False, False), (12, 2**31, True, False)]:
        r = gl_oracle.powers_of_g(ld, off, inv, bitrev, 40)
        for i in range(40):
            raw = (i + off) % 2**32
            j = br.bitrev(raw & ((1 << ld) - 1), ld) if bitrev else raw
            e = ((j << (30 - ld)) % 2**32) % 2**30
            assert int(r[i]) == pow(g_inv if inv else 7, e, P), (ld, off, inv, bitrev, i)


def test_powers_by_val_bf():
    base = 0x1234567890ABCDEF
    for off, bitrev in [(0, False), (9, False), (2**32 - 3, False), (5, True)]:
        r = gl_oracle.powers_by_val_bf(base, off, bitrev, 33)
        for i in range(33):
            e = ((br.bitrev(i, 32) if bitrev else i) + off) % 2**32
            assert int(r[i]) == pow(base, e, P), (off, bitrev, i)


def ext_mul_py(x, y):
    return ((x[0] * y[0] + 7 * x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)


def ext_pow_py(x, e):
    r = (1, 0)
    while e:
        if e & 1:
            r = ext_mul_py(r, x)
        x = ext_mul_py(x, x)
        e >>= 1
    return r


def test_powers_by_val_ef():
    # 7 is a quadratic non-residue, so F_p[u]/(u^2 - 7) is a field (goldilocks_extension.cuh:11)
    assert pow(7, (P - 1) // 2, P) == P - 1
    base = (0xDEADBEEF12345678, 0x0123456789ABCDEF)
    for off, bitrev in [(0, False), (11, True), (2**32 - 1, False)]:
        r = gl_oracle.powers_by_val_ef(base[0], base[1], off, bitrev, 17)
        for i in range(17):
            e = ((br.bitrev(i, 32) if bitrev else i) + off) % 2**32
            assert (int(r[i]), int(r[17 + i])) == ext_pow_py(base, e), (off, bitrev, i)
    # u^2 = 7: powers of u alternate between the planes
    r = gl_oracle.powers_by_val_ef(0, 1, 0, False, 5)
    assert [(int(r[i]), int(r[5 + i])) for i in range(5)] == [(1, 0), (0, 1), (7, 0), (0, 7), (49, 0)]


def test_omega_shift():
    rng = np.random.default_rng(4)
    src = rng.integers(0, 2**64, size=50, dtype=np.uint64)
    # shift 1 is multiplication by the powers of omega (a relation between the two functions)
    for ld, off, inv in [(10, 0, False), (10, 77, True), (30, 2**32 - 20, False)]:
        got = gl_oracle.omega_shift(src, ld, off, inv, 1)
        pw = gl_oracle.powers_of_w(ld, off, inv, False, 50)
        assert all(int(g) == int(s) % P * int(w) % P for g, s, w in zip(got, src, pw))
    for shift in (3, 2**31 + 5):
        got = gl_oracle.omega_shift(src, 12, 9, False, shift)
        for i in range(50):
            e = (((shift * (i + 9)) % 2**32) << 18) % 2**32 % 2**30
            assert int(got[i]) == int(src[i]) % P * pow(W30, e, P) % P


# ---- barycentric evaluation (crates/boojum-cuda/native/barycentric.cu:11-158) ----
def _ext_horner(coeffs, x):
    acc = (0, 0)
    for c in reversed(coeffs):
        acc = ext_mul_py(acc, x)
        acc = ((acc[0] + c[0]) % P, (acc[1] + c[1]) % P)
    return acc


def test_barycentric_equals_direct_evaluation():
    """sum_i L_i y_i equals the polynomial's value at X, computed by Horner in big integers,
    for base and extension points and both coefficient orders."""
    rng = np.random.default_rng(21)

    def rnd():
        return int(rng.integers(0, P, dtype=np.uint64))

    for log_n in (0, 1, 4, 6):
        n = 1 << log_n
        w = pow(W30, 1 << (30 - log_n), P)
        coset = 7 * pow(W30, 12345, P) % P
        mono = [(rnd(), rnd()) for _ in range(n)]  # extension coefficients
        mono_b = [(m[0], 0) for m in mono]
        pts = [coset * pow(w, i, P) % P for i in range(n)]
        ys_e = [_ext_horner(mono, (pt, 0)) for pt in pts]
        ys_b = [_ext_horner(mono_b, (pt, 0))[0] for pt in pts]
        xb, xe = rnd(), (rnd(), rnd())
        for bitrev in (False, True):
            # storage position q holds the value at domain point bitrev(q) (or q)
            src = [br.bitrev(q, log_n) if bitrev else q for q in range(n)]
            yb = np.array([ys_b[i] for i in src], dtype=np.uint64)
            ye = np.array([ys_e[i][0] for i in src] + [ys_e[i][1] for i in src], dtype=np.uint64)
            cb = gl_oracle.barycentric_coeffs(xb, coset, n, bitrev)
            assert int(gl_oracle.barycentric_eval(0, yb, n, 1, cb)[0]) == _ext_horner(mono_b, (xb, 0))[0]
            ce = gl_oracle.barycentric_coeffs(xe, coset, n, bitrev, ext=True)
            got = gl_oracle.barycentric_eval(1, yb, n, 1, ce)
            assert (int(got[0]), int(got[1])) == _ext_horner(mono_b, xe), (log_n, bitrev)
            got = gl_oracle.barycentric_eval(2, ye, n, 1, ce)
            assert (int(got[0]), int(got[1])) == _ext_horner(mono, xe), (log_n, bitrev)


def test_barycentric_on_lde_cosets():
    """The interpolant of evaluations on the base domain, evaluated at X from the evaluations
    themselves and from each bit-reversed LDE coset (shift 7 w_{Bn}^bitrev(c)), agrees."""
    rng = np.random.default_rng(22)
    log_n, lb = 10, 2
    n, B = 1 << log_n, 1 << lb
    ev = rng.integers(0, P, size=n, dtype=np.uint64)
    lde = gl_oracle.lde(ev, log_n, 1, lb)  # coset c at [c n, (c + 1) n), bit-reversed
    x = int(rng.integers(0, P, dtype=np.uint64))
    want = int(gl_oracle.barycentric_eval(0, ev, n, 1, gl_oracle.barycentric_coeffs(x, 1, n, False))[0])
    for c in range(B):
        shift = 7 * pow(W30, br.bitrev(c, lb) << (30 - log_n - lb), P) % P
        coeffs = gl_oracle.barycentric_coeffs(x, shift, n, True)
        assert int(gl_oracle.barycentric_eval(0, lde[c * n:(c + 1) * n], n, 1, coeffs)[0]) == want, c


# ---- Poseidon2 (width 12) Merkle hashing, crates/boojum-cuda/native/poseidon2/gl/ ----
P2_MU = [4, 14, 11, 8, 0, 5, 2, 9, 13, 6, 3, 12]


def _m4_chain(x):
    """M4 as the reference evaluates it, an addition chain (poseidon2_gl.cuh:9-25), here on
    Python integers: the matrix it computes must be the paper's M4 (eprint 2023/323, App. B)."""
    t0 = x[0] + x[1]
    t1 = x[2] + x[3]
    t2 = 2 * x[1] + t1
    t3 = 2 * x[3] + t0
    t4 = 4 * t1 + t3
    t5 = 4 * t0 + t2
    return [t3 + t5, t5, t2 + t4, t4]


def test_poseidon2_linear_layers():
    m4 = [[5, 7, 1, 3], [4, 6, 1, 1], [1, 3, 5, 7], [1, 1, 4, 6]]
    for j in range(4):  # the chain's columns are the paper's matrix
        e = [int(i == j) for i in range(4)]
        assert _m4_chain(e) == [m4[r][j] for r in range(4)]
    # M_E = (I_3 + J_3) (x) M4, M_I = J_12 + diag(2^mu), as integer matrices mod p
    me = [[(1 + (r // 4 == c // 4)) * m4[r % 4][c % 4] for c in range(12)] for r in range(12)]
    mi = [[1 + (r == c) * (1 << P2_MU[r]) for c in range(12)] for r in range(12)]
    rng = np.random.default_rng(21)
    for trial in range(40):
        x = [int(v) for v in rng.integers(0, P, 12, dtype=np.uint64)]
        if trial == 0:
            x = [P - 1] * 12
        for internal, m in ((0, me), (1, mi)):
            want = [sum(m[r][c] * x[c] for c in range(12)) % P for r in range(12)]
            assert [int(v) for v in gl_oracle.poseidon2_matrix(internal, x)] == want, (internal, trial)
    # the reference's M_E formulation: M4 per tile, then each lane plus the sum over tiles
    x = [int(v) for v in rng.integers(0, P, 12, dtype=np.uint64)]
    y = sum((_m4_chain(x[4 * t:4 * t + 4]) for t in range(3)), [])
    acc = [sum(y[4 * t + j] for t in range(3)) for j in range(4)]
    assert [int(v) for v in gl_oracle.poseidon2_matrix(0, x)] == [(acc[i % 4] + y[i]) % P for i in range(12)]


def test_poseidon2_permutation_rounds():
    """x^7 is a permutation of F_p (gcd(7, p - 1) = 1); the permutation equals the paper's round
    sequence evaluated on Python integers with the matrices above."""
    from math import gcd

    assert gcd(7, P - 1) == 1
    m4 = [[5, 7, 1, 3], [4, 6, 1, 1], [1, 3, 5, 7], [1, 1, 4, 6]]
    me = [[(1 + (r // 4 == c // 4)) * m4[r % 4][c % 4] for c in range(12)] for r in range(12)]
    mi = [[1 + (r == c) * (1 << P2_MU[r]) for c in range(12)] for r in range(12)]

    def mat(m, x):
        return [sum(m[r][c] * x[c] for c in range(12)) % P for r in range(12)]

    rng = np.random.default_rng(22)
    for trial in range(4):
        rc = rng.integers(0, P, 360, dtype=np.uint64)
        x = [int(v) for v in rng.integers(0, P, 12, dtype=np.uint64)]
        got = gl_oracle.poseidon2_permutation(x, rc)
        s = mat(me, x)
        for r in range(30):
            if r < 4 or r >= 26:
                s = mat(me, [pow((s[i] + int(rc[12 * r + i])) % P, 7, P) for i in range(12)])
            else:
                s = mat(mi, [pow((s[0] + int(rc[12 * r])) % P, 7, P)] + s[1:])
        assert [int(v) for v in got] == s, trial


def test_poseidon2_sponge_layouts():
    """Leaves, nodes and the tree against Python loops over the oracle's permutation: leaf gid
    reads values[gid * rows + r + c * rows * count] column by column (the reference's CPU check,
    crates/boojum-cuda/src/poseidon2.rs:575-590), 8 per permutation in overwrite mode, zero-filled;
    a node hashes left || right; intermediate capacities chain partial column sets."""
    rng = np.random.default_rng(23)
    rc = rng.integers(0, P, 360, dtype=np.uint64)
    for rows, cols, count in ((1, 5, 6), (2, 7, 4), (4, 4, 3), (8, 3, 2)):
        values = raw_u64(rng, rows * cols * count)
        got = gl_oracle.poseidon2_leaves(values, rows, cols, count, rc)
        for gid in range(count):
            seq = [int(values[gid * rows + o % rows + (o // rows) * rows * count]) % P for o in range(rows * cols)]
            st = np.zeros(12, dtype=np.uint64)
            for off in range(0, len(seq), 8):
                chunk = seq[off:off + 8] + [0] * max(0, off + 8 - len(seq))
                st[:8] = chunk
                st = gl_oracle.poseidon2_permutation(st, rc)
            assert [int(got[gid + k * count]) for k in range(4)] == [int(v) for v in st[:4]], (rows, cols, gid)
    # two calls over column halves (intermediate capacity) equal one call over all columns, when
    # the first half fills whole permutations
    rows, count = 2, 5
    values = raw_u64(rng, rows * 8 * count)
    whole = gl_oracle.poseidon2_leaves(values, rows, 8, count, rc)
    half = rows * 4 * count
    mid = gl_oracle.poseidon2_leaves(values[:half], rows, 4, count, rc, store_intermediate=True)
    done = gl_oracle.poseidon2_leaves(values[half:], rows, 4, count, rc, load_intermediate=True, results=mid)
    assert np.array_equal(done, whole)
    # nodes: word k of the left child at values[2 (gid + k count)], the right child next to it
    count = 4
    children = raw_u64(rng, 8 * count) % np.uint64(P)
    got = gl_oracle.poseidon2_nodes(children, count, rc)
    for gid in range(count):
        st = np.zeros(12, dtype=np.uint64)
        for k in range(4):
            st[k] = children[2 * (gid + k * count)]
            st[4 + k] = children[2 * (gid + k * count) + 1]
        assert np.array_equal(got[gid::count][:4], gl_oracle.poseidon2_permutation(st, rc)[:4])
    # tree: leaf digests, then each node layer hashes adjacent pairs of the layer below
    leaves, layers = 8, 4
    values = raw_u64(rng, leaves * 2 * 3)
    tree = gl_oracle.poseidon2_merkle_tree(values, 1, leaves, layers, rc)
    layer = [[int(tree[g + k * leaves]) for k in range(4)] for g in range(leaves)]
    at, n = 4 * leaves, leaves
    for _ in range(layers - 1):
        n //= 2
        nxt = []
        for g in range(n):
            st = np.array(layer[2 * g] + layer[2 * g + 1] + [0] * 4, dtype=np.uint64)
            nxt.append([int(v) for v in gl_oracle.poseidon2_permutation(st, rc)[:4]])
        assert [[int(tree[at + g + k * n]) for k in range(4)] for g in range(n)] == nxt
        at += 4 * n
        layer = nxt
    assert not tree[at:].any()
