"""Independent big-integer restatement of the transform definitions.

Used to pin the C oracle: everything here is computed straight from the definitions with
Python integers (no shared code with oracle/gl_oracle.c), O(n^2), so it is only run for
small n.

Definitions (crates/boojum-cuda/src/ntt.rs:527-583, crates/shivini/src/utils.rs:5-16):
  omega_{2^k}   = (7^((p-1)/2^32))^(2^(32-k))
  coset s       = 1 if log_ext == 0 else 7 * omega_{n*2^e}^c
  forward       X_k = sum_j s^j x_j omega_n^(j k)
  inverse       x_j = n^-1 s^-j sum_k X_k omega_n^(-j k)
  n2b: natural input, bit-reversed output; b2n: bit-reversed input, natural output.
"""
P = 2**64 - 2**32 + 1


def omega(log_size):
    w = pow(7, (P - 1) >> 32, P)
    for _ in range(log_size, 32):
        w = w * w % P
    return w


def bitrev(i, bits):
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def coset_shift(log_n, log_ext, coset_idx):
    if log_ext == 0:
        return 1
    return 7 * pow(omega(log_n + log_ext), coset_idx, P) % P


def dft(x, log_n, s=1, inverse=False):
    n = 1 << log_n
    w = omega(log_n)
    if inverse:
        w = pow(w, P - 2, P)
        s_inv = pow(s, P - 2, P)
        n_inv = pow(n, P - 2, P)
        out = []
        for j in range(n):
            acc = 0
            wj = pow(w, j, P)
            t = 1
            for k in range(n):
                acc = (acc + (x[k] % P) * t) % P
                t = t * wj % P
            out.append(acc * n_inv % P * pow(s_inv, j, P) % P)
        return out
    xs = [(x[j] % P) * pow(s, j, P) % P for j in range(n)]
    out = []
    for k in range(n):
        wk = pow(w, k, P)
        acc = 0
        t = 1
        for j in range(n):
            acc = (acc + xs[j] * t) % P
            t = t * wk % P
        out.append(acc)
    return out


def batch_mode(x, log_n, bitrev_inputs, inverse, log_ext=0, coset_idx=0):
    """One column of batch_ntt_internal, from the definitions."""
    n = 1 << log_n
    if bitrev_inputs:
        x = [x[bitrev(i, log_n)] for i in range(n)]
    y = dft(list(x), log_n, coset_shift(log_n, log_ext, coset_idx), inverse)
    if not bitrev_inputs:
        y = [y[bitrev(i, log_n)] for i in range(n)]
    return y


def horner(coeffs, point):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * point + c) % P
    return acc


def lde_point(log_n, log_blowup, position):
    """Domain point held at `position` of a column's bit-reversed LDE (coset-major)."""
    big = log_n + log_blowup
    return 7 * pow(omega(big), bitrev(position, big), P) % P
