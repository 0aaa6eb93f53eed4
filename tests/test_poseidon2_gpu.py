"""GPU parity of the Poseidon2 Merkle hashing (ba_poseidon2_gl_*, mirroring
crates/boojum-cuda/src/poseidon2.rs:165-316) against the oracle, bit-exact: leaves in every
layout (rows per hash, columns, zero-filled last chunk, intermediate capacities), nodes, whole
trees over the cosets of a GPU LDE, and the reference's preconditions.  boojum's round constants
are not in the reference tree, so the constants here are seeded field elements; the oracle's
permutation is pinned separately (tests/test_oracle.py) and the same constants go to both."""
import numpy as np
import pytest

import gl_oracle
from gpu_util import log, raw_u64

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ba():
    import boojum_amd
    assert boojum_amd.device_count() >= 1, "no GPU visible"
    ctx = boojum_amd.Context.create(12, 12)
    boojum_amd._test_ctx = ctx
    yield boojum_amd
    ctx.destroy()


@pytest.fixture(scope="module")
def rc(ba):
    constants = gl_oracle.fill_uniform(360, 2023)
    ba.poseidon2.set_round_constants(ba._test_ctx, constants)
    return constants


def test_requires_round_constants(ba):
    """A context without installed constants refuses every hashing call."""
    ctx = ba.Context.create(12, 12)  # now the current context of the device
    v, r = ba.DeviceAllocation(64), ba.DeviceAllocation(8)
    try:
        with pytest.raises(ba.BoojumAmdError) as e:
            ba.poseidon2.compute_leaf_hashes(v, r, 0)
        assert e.value.code == 3  # BA_ERROR_NOT_INITIALIZED
        with pytest.raises(ba.BoojumAmdError) as e:
            ba.poseidon2.build_merkle_tree(v, ba.DeviceAllocation(16), 0, None, 1)
        assert e.value.code == 3
    finally:
        v.free()
        r.free()
        ctx.destroy()


@pytest.mark.parametrize("log_rows,cols,count", [(0, 8, 1000), (0, 5, 777), (1, 3, 1024), (2, 7, 513),
                                                 (3, 1, 300), (0, 1, 4096), (1, 40, 2048), (0, 0, 64)])
def test_leaves(ba, rc, log_rows, cols, count):
    rows = 1 << log_rows
    rng = np.random.default_rng(7 * count + cols)
    values = raw_u64(rng, rows * cols * count)
    want = gl_oracle.poseidon2_leaves(values, rows, cols, count, rc)
    d_v = ba.DeviceAllocation.from_host(values) if values.size else ba.DeviceAllocation(0)
    d_r = ba.DeviceAllocation(4 * count)
    try:
        log(f"poseidon2 leaves rows={rows} cols={cols} count={count}")
        ba.poseidon2.compute_leaf_hashes(d_v, d_r, log_rows)
        ba.synchronize()
        assert np.array_equal(d_r.to_host(), want)
    finally:
        d_v.free()
        d_r.free()


def test_leaves_intermediate(ba, rc):
    """Hashing a column subset with store_intermediate, then the rest with load_intermediate,
    equals one call over all columns (the first subset fills whole permutations)."""
    rows, count = 2, 3000
    rng = np.random.default_rng(5)
    values = raw_u64(rng, rows * 12 * count)
    want = gl_oracle.poseidon2_leaves(values, rows, 12, count, rc)
    split = rows * 4 * count  # 4 columns: 8 values per leaf
    d_a = ba.DeviceAllocation.from_host(values[:split])
    d_b = ba.DeviceAllocation.from_host(values[split:])
    d_r = ba.DeviceAllocation(4 * count)
    try:
        ba.poseidon2.compute_leaf_hashes(d_a, d_r, 1, store_intermediate=True)
        mid = gl_oracle.poseidon2_leaves(values[:split], rows, 4, count, rc, store_intermediate=True)
        ba.synchronize()
        assert np.array_equal(d_r.to_host(), mid)
        ba.poseidon2.compute_leaf_hashes(d_b, d_r, 1, load_intermediate=True)
        ba.synchronize()
        assert np.array_equal(d_r.to_host(), want)
    finally:
        for d in (d_a, d_b, d_r):
            d.free()


@pytest.mark.parametrize("count", [1, 3, 256, 5000])
def test_nodes(ba, rc, count):
    rng = np.random.default_rng(count)
    values = raw_u64(rng, 8 * count)
    want = gl_oracle.poseidon2_nodes(values, count, rc)
    d_v, d_r = ba.DeviceAllocation.from_host(values), ba.DeviceAllocation(4 * count)
    try:
        ba.poseidon2.compute_node_hashes(d_v, d_r)
        ba.synchronize()
        assert np.array_equal(d_r.to_host(), want)
    finally:
        d_v.free()
        d_r.free()


@pytest.mark.parametrize("log_n,cols,lb,log_rows,layers", [(10, 4, 1, 0, 11), (12, 3, 2, 1, 8), (14, 2, 1, 2, 5)])
def test_merkle_tree_of_lde_cosets(ba, rc, log_n, cols, lb, log_rows, layers):
    """The consumer of the LDE: a GPU LDE written coset-major (each coset a column-major matrix,
    shivini's layout), then one tree per coset; against the oracle LDE and oracle tree."""
    n, B = 1 << log_n, 1 << lb
    rng = np.random.default_rng(log_n)
    x = raw_u64(rng, cols * n)
    lde = gl_oracle.lde(x, log_n, cols, lb, nthreads=8)  # column-major, cosets of a column adjacent
    d_x = ba.DeviceAllocation.from_host(x)
    d_out = ba.DeviceAllocation(B * cols * n)
    leaves = n >> log_rows
    d_tree = ba.DeviceAllocation(8 * leaves)
    try:
        ba.lde.lde_batch(ba._test_ctx, d_x, log_n, cols, lb, d_out, out_col_stride=n, out_coset_stride=cols * n)
        for c in range(B):
            log(f"poseidon2 tree over LDE coset {c}: 2^{log_n} x {cols}, rows/hash {1 << log_rows}, {layers} layers")
            coset = np.concatenate([lde[k * B * n + c * n:k * B * n + (c + 1) * n] for k in range(cols)])
            want = gl_oracle.poseidon2_merkle_tree(coset, log_rows, leaves, layers, rc)
            d_tree.copy_from_host(np.zeros(8 * leaves, dtype=np.uint64))
            ba.poseidon2.build_merkle_tree(d_out.sub(c * cols * n, cols * n), d_tree, log_rows, None, layers)
            ba.synchronize()
            assert np.array_equal(d_tree.to_host(), want), c
    finally:
        for d in (d_x, d_out, d_tree):
            d.free()


def test_merkle_tree_nodes_layers(ba, rc):
    """build_merkle_tree_nodes over precomputed leaf digests, layer by layer."""
    leaves = 64
    rng = np.random.default_rng(9)
    digests = raw_u64(rng, 4 * leaves) % np.uint64(0xFFFFFFFF00000001)
    d_v, d_r = ba.DeviceAllocation.from_host(digests), ba.DeviceAllocation(4 * leaves)
    try:
        ba.poseidon2.build_merkle_tree_nodes(d_v, d_r, 6)
        ba.synchronize()
        got = d_r.to_host()
        layer, at, count = digests, 0, leaves
        for _ in range(6):
            count //= 2
            layer = gl_oracle.poseidon2_nodes(layer, count, rc)
            assert np.array_equal(got[at:at + 4 * count], layer)
            at += 4 * count
    finally:
        d_v.free()
        d_r.free()


def test_bad_arguments(ba, rc):
    v, r = ba.DeviceAllocation(1024), ba.DeviceAllocation(64)
    try:
        bad_calls = [
            lambda: ba.poseidon2.compute_leaf_hashes(v.sub(0, 1024), r.sub(0, 6), 0),        # results % 4
            lambda: ba.poseidon2.compute_leaf_hashes(v.sub(0, 1000), r.sub(0, 64), 0),       # 16 leaves do not divide
            lambda: ba.poseidon2.compute_leaf_hashes(v.sub(0, 48), r.sub(0, 64), 0, store_intermediate=True),
            lambda: ba.poseidon2.compute_node_hashes(v.sub(0, 16), r.sub(0, 4)),             # values != 2 results
            lambda: ba.poseidon2.compute_node_hashes(v.sub(0, 12), r.sub(0, 6)),             # not whole pairs
            lambda: ba.poseidon2.build_merkle_tree_nodes(v.sub(0, 24), r.sub(0, 24), 1),     # not RATE << layer
            lambda: ba.poseidon2.build_merkle_tree_nodes(v.sub(0, 16), r.sub(0, 16), 3),     # runs out of layers
            lambda: ba.poseidon2.build_merkle_tree(v.sub(0, 64), r.sub(0, 64), 0, None, 0),  # zero layers
            lambda: ba.poseidon2.build_merkle_tree(v.sub(0, 64), r.sub(0, 64), 0, None, 5),  # 2^4 > 8 leaves
            lambda: ba.poseidon2.build_merkle_tree(v.sub(0, 60), r.sub(0, 64), 0, None, 2),  # 8 leaves do not divide
        ]
        for i, call in enumerate(bad_calls):
            with pytest.raises(ba.BoojumAmdError) as e:
                call()
            assert e.value.code == 1, i
        ba.poseidon2.build_merkle_tree_nodes(v.sub(0, 16), r.sub(0, 16), 0)  # zero layers: nothing to do
    finally:
        v.free()
        r.free()
