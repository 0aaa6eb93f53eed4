"""boojum_cuda::poseidon2::GpuTreeHasher mirror for the Goldilocks Poseidon2 hasher
(crates/boojum-cuda/src/poseidon2.rs:165-316) over the C-ABI.  Slices are DeviceSlices of u64
words; digests are 4 planes (word k of digest g at results[g + k * count]).  The current
device's context needs its round constants first (set_round_constants)."""
import numpy as np

from ._lib import check, lib

ROUNDS, WIDTH, RATE, CAPACITY = 30, 12, 8, 4


def _h(stream):
    return stream.handle if stream is not None else None


def set_round_constants(ctx, round_constants):
    """round_constants: 30 x 12 field elements, row-major by round (boojum's
    ALL_ROUND_CONSTANTS_AS_FIELD_ELEMENTS)."""
    rc = np.ascontiguousarray(round_constants, dtype=np.uint64).reshape(-1)
    assert rc.size == ROUNDS * WIDTH
    check(lib().ba_poseidon2_gl_set_round_constants(ctx.handle, rc.ctypes.data), "poseidon2 set_round_constants")


def compute_leaf_hashes(values, results, log_rows_per_hash, load_intermediate=False, store_intermediate=False,
                        stream=None):
    check(lib().ba_poseidon2_gl_compute_leaf_hashes(values.ptr, values.len, results.ptr, results.len,
                                                    log_rows_per_hash, int(load_intermediate),
                                                    int(store_intermediate), _h(stream)), "compute_leaf_hashes")


def compute_node_hashes(values, results, stream=None):
    check(lib().ba_poseidon2_gl_compute_node_hashes(values.ptr, values.len, results.ptr, results.len, _h(stream)),
          "compute_node_hashes")


def build_merkle_tree_leaves(values, results, log_rows_per_hash, load_intermediate=False, store_intermediate=False,
                             stream=None):
    check(lib().ba_poseidon2_gl_build_merkle_tree_leaves(values.ptr, values.len, results.ptr, results.len,
                                                         log_rows_per_hash, int(load_intermediate),
                                                         int(store_intermediate), _h(stream)),
          "build_merkle_tree_leaves")


def build_merkle_tree_nodes(values, results, layers_count, stream=None):
    check(lib().ba_poseidon2_gl_build_merkle_tree_nodes(values.ptr, values.len, results.ptr, results.len,
                                                        layers_count, _h(stream)), "build_merkle_tree_nodes")


def build_merkle_tree(values, results, log_rows_per_hash, stream, layers_count):
    """Same argument order as GpuTreeHasher::build_merkle_tree."""
    check(lib().ba_poseidon2_gl_build_merkle_tree(values.ptr, values.len, results.ptr, results.len,
                                                  log_rows_per_hash, layers_count, _h(stream)), "build_merkle_tree")
