"""CPU-side checks of the C-ABI library: it loads without a GPU and exports every function
include/boojum_amd.h declares, and the Python binding declares each of them."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "boojum_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ba_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("ba_batch_ntt_internal", "ba_batch_ntt_in_place", "ba_batch_ntt_out_of_place",
                     "ba_context_create", "ba_context_destroy", "ba_bit_reverse", "ba_bit_reverse_in_place"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from boojum_amd import _lib

    if not os.path.exists(_lib.library_path):
        pytest.fail("libboojum_amd.so not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.library_path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from boojum_amd import _lib

    assert set(declared_functions()) == set(_lib.SIGNATURES), set(declared_functions()) ^ set(_lib.SIGNATURES)


def test_no_gpu_calls_fail_cleanly_without_device():
    """Without a GPU the runtime reports zero devices instead of crashing the process."""
    import boojum_amd

    n = boojum_amd.device_count()
    assert n >= 0


def test_job_struct_matches_abi():
    """ba_job is 80 bytes (8 u32 + 3 pointers + 3 u64), as batch.hip asserts."""
    from boojum_amd._lib import Job

    assert ctypes.sizeof(Job) == 80
    assert [f[0] for f in Job._fields_][:3] == ["kind", "log_n", "num_cols"]
