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


def test_struct_layouts_match_header(tmp_path):
    """sizeof/offsetof of every struct in include/boojum_amd.h, from a C compiler, equal the
    ctypes mirrors the Python binding passes across the ABI."""
    import shutil
    import subprocess
    from boojum_amd import _lib

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"ba_job": _lib.Job, "ba_tuned_plan": _lib.TunedPlan, "ba_serve_request": _lib.ServeRequest,
               "ba_serve_result": _lib.ServeResult, "ba_serve_stats": _lib.ServeStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'  printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("  return 0;\n}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-o", str(exe), str(src)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                         text=True).stdout.split("\n") if l)
    for cname, py in structs.items():
        assert int(got[cname]) == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)


def test_tuning_candidates_host_enumeration():
    """The library's candidate splits (host call, no device) equal the Python mirror's, and every
    one covers log_n with a 9..12-bit contiguous pass under 4..10-bit strided passes."""
    from boojum_amd import tune

    for log_n in range(1, 29):
        lib_c = tune.library_candidates(log_n)
        assert lib_c == tune.candidate_splits(log_n), log_n
        for bits in lib_c:
            assert sum(bits) == log_n
            if log_n > 12:
                assert 9 <= bits[0] <= 12 and all(4 <= b <= 10 for b in bits[1:]) and len(bits) <= 8
            else:
                assert bits == [log_n]
    assert [12, 8] in tune.library_candidates(20)  # the built-in 2^20 plan is among them
    assert [12, 10] in tune.library_candidates(22)  # two passes with a 10-bit strided pass


def test_rust_sys_crate_declares_header():
    """The Rust -sys crate (not compiled here: no cargo) declares exactly the header's functions,
    and its #[repr(C)] structs list the header's fields in order."""
    src = open(os.path.join(ROOT, "zksync-crypto-gpu_amd", "rust", "boojum-amd-sys", "src", "lib.rs")).read()
    rust_fns = set(re.findall(r"pub fn (ba_[a-z0-9_]+)\(", src))
    assert rust_fns == set(declared_functions()), rust_fns ^ set(declared_functions())
    header = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    for name in ("ba_job", "ba_tuned_plan", "ba_serve_request", "ba_serve_result", "ba_serve_stats"):
        body = re.search(r"typedef struct \{([^{}]*)\}\s*%s;" % name, header).group(1)
        c_fields = re.findall(r"(\w+)(?:\[\d+\])?\s*[,;]", body)
        rust_body = re.search(r"pub struct %s \{(.*?)\}" % name, src, flags=re.S).group(1)
        assert re.findall(r"pub (\w+):", rust_body) == c_fields, name


def _lde_rc(lib, evals, in_stride, mono, out, col_stride, coset_stride, log_n, cols, log_b):
    return lib.ba_lde_batch(None, evals, in_stride, mono, out, col_stride, coset_stride, log_n, cols, log_b, None)


def test_lde_output_layout_check_is_exact():
    """ba_lde_batch checks its arguments on the host before it needs a context: with no context
    a valid layout reports BA_ERROR_NOT_INITIALIZED (3) and an overlapping one
    BA_ERROR_INVALID_VALUE (1).  The output vectors k*col_stride + c*coset_stride must be
    pairwise disjoint; the library's answer equals a brute-force check over every pair of
    vectors, over seeded small layouts (interleaved, coset-major, column-major, overlapping)."""
    import random

    from boojum_amd import _lib

    lib = _lib.lib()
    rng = random.Random(5)
    evals, out = 1 << 40, 1 << 44  # addresses only: nothing is dereferenced
    seen = {1: 0, 3: 0}
    for _ in range(3000):
        log_n = rng.randint(1, 4)
        n = 1 << log_n
        cols = rng.randint(1, 6)
        log_b = rng.randint(1, 3)
        b = 1 << log_b
        col_stride = rng.choice([0, rng.randint(0, 4 * n * b), n * b, n, 2 * n, 3 * n])
        coset_stride = rng.choice([0, rng.randint(0, 4 * n * b), n, n * cols, 2 * n, 3 * n])
        starts = sorted(k * col_stride + c * coset_stride for k in range(cols) for c in range(b))
        disjoint = all(y - x >= n for x, y in zip(starts, starts[1:]))
        rc = _lde_rc(lib, evals, n, None, out, col_stride, coset_stride, log_n, cols, log_b)
        assert rc == (3 if disjoint else 1), (log_n, cols, log_b, col_stride, coset_stride)
        seen[rc] += 1
    assert seen[1] > 100 and seen[3] > 100


def test_lde_huge_layout_check_allocates_nothing():
    """A legal call with 2^20 columns at blowup 2^20 (2^40 output vectors) is checked in time
    linear in the smaller count, with no host list of starts: the old sorted list would have
    needed 8 TB.  Outputs overlapping the evaluations are rejected."""
    import time

    from boojum_amd import _lib

    lib = _lib.lib()
    log_n, log_b, cols = 1, 20, 1 << 20
    n = 1 << log_n
    t0 = time.time()
    rc = _lde_rc(lib, 1 << 40, n, None, 1 << 44, n << log_b, n, log_n, cols, log_b)
    assert rc == 3 and time.time() - t0 < 5
    # coset stride n - 1: neighbouring cosets overlap
    assert _lde_rc(lib, 1 << 40, n, None, 1 << 44, n << log_b, n - 1, log_n, cols, log_b) == 1
    # the evaluations inside the output range
    assert _lde_rc(lib, (1 << 44) + 8, n, None, 1 << 44, n << log_b, n, log_n, cols, log_b) == 1
    # an output range past the end of the address space
    assert _lde_rc(lib, 1 << 40, n, None, (1 << 64) - (1 << 20), n << log_b, n, log_n, cols, log_b) == 1


def _rust_fn_params(src, name):
    """(generic parameters, [(param, type)], return type) of `pub fn name` in Rust source text."""
    m = re.search(r"pub fn %s(<[^>]*>)?\((.*?)\)\s*->\s*([^{]+)\{" % name, src, flags=re.S)
    assert m, name
    params = []
    depth, cur = 0, ""
    for ch in m.group(2):
        depth += ch in "<([" and 1 or (ch in ">)]" and -1 or 0)
        if ch == "," and depth == 0:
            params.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        params.append(cur)
    out = [tuple(x.strip() for x in p.split(":", 1)) for p in params if p.strip()]
    return (m.group(1) or ""), out, " ".join(m.group(3).split())


def test_rust_ntt_signatures_keep_boojum_cuda_call_sites():
    """boojum-amd's ntt functions keep boojum_cuda's parameter names and order
    (crates/boojum-cuda/src/ntt.rs:179-191, 294-308, 338-349), take era_cudart's
    `&CudaStream` from the boojum-amd-cudart crate, return `CudaResult<()>`, and are generic over
    the element type (`F: GoldilocksRepr`, implemented for boojum's GoldilocksField under the
    `boojum` feature), so shivini's calls with `DeviceSlice<GoldilocksField>` compile unchanged
    once the dependencies are renamed.  Not compiled here: there is no cargo."""
    rust = os.path.join(ROOT, "zksync-crypto-gpu_amd", "rust")
    ntt = open(os.path.join(rust, "boojum-amd", "src", "ntt.rs")).read()
    lib = open(os.path.join(rust, "boojum-amd", "src", "lib.rs")).read()
    want = {
        "batch_ntt_internal": ["inputs_ptr_in", "outputs_ptr", "log_n", "num_ntts", "stride_between_input_arrays_in",
                               "stride_between_output_arrays", "bitrev_inputs", "inverse", "log_extension_degree",
                               "coset_index", "stream"],
        "batch_ntt_out_of_place": ["inputs_matrix", "outputs_matrix", "log_n", "num_ntts", "inputs_offset",
                                   "outputs_offset", "stride_between_input_arrays", "stride_between_output_arrays",
                                   "bitrev_inputs", "inverse", "log_extension_degree", "coset_index", "stream"],
        "batch_ntt_in_place": ["inputs_matrix", "log_n", "num_ntts", "inputs_offset", "stride_between_input_arrays",
                               "bitrev_inputs", "inverse", "log_extension_degree", "coset_index", "stream"],
    }
    for name, params in want.items():
        gen, got, ret = _rust_fn_params(ntt, name)
        assert [p for p, _ in got] == params, name
        assert gen == "<F: GoldilocksRepr>", name
        assert ret == "CudaResult<()>", name
        types = dict(got)
        assert types["stream"] == "&CudaStream"
        for p in ("log_n", "num_ntts", "log_extension_degree", "coset_index"):
            assert types[p] == "u32", (name, p)
        for p in ("bitrev_inputs", "inverse"):
            assert types[p] == "bool", (name, p)
    assert dict(_rust_fn_params(ntt, "batch_ntt_out_of_place")[1])["inputs_matrix"] == "&DeviceSlice<F>"
    assert dict(_rust_fn_params(ntt, "batch_ntt_in_place")[1])["inputs_matrix"] == "&mut DeviceSlice<F>"
    assert "use crate::cudart::stream::CudaStream;" in ntt and "use crate::cudart::slice::DeviceSlice;" in ntt
    assert "unsafe impl GoldilocksRepr for boojum::field::goldilocks::GoldilocksField {}" in lib
    # the era_cudart paths shivini's LDE code imports (crates/shivini/src/primitives/ntt.rs:1-30)
    cud = os.path.join(rust, "boojum-amd-cudart", "src")
    for mod, items in {"stream": ["pub struct CudaStream", "pub struct CudaStreamWaitEventFlags",
                                  "pub const DEFAULT: CudaStream", "#[repr(transparent)]"],
                       "slice": ["pub struct DeviceSlice<T>([T]);", "#[repr(transparent)]",
                                 "pub unsafe fn from_mut_slice"],
                       "result": ["pub type CudaResult<T> = Result<T, CudaError>;"],
                       "event": ["pub struct CudaEvent"],
                       "memory": ["pub fn memory_copy_async", "pub struct DeviceAllocation<T>"]}.items():
        text = open(os.path.join(cud, mod + ".rs")).read()
        for item in items:
            assert item in text, (mod, item)
