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
    one covers log_n with a 9..12-bit contiguous pass under 4..8-bit strided passes."""
    from boojum_amd import tune

    for log_n in range(1, 29):
        lib_c = tune.library_candidates(log_n)
        assert lib_c == tune.candidate_splits(log_n), log_n
        for bits in lib_c:
            assert sum(bits) == log_n
            if log_n > 12:
                assert 9 <= bits[0] <= 12 and all(4 <= b <= 8 for b in bits[1:]) and len(bits) <= 8
            else:
                assert bits == [log_n]
    assert [12, 8] in tune.library_candidates(20)  # the built-in 2^20 plan is among them


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
