"""The C-ABI boundary on the CPU: the library loads, exports every function
include/*.h declares, hashes like util::hash (test/SHA1Tests.cpp:58-77),
reports errors with status + message, and the product never touches the
oracle."""
import ctypes as C
import os
import re
import subprocess

import pytest

import skelcl_amd as skl
from skelcl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("skelcl_rt.h", "skelcl_skeletons.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(skl_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (skl_[a-z0-9_]+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    assert len(declared_functions()) > 50


def test_ctypes_binds_every_declared_symbol():
    L = _lib.lib()
    for n in declared_functions():
        assert hasattr(L, n), n
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


@pytest.mark.parametrize("text,digest", [
    ("", "da39a3ee5e6b4b0d3255bfef95601890afd80709"),
    ("The quick brown fox jumps over the lazy dog", "2fd4e1c67a2d28fced849ee1bb76e7391b93eb12"),
    ("The quick brown fox jumps over the lazy cog", "de9f2c7fd25e1b3afad3e85a0bd17d9b100db4b3"),
])
def test_sha1_kats(text, digest):  # SHA1Tests.cpp:58-77
    assert skl.sha1(text.encode()) == digest


def test_sha1_matches_hashlib_on_block_boundaries():
    import hashlib
    for n in (55, 56, 57, 63, 64, 65, 119, 120, 1000):
        data = bytes(range(256)) * 4
        assert skl.sha1(data[:n]) == hashlib.sha1(data[:n]).hexdigest()


def test_calls_before_init_fail_with_status():
    if _lib.lib().skl_is_initialized():
        pytest.skip("library already initialised in this process")
    p = C.c_void_p()
    st = _lib.lib().skl_malloc(C.byref(p), 16)
    assert st == _lib.SKL_ERR_NOT_INIT
    assert b"skl_init" in _lib.lib().skl_last_error()


def test_init_without_gpu_reports_no_device():
    import torch  # noqa: F401  (device count only; does not initialise HIP here)
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    st = _lib.lib().skl_init(1, 0)
    assert st == _lib.SKL_ERR_NO_DEVICE
    assert _lib.lib().skl_last_error()


def test_program_compile_needs_no_gpu():
    code, size, log = C.c_void_p(), C.c_size_t(), C.c_void_p()
    src = b'extern "C" __global__ void k(float* o){ o[threadIdx.x] = 1.0f; }'
    _lib.check(_lib.lib().skl_program_compile(src, None, C.byref(code), C.byref(size),
                                              C.byref(log)))
    assert size.value > 1000
    blob = C.string_at(code.value, 4)
    assert blob == b"\x7fELF"  # a gfx950 code object
    _lib.lib().skl_free_string(code)
    _lib.take_string(log)


def test_product_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "skelcl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert "import oracle" not in text and "liboracle" not in text, f
                assert "from oracle" not in text, f


def test_hip_extension_required():
    """Loading fails loudly without the library (no silent CPU path)."""
    import importlib
    saved = _lib.LIB_PATH
    try:
        _lib.LIB_PATH = "/nonexistent/libskelcl_amd.so"
        old = _lib._lib
        _lib._lib = None
        with pytest.raises(RuntimeError):
            _lib.lib()
    finally:
        _lib.LIB_PATH = saved
        _lib._lib = old
        importlib.invalidate_caches()


def test_cpp_api_headers_compile():
    """The drop-in C++ headers and the restated reference test suite compile
    with g++ (host only)."""
    src = os.path.join(ROOT, "tests", "cpp", "skeleton_tests.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "skelcl_amd", "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_segment_offsets_checked_on_host():
    """Host offsets are validated before anything reaches the device."""
    import skelcl_amd as skl
    from skelcl_amd._lib import SkelError
    v = skl._segment_offsets([0, 3, 3, 10], 10)
    assert len(v) == 4
    for bad, n in (([0, 5, 4], 10), ([0, 11], 10), ([-1, 2], 10), ([], 10)):
        with pytest.raises(SkelError):
            skl._segment_offsets(bad, n)


def _compile(src):
    code, size, log = C.c_void_p(), C.c_size_t(), C.c_void_p()
    _lib.check(_lib.lib().skl_program_compile(src, None, C.byref(code), C.byref(size),
                                              C.byref(log)))
    blob = C.string_at(code.value, size.value)
    _lib.lib().skl_free_string(code)
    _lib.take_string(log)
    return blob


def test_code_object_cache_is_opt_in(tmp_path, monkeypatch):
    """SKELCL_SAVE_BINARY / SKELCL_LOAD_BINARY act only when they are "YES"
    (src/Program.cpp:139,256): nothing is written or read otherwise; with
    SAVE=YES one code object is written under SKELCL_CACHE_DIR; with
    LOAD=YES a compile returns that file's bytes."""
    src = b'extern "C" __global__ void k(int* o){ o[threadIdx.x] = 7; }'
    monkeypatch.setenv("SKELCL_CACHE_DIR", str(tmp_path))
    for v in (None, "NO", "yes", "1"):
        for name in ("SKELCL_SAVE_BINARY", "SKELCL_LOAD_BINARY"):
            if v is None:
                monkeypatch.delenv(name, raising=False)
            else:
                monkeypatch.setenv(name, v)
        _compile(src)
        assert list(tmp_path.iterdir()) == [], v
    monkeypatch.setenv("SKELCL_SAVE_BINARY", "YES")
    monkeypatch.delenv("SKELCL_LOAD_BINARY")
    real = _compile(src)
    files = list(tmp_path.iterdir())
    assert len(files) == 1 and files[0].suffix == ".co"
    assert files[0].read_bytes() == real
    # plant a marker: only LOAD=YES returns it
    files[0].write_bytes(b"\x7fELF-cached-marker")
    monkeypatch.delenv("SKELCL_SAVE_BINARY")
    assert _compile(src) == real
    monkeypatch.setenv("SKELCL_LOAD_BINARY", "YES")
    assert _compile(src) == b"\x7fELF-cached-marker"
    # another source has another key: compiled, not loaded
    other = _compile(src.replace(b"7", b"8"))
    assert other.startswith(b"\x7fELF") and other != b"\x7fELF-cached-marker"
