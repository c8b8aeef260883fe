"""The C++ drop-in on the CPU: the reference's own example programs
(examples/*/main.cpp) compile unchanged against this package's headers, read
where they lie in the reference tree (skipped when it is absent), and the
headers compile as C++11 (the reference's dialect) and C++17 with warnings as
errors."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("SKELCL_REFERENCE", "/root/reference")
INC = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "skelcl_amd", "include")]
EXAMPLES = ["dot_product", "saxpy", "matrix_mult", "gauss", "mandelbrot", "montecarlo"]
HEADERS = ["SkelCL.h", "Vector.h", "Matrix.h", "Map.h", "Zip.h", "Reduce.h", "Scan.h",
           "MapOverlap.h", "AllPairs.h", "IndexVector.h", "IndexMatrix.h", "Distributions.h",
           "Local.h", "Out.h", "Source.h", "Autodiff.h", "detail/Util.h", "detail/Device.h",
           "detail/DeviceList.h", "detail/DeviceProperties.h", "detail/Macros.h"]


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "examples")),
                    reason="the reference tree is not present")
@pytest.mark.parametrize("example", EXAMPLES)
def test_reference_example_compiles_unchanged(example):
    src = os.path.join(REF, "examples", example, "main.cpp")
    pvs = os.path.join(REF, "libraries", "pvsutil", "include")  # its logging/CLI helpers
    r = subprocess.run(["g++", "-std=c++11", "-fsyntax-only", *INC, "-I", pvs, src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.parametrize("std", ["c++11", "c++17"])
def test_headers_compile_standalone(tmp_path, std):
    """Each public header compiles on its own, without warnings."""
    for h in HEADERS:
        src = tmp_path / "t.cpp"
        src.write_text(f"#include <SkelCL/{h}>\nint main() {{ return 0; }}\n")
        r = subprocess.run(["g++", f"-std={std}", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                            *INC, str(src)], capture_output=True, text=True)
        assert r.returncode == 0, (h, r.stderr[-3000:])


def test_user_struct_type_names(tmp_path):
    """typeToString / TypeName strip namespaces from demangled names, so a
    user struct and skelcl::float4 reach kernel source by their plain names
    (detail/Util.h:81-98)."""
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include <SkelCL/SkelCL.h>
#include <cstdio>
SKELCL_COMMON_DEFINITION(typedef struct { unsigned char r; unsigned char g; } Pixel;)
namespace ns { struct Deep { int a; }; }
int main() {
  std::printf("%s|%s|%s|%s|%s\n", skelcl::TypeName<Pixel>::get(),
              skelcl::TypeName<skelcl::float4>::get(), skelcl::TypeName<ns::Deep>::get(),
              skelcl::TypeName<unsigned int>::get(),
              skelcl::detail::util::typeToString<double>().c_str());
  char* t = nullptr;
  skl_common_definitions(&t);
  std::printf("%s", t);
  skl_free_string(t);
  return 0;
}
''')
    exe = tmp_path / "t"
    r = subprocess.run(["g++", "-std=c++11", *INC, str(src), "-o", str(exe),
                        "-L", os.path.join(ROOT, "skelcl_amd"), "-lskelcl_amd",
                        "-Wl,-rpath," + os.path.join(ROOT, "skelcl_amd")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60).stdout
    first, rest = out.split("\n", 1)
    assert first == "Pixel|float4|Deep|unsigned int|double"
    assert "typedef struct { unsigned char r; unsigned char g; } Pixel;" in rest
