"""Build recipe of libskelcl_amd.so (hipcc, gfx950).

  python -m skelcl_amd.build          # or skelcl_amd.build.build()

Steps: embed the kernel templates (csrc/kernels/*.hip) as C strings, compile
every csrc/*.cpp and csrc/native/*.hip with hipcc --offload-arch=gfx950 and
link the shared library in-tree next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build", "skelcl_amd")
LIB = os.path.join(HERE, "libskelcl_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SKELCL_ARCH", "gfx950")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off in every device code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]

TEMPLATES = ["prelude", "map", "reduce", "scan", "mapoverlap", "allpairs"]

# Per-source options.  matmul.hip: MFMA accumulators in the VGPR form; with
# the default (AGPR) form the register allocator gives each MFMA a result
# register other than its accumulator input and copies the accumulators
# round the k-loop (hundreds of v_accvgpr moves per slice).
EXTRA = {"matmul.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
         "batch.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"],
         # the weighted-stencil forward fuses only where it calls fmaf
         "stencil.hip": ["-ffp-contract=off"]}


def _embed() -> str:
    out = os.path.join(BUILD, "embedded_kernels.cpp")
    parts = ["// generated by skelcl_amd/build.py from csrc/kernels/*.hip\n"]
    for name in TEMPLATES:
        with open(os.path.join(CSRC, "kernels", name + ".hip")) as f:
            text = f.read()
        assert ")SKLRAW\"" not in text
        parts.append(f'extern "C" const char* skl_tmpl_{name} = R"SKLRAW({text})SKLRAW";\n')
    data = "".join(parts)
    old = None
    if os.path.exists(out):
        with open(out) as f:
            old = f.read()
    if old != data:
        with open(out, "w") as f:
            f.write(data)
    return out


def _sources() -> list[str]:
    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".cpp")]
    nat = os.path.join(CSRC, "native")
    srcs += [os.path.join(nat, f) for f in sorted(os.listdir(nat)) if f.endswith(".hip")]
    return srcs


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    lang = ["-x", "hip"]
    cmd = [HIPCC, *COMMON, *EXTRA.get(os.path.basename(src), []), *lang, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources() + [_embed()]
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *COMMON, "-shared", "-o", tmp, *objs, "-lhiprtc", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB




def build_cpp_tests(verbose: bool = False) -> str:
    """Compile tests/cpp (the reference's skeleton, container and device
    tests restated against the C++ API) with g++ against libskelcl_amd.so;
    host code only."""
    srcs = [os.path.join(ROOT, "tests", "cpp", f) for f in ("skeleton_tests.cpp", "api_tests.cpp")]
    out = os.path.join(ROOT, "tests", "cpp", "skelcl_cpp_tests")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(HERE, "include"), "-o", out, *srcs, "-L", HERE, "-lskelcl_amd",
           "-Wl,-rpath,$ORIGIN/../../skelcl_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ tests failed to compile:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {out}")
    return out


# The reference's example programs (examples/*/main.cpp), compiled unchanged
# against this package's headers and library -- the drop-in check on the
# reference's own callers.  They are read where they lie (never copied) with
# the reference's pvsutil helper library (libraries/pvsutil/src/*.cpp, 8
# plain C++ files) for their logging and command-line parsing; the binaries
# go to tests/cpp/_ref_examples/ (git-ignored).  Nothing is built when the
# reference tree is absent.
REFERENCE = os.environ.get("SKELCL_REFERENCE", "/root/reference")
REF_EXAMPLES = ("dot_product", "saxpy", "matrix_mult", "montecarlo", "gauss", "mandelbrot")
REF_EXAMPLE_DIR = os.path.join(ROOT, "tests", "cpp", "_ref_examples")


def build_reference_examples(verbose: bool = False) -> list[str]:
    ex_dir = os.path.join(REFERENCE, "examples")
    pvs = os.path.join(REFERENCE, "libraries", "pvsutil")
    if not (os.path.isdir(ex_dir) and os.path.isdir(pvs)):
        if verbose:
            print(f"reference examples not built: {REFERENCE} is absent")
        return []
    objdir = os.path.join(ROOT, "_build", "pvsutil")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(REF_EXAMPLE_DIR, exist_ok=True)
    inc_pvs = os.path.join(pvs, "include")
    srcs = sorted(f for f in os.listdir(os.path.join(pvs, "src")) if f.endswith(".cpp"))

    def run(cmd):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")

    objs = []
    for f in srcs:
        o = os.path.join(objdir, f[:-4] + ".o")
        run(["g++", "-std=c++11", "-O1", "-c", "-I", inc_pvs, os.path.join(pvs, "src", f),
             "-o", o])
        objs.append(o)
    built = []
    for ex in REF_EXAMPLES:
        out = os.path.join(REF_EXAMPLE_DIR, ex)
        run(["g++", "-std=c++11", "-O1", "-I", os.path.join(ROOT, "include"),
             "-I", os.path.join(HERE, "include"), "-I", inc_pvs,
             os.path.join(ex_dir, ex, "main.cpp"), *objs, "-L", HERE, "-lskelcl_amd",
             "-Wl,-rpath,$ORIGIN/../../../skelcl_amd", "-o", out])
        built.append(out)
    if verbose:
        print(f"built {len(built)} reference examples in {REF_EXAMPLE_DIR}")
    return built


if __name__ == "__main__":
    build(verbose=True)
    build_cpp_tests(verbose=True)
    build_reference_examples(verbose=True)
    sys.exit(0)
