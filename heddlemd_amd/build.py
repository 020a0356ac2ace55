"""Build recipe for libheddle_hip.so (the analogue of the reference's build.rs,
which compiles kernels/*.cu with nvcc and embeds the PTX).

Steps:
  1. hipcc --genco compiles every kernel (csrc/kernels/all_kernels.hip) into
     one gfx950 code object;
  2. the code object is embedded into the library with an assembler .incbin
     (symbol hmd_code_object) and loaded at run time by hipModuleLoadData;
  3. the C++17 host layer (csrc/host/*.cpp) is compiled with g++ against the
     HIP runtime API and linked into one shared library with a C ABI
     (include/heddle_hip.h).

Run ``python -m heddlemd_amd.build`` or call ``build()``.  The compile works
without a GPU (hipcc cross-compiles gfx950).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libheddle_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIPCC_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "--genco",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", f"-I{ROOT}/include"]

HOST_SOURCES = ["context.cpp", "forcefield.cpp", "integrator.cpp", "nn.cpp", "io.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]}")
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
    bundle = os.path.join(BUILD, "kernels.bundle")
    _run([hipcc, *HIPCC_FLAGS, "-o", bundle, os.path.join(CSRC, "kernels", "all_kernels.hip")])
    # --genco emits an offload bundle; embed the raw gfx950 ELF code object.
    hsaco = os.path.join(BUILD, "kernels.hsaco")
    bundler = os.path.join(ROCM, "llvm", "bin", "clang-offload-bundler")
    _run([bundler, "--unbundle", "--type=o",
          f"--targets=hipv4-amdgcn-amd-amdhsa--{ARCH}",
          f"--input={bundle}", f"--output={hsaco}"])
    blob = os.path.join(BUILD, "kernels_blob.S")
    with open(blob, "w") as f:
        f.write(
            '\t.section .rodata.hmd_code_object,"a",@progbits\n'
            "\t.p2align 12\n"
            "\t.globl hmd_code_object\n"
            "\t.type hmd_code_object,@object\n"
            "hmd_code_object:\n"
            f'\t.incbin "{hsaco}"\n'
            "\t.byte 0\n"
            '\t.section .note.GNU-stack,"",@progbits\n')
    objs = [os.path.join(BUILD, "kernels_blob.o")]
    _run(["gcc", "-c", "-o", objs[0], blob])
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
        _run(["g++", *CXX_FLAGS, "-c", "-o", obj, os.path.join(CSRC, "host", src)])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-o", tmp, *objs, f"-L{ROCM}/lib", "-lamdhip64",
          f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
