"""Build the native parts of mithra_amd in-tree (no JIT cache, nothing installed).

  mithra_amd/lib/libmithra_gpu.so   HIP kernels + device C-ABI (include/mithra_gpu.h), gfx950
  mithra_amd/lib/libmithra.so       C++ host drop-in: job-file parser, Solver/FdTd set-up and
                                    main loop over the device C-ABI (include/mithra_host.h)
  mithra_amd/bin/mithra             the `mithra <job-file>` command

Every HIP translation unit is compiled with -ffp-contract=off (the FP64 path reproduces the
reference's arithmetic order, which has no fused multiply-adds) and with packed FP32
instructions disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GPU_SRC = os.path.join(HERE, "csrc", "gpu")
HOST_SRC = os.path.join(HERE, "csrc", "host")
LIB = os.path.join(HERE, "lib")
BIN = os.path.join(HERE, "bin")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("MITHRA_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = ["-std=c++17", "-O3", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics",
             "-Wno-unused-result"]
CXX = shutil.which("g++") or "g++"
CXX_FLAGS = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-variable"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def _sources(d, exts):
    return sorted(os.path.join(d, f) for f in os.listdir(d) if os.path.splitext(f)[1] in exts)


def build_gpu(jobs=8):
    os.makedirs(os.path.join(OBJ, "gpu"), exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    srcs = _sources(GPU_SRC, {".hip"})
    objs = [os.path.join(OBJ, "gpu", os.path.basename(s) + ".o") for s in srcs]
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([HIPCC, *HIP_FLAGS, "-c", so[0], "-o", so[1]]), zip(srcs, objs)))
    out = os.path.join(LIB, "libmithra_gpu.so")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", *objs, "-o", out])
    return out


def build_host(jobs=8):
    srcs = _sources(HOST_SRC, {".cpp"})
    if not srcs:
        return None
    os.makedirs(os.path.join(OBJ, "host"), exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    lib_srcs = [s for s in srcs if os.path.basename(s) != "main.cpp"]
    objs = [os.path.join(OBJ, "host", os.path.basename(s) + ".o") for s in lib_srcs]
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + HOST_SRC]
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda so: _run([CXX, *CXX_FLAGS, *inc, "-c", so[0], "-o", so[1]]), zip(lib_srcs, objs)))
    lib = os.path.join(LIB, "libmithra.so")
    _run([CXX, "-shared", *objs, "-L" + LIB, "-lmithra_gpu", "-Wl,-rpath,$ORIGIN", "-o", lib])
    main = os.path.join(HOST_SRC, "main.cpp")
    if os.path.exists(main):
        _run([CXX, *CXX_FLAGS, *inc, main, "-L" + LIB, "-lmithra", "-lmithra_gpu",
              "-Wl,-rpath,$ORIGIN/../lib", "-o", os.path.join(BIN, "mithra")])
    return lib


def build_oracle():
    """The CPU oracle (test infrastructure) and, where the reference sources exist, the
    reference build the fixtures come from."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "cpu_ref"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])


def build_all():
    build_gpu()
    build_host()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", os.path.join(LIB, "libmithra_gpu.so"))
