"""Build recipe for the package's native code (gfx950 HIP library, host facade driver).

Outputs, in-tree so they travel with the source tree:
  3ddg_amd/libdg3d.so   C ABI of include/dg3d.h (HIP kernels + host setup)
  3ddg_amd/dg3d_driver  C++ driver with the reference's Grid/Block/Element call surface
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# packed-FP32 instructions are disabled for every HIP translation unit; multiply-add fusion
# follows the source expressions only (-ffp-contract=on), so every kernel variant computes
# bitwise-identical results
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=on"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result", f"-I{ROOT}/include", f"-I{CSRC}"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _run_parallel(cmds, jobs):
    """Run the commands, at most `jobs` at a time; raise if any fails (after all have ended)."""
    pending = list(cmds)
    running = []
    failed = []
    while pending or running:
        while pending and len(running) < jobs:
            cmd = pending.pop(0)
            print("+", " ".join(cmd), flush=True)
            running.append((cmd, subprocess.Popen(cmd)))
        cmd, proc = running.pop(0)
        if proc.wait() != 0:
            failed.append(cmd)
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])


# The kernel source is compiled once per order (the stage kernels of order n-1 with
# -DDG3D_TU_ORDER=n) and once for the order dispatch and the order-independent kernels.
KERNEL_UNITS = [("common", [])] + [(f"n{n}", [f"-DDG3D_TU_ORDER={n}"]) for n in range(2, 9)]


def build_library(verbose_resources=False, jobs=None):
    objdir = os.path.join(PKG, "_obj")
    os.makedirs(objdir, exist_ok=True)
    jobs = jobs or max(1, min(8, os.cpu_count() or 1))
    extra = ["-Rpass-analysis=kernel-resource-usage"] if verbose_resources else []
    cmds, objs = [], []
    for tag, defs in KERNEL_UNITS:
        obj = os.path.join(objdir, f"dg3d_kernels_{tag}.o")
        cmds.append([HIPCC, "-c", f"--offload-arch={ARCH}", *COMMON, *DEVICE_FLAGS, *defs, *extra,
                     os.path.join(CSRC, "dg3d_kernels.hip"), "-o", obj])
        objs.append(obj)
    for src in ["dg3d_capi.cpp", "dg3d_batch.cpp", "dg3d_setup.cpp", "basis.cpp"]:
        obj = os.path.join(objdir, src + ".o")
        cmds.append([HIPCC, "-x", "c++", "-c", *COMMON, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                     os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run_parallel(cmds, jobs)
    out = os.path.join(PKG, "libdg3d.so")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", *objs, "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_driver():
    host = os.path.join(CSRC, "host")
    if not os.path.isdir(host):
        return None
    srcs = sorted(os.path.join(host, f) for f in os.listdir(host) if f.endswith(".cpp"))
    if not srcs:
        return None
    out = os.path.join(PKG, "dg3d_driver")
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT}/include", f"-I{host}", *srcs,
          f"-L{PKG}", "-ldg3d", "-Wl,-rpath,$ORIGIN", "-ldl", "-o", out])
    return out


def build_stream_probe():
    """The streaming-ceiling probe that bench.py runs beside the roofline (a program of its own)."""
    out = os.path.join(PKG, "dg3d_stream_probe")
    _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", *DEVICE_FLAGS,
          os.path.join(CSRC, "stream_probe.hip"), "-o", out])
    return out


def build_all(verbose_resources=False):
    lib = build_library(verbose_resources)
    drv = build_driver()
    build_stream_probe()
    return lib, drv


if __name__ == "__main__":
    build_all(verbose_resources="-v" in sys.argv)
