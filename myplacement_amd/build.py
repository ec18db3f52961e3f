"""Builds the in-tree shared libraries.

  myplacement_amd/libmp_gp_hip.so   HIP kernels for gfx950 + host C++ + the C ABI (hipcc)
  oracle/liboracle.so               CPU checker (g++), test infrastructure
  oracle/_ref/libmpref.so           reference sources + shim (only where the reference tree exists)
  myplacement_amd/libmp_gp_hip_hostsan.so
                                    the same with AddressSanitizer and UndefinedBehaviorSanitizer on
                                    the host code, for the CPU tests of the C ABI

Run as `python -m myplacement_amd.build` or through __graft_entry__.build().
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmp_gp_hip.so")
HOSTSAN_LIB = os.path.join(HERE, "libmp_gp_hip_hostsan.so")
SOURCES = ["gp_kernels.hip", "gp_backend.cpp", "gp_host.cpp", "gp_session.cpp", "gp_batch.cpp", "gp_quadratic.hip", "gp_serve.cpp",
           "gp_detailed.hip"]
ARCH = os.environ.get("MP_GP_ARCH", "gfx950")
REFERENCE = "/root/reference"


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def hip_flags() -> list:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-munsafe-fp-atomics",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
        # the host replays the device's step decision bit for bit: no fused multiply-adds on the
        # host side whatever the target (device code sets contraction per kernel)
        "-Xarch_host", "-ffp-contract=off",
    ]


def hostsan_flags() -> list:
    """AddressSanitizer and UndefinedBehaviorSanitizer on the host code only (the device code is
    compiled as usual: the host registers it); a process that loads the library preloads the
    sanitizer runtime (sanitizer_runtime()).  The library is for CPU tests of the C ABI."""
    return ("-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined "
            "-Xarch_host -fno-sanitize=vptr -Xarch_host -fno-sanitize-recover=undefined "
            "-Xarch_host -fno-omit-frame-pointer -Xarch_host -g").split()


def sanitizer_runtime() -> str:
    """The shared AddressSanitizer runtime of hipcc's clang (it carries the UBSan handlers too)."""
    r = subprocess.run([hipcc(), "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path):
        return path
    rocm = os.path.dirname(os.path.dirname(os.path.realpath(hipcc())))
    for root, _, files in os.walk(os.path.join(rocm, "llvm", "lib", "clang")):
        if "libclang_rt.asan-x86_64.so" in files:
            return os.path.join(root, "libclang_rt.asan-x86_64.so")
    return ""


def build_hip(verbose: bool = False, hostsan: bool = False) -> str:
    objs = []
    obj_dir = os.path.join(ROOT, "_build", "hostsan" if hostsan else "")
    os.makedirs(obj_dir, exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
        extra = hostsan_flags() if hostsan else []
        cmd = [hipcc()] + hip_flags() + extra + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, cmd, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip())
        if p.returncode != 0:
            failed.append(f"{src}:\n{out}")
        elif verbose and out:
            print(out)
    if failed:
        raise RuntimeError("HIP build failed:\n" + "\n".join(failed))
    lib = HOSTSAN_LIB if hostsan else LIB
    tmp = lib + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("HIP link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, lib)
    return lib


def build_oracle() -> None:
    targets = ["oracle"]
    if os.path.isdir(REFERENCE):
        targets.append("ref")
    r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")] + targets, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def build(verbose: bool = False) -> None:
    import threading
    errors = []

    def hostsan():
        try:
            build_hip(verbose, hostsan=True)
        except Exception as e:  # reported after the product build
            errors.append(e)

    side = threading.Thread(target=hostsan)
    side.start()
    build_hip(verbose)
    side.join()
    if errors:
        raise errors[0]
    build_oracle()


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB)
