"""Build recipe for libturbo_cpp_amd.so (HIP kernels for gfx950 + host C++).

Usage: python -m turbo_cpp_amd.build   (or turbo_cpp_amd.build.build()).
Objects go to turbo_cpp_amd/_build/, the library to turbo_cpp_amd/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libturbo_cpp_amd.so")
# VALU issue calibration for bench.py's roofline (not part of the decoder library)
CALIB_SRC = os.path.join(CSRC, "calib", "valu_rate.hip")
CALIB_LIB = os.path.join(HERE, "libturbo_calib.so")
API_TEST = os.path.join(OUT_DIR, "api_test")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HOST_SOURCES = ["host/codec.cpp", "host/api.cpp"]
DEVICE_SOURCES = ["device/exact_kernel.hip", "device/window_log.hip", "device/window_log_lean.hip",
                  "device/window_lin.hip", "device/window_wide2.hip", "device/window_wide3.hip",
                  "device/window_exact.hip", "device/c_api.hip",
                  "device/rx_kernels.hip", "device/rx_api.hip", "device/tx_kernels.hip", "device/channel.hip",
                  "device/diff_kernel.hip",
                  "device/diff_api.hip", "device/noalloc.hip"]

COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include")]
# Packed FP32 (v_pk_*_f32) stays off for all device code.
DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-munsafe-fp-atomics"]


def _obj(src: str) -> str:
    return os.path.join(OUT_DIR, src.replace("/", "_") + ".o")


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = _obj(src)
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + COMMON + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = HOST_SOURCES + DEVICE_SOURCES
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE_FLAGS + ["-shared", CALIB_SRC, "-o", CALIB_LIB + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"calibration build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(CALIB_LIB + ".tmp", CALIB_LIB)
    # C++ API test program (tests/cpp/api_test.cpp) linked against the library
    test_src = os.path.join(ROOT, "tests", "cpp", "api_test.cpp")
    if os.path.exists(test_src):
        cmd = ["g++"] + COMMON + ["-pthread", test_src, "-o", API_TEST, "-L", HERE, "-lturbo_cpp_amd",
                                  "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"api_test build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
