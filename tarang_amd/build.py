"""Build libtarang_hip.so (gfx950) in-tree with hipcc, plus the C++ host driver.

Usage: python -m tarang_amd.build [--jobs N] [--debug]
Every .hip translation unit is compiled for gfx950 only, with packed-FP32 instruction selection
disabled, then linked into tarang_amd/libtarang_hip.so.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtarang_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["th_api.hip", "th_generic.hip", "th_reduce.hip", "th_ic.hip", "th_fused.hip", "th_fused_f32.hip",
           "th_fused_f64.hip", "th_force.hip", "th_probe.hip"]
HOST_SOURCES = ["th_host.cpp", "th_io.cpp"]

COMMON = ["-std=c++17", "-fPIC", "-O3", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
          "-Wno-unused-result", "-Wno-pass-failed"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + "\n")
        raise RuntimeError("compile failed: " + cmd[-1])
    return r.stdout


def _stale(src, obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in [src] + deps)


def build(jobs=8, force=False, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    headers = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    headers.append(os.path.join(ROOT, "include", "tarang_hip.h"))
    jobs_list = []
    objs = []
    for s in SOURCES + HOST_SOURCES:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        obj = os.path.join(BUILD, s + ".o")
        objs.append(obj)
        if force or _stale(src, obj, headers):
            cmd = [HIPCC] + COMMON + ["-c", "-o", obj, src]
            if s.endswith(".cpp"):
                cmd = [HIPCC, "-std=c++17", "-fPIC", "-O2", "--offload-arch=gfx950", "-I" + os.path.join(ROOT, "include"),
                       "-I" + CSRC, "-c", "-o", obj, src]
            jobs_list.append(cmd)
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(_run, jobs_list):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # C++ driver over the host classes (links the library in place, rpath $ORIGIN)
    drv = os.path.join(HERE, "tarang_run")
    _run([HIPCC, "-std=c++17", "-O2", "--offload-arch=gfx950", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
          "-o", drv + ".tmp", os.path.join(CSRC, "tarang_run.cpp"), "-L" + HERE, "-ltarang_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(drv + ".tmp", drv)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.force, a.v))
