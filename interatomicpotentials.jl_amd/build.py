"""Builds libipgpu.so (HIP kernels for gfx950 + the C ABI) in-tree.

Usage: python interatomicpotentials.jl_amd/build.py [-j N] [--verbose]

Every .hip file under csrc/ is compiled with hipcc for gfx950 into
_build/*.o and linked into lib/libipgpu.so.  All HIP code is compiled with
packed FP32 instructions disabled (-packed-fp32-ops).
"""
import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libipgpu.so")
ARCH = os.environ.get("IPG_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(REPO, "include"),
    "-I" + CSRC,
    "-Wno-unused-command-line-argument",
]


def _includes(path, seen=None):
    """Local headers a source includes, followed recursively (quoted
    #include lines resolved against csrc/ and include/)."""
    seen = set() if seen is None else seen
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line.startswith("#include \""):
                continue
            name = line.split('"')[1]
            for d in (CSRC, os.path.join(REPO, "include")):
                h = os.path.join(d, name)
                if os.path.exists(h) and h not in seen:
                    seen.add(h)
                    _includes(h, seen)
                    break
    return seen


def _stale(src, obj, headers):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(f) > t for f in [src] + sorted(_includes(src)))


def _compile(src, verbose, incremental=False, headers=()):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if incremental and not _stale(src, obj, list(headers)):
        return obj, ""
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (src, r.stdout, r.stderr))
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "target-feature" not in l)
    return obj, err


def build(jobs=None, verbose=False, incremental=False):
    """Compiles every source (incremental=True: only those older than their
    .o or than any header) and links the library."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    headers = sorted(glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(REPO, "include", "*.h")))
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda s: _compile(s, verbose, incremental, headers), srcs):
            objs.append(obj)
            if err.strip() and verbose:
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--incremental", action="store_true", help="skip sources whose object is up to date")
    a = ap.parse_args()
    print(build(a.j, a.verbose, a.incremental))
