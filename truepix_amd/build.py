"""Build libtruepix_hip.so (gfx950) in-tree with hipcc.

    python -m truepix_amd.build            # incremental
    python -m truepix_amd.build --clean    # rebuild everything

Every .hip under csrc/ is compiled to an object with hipcc and linked into one
shared library exporting the C ABI declared in include/truepix_hip.h.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libtruepix_hip.so")
ARCH = os.environ.get("TRUEPIX_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-Wall", "-Wno-unused-function",
    # packed-FP32 instructions stay off (none of this code is floating point,
    # and the compiler must not form them from anything that is)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


CXX = shutil.which("g++") or "g++"
# host-only sources (.cpp): the transcript hash and the proof writer
CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-pthread", "-Wall", "-Wno-unused-function"]


def _sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _headers_mtime():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    return max((os.path.getmtime(h) for h in hs), default=0.0)


def _compile(src: str, force: bool) -> str:
    s = os.path.join(CSRC, src)
    o = os.path.join(OBJ, src.rsplit(".", 1)[0] + (".o" if src.endswith(".hip") else ".host.o"))
    if (not force and os.path.isfile(o)
            and os.path.getmtime(o) >= max(os.path.getmtime(s), _headers_mtime())):
        return o
    if src.endswith(".cpp"):
        cmd = [CXX] + CXXFLAGS + ["-c", s, "-o", o]
    else:
        cmd = [HIPCC] + CFLAGS + ["-c", s, "-o", o]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, out))
    return o


def build(force: bool = False, verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args(argv)
    build(force=a.clean)


if __name__ == "__main__":
    sys.exit(main())
