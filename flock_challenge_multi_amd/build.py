"""Build the gfx950 device library `libflock_gpu.so` in-tree with hipcc.

Every `.hip` source under csrc/ is compiled for gfx950 and linked into one
shared library next to this file.  Flags:
  -Xclang -target-feature -Xclang -packed-fp32-ops   keeps hipcc from emitting
      packed fp32 ops (none of this integer code needs them);
  -munsafe-fp-atomics is not used: the only atomics are integer XORs.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libflock_gpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src, obj):
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature; harmless.
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return err


def _compile_host(src, obj):
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-Wall", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {src}:\n{r.stderr}")
    return r.stderr


def build(verbose=False):
    """.hip sources go through hipcc for gfx950; .cpp sources (host-only code,
    the Fiat-Shamir transcript) through g++; one shared library."""
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    hsrcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".cpp"))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    hobjs = [os.path.join(objdir, os.path.basename(s)[:-4] + ".host.o") for s in hsrcs]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs) + len(hsrcs))) as ex:
        errs = list(ex.map(_compile, srcs, objs)) + list(ex.map(_compile_host, hsrcs, hobjs))
    objs = objs + hobjs
    if verbose:
        for e in errs:
            if e.strip():
                print(e)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


def build_host_mirror_test():
    """Compile tests/cpp/host_mirror_test (the C++ host mirror's parity program,
    linked against this library and the oracle).  Needs oracle/liboracle.so."""
    root = os.path.dirname(HERE)
    exe = os.path.join(root, "tests", "cpp", "host_mirror_test")
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(root, "include"), "-I", os.path.join(root, "oracle"),
           os.path.join(root, "tests", "cpp", "host_mirror_test.cpp"), "-o", exe,
           "-L", HERE, "-lflock_gpu", "-L", os.path.join(root, "oracle"), "-loracle",
           "-Wl,-rpath,$ORIGIN/../../flock_challenge_multi_amd", "-Wl,-rpath,$ORIGIN/../../oracle"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host mirror test build failed:\n{r.stderr}")
    return exe


def build_witness_lane_test():
    """Compile tests/cpp/witness_lane_test (host check of the per-compression
    witness code in csrc/witness_lane.hpp against the oracle).  Needs
    oracle/liboracle.so."""
    root = os.path.dirname(HERE)
    exe = os.path.join(root, "tests", "cpp", "witness_lane_test")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", "-I", os.path.join(root, "oracle"),
           "-I", CSRC, os.path.join(root, "tests", "cpp", "witness_lane_test.cpp"), "-o", exe,
           "-L", os.path.join(root, "oracle"), "-loracle", "-Wl,-rpath,$ORIGIN/../../oracle"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"witness lane test build failed:\n{r.stderr}")
    return exe


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
