"""Build libdfg.so (HIP kernels + C ABI) for gfx950, in-tree.

    python guiding-df_amd/build.py            # compile every source, link
    python guiding-df_amd/build.py --verbose

Objects go to guiding-df_amd/_build/, the library to guiding-df_amd/libdfg.so,
the memory-rate microbenchmark to guiding-df_amd/microbench/gather_rate.
"""
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdfg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["k_query.hip", "k_train.hip", "k_plain.hip", "k_records.hip", "api.cpp", "serve.cpp", "records.cpp", "noalloc.hip", "k_train_fused.hip", "k_serve.hip"]

# -packed-fp32-ops: keep the compiler from forming packed fp32 VALU ops.
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile_cmd(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    return obj, [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    t0 = time.time()
    procs = []
    objs = []
    for src in SOURCES:
        obj, cmd = _compile_cmd(src)
        objs.append(obj)
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and "unknown target feature" not in l]
        if p.returncode != 0:
            failed.append(src)
            print("\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed for: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    # the memory-rate microbenchmark (a standalone program, not part of the library)
    mb = os.path.join(HERE, "microbench")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", os.path.join(mb, "gather_rate.hip"), "-o", os.path.join(mb, "gather_rate")]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if p.returncode != 0:
        print(p.stdout.decode(errors="replace"), file=sys.stderr)
        raise RuntimeError("hipcc failed for microbench/gather_rate.hip")
    if verbose:
        print("built %s in %.1fs" % (LIB, time.time() - t0))
    return LIB


if __name__ == "__main__":
    build(verbose="--verbose" in sys.argv)
