"""Build recipe for the native library (librover.so) and the CPU oracle.

librover.so = HIP kernels for gfx950 (hipcc) + the C host library (gcc),
linked in-tree under rover_amd/lib/ so it travels with the source tree.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "librover.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ROVER_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every float op rounds separately, as the reference's
# scalar C does; packed fp32 ops are kept off for gfx950.
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
C_FLAGS = ["-std=gnu11", "-O2", "-fPIC", "-ffp-contract=off", "-Wall",
           "-Wno-unused-function"]

HIP_SOURCES = ["rover_search.hip", "rover_learned.hip", "rover_build.hip",
               "rover_build_knn.hip", "rv_radix.hip"]
C_SOURCES = ["rover_host.c", "rover_learned_host.c", "rover_sha256.c"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True)
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    if verbose and out.strip():
        print(out)


def build_native(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objs, cmds = [], []
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    for src in HIP_SOURCES + C_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(LIBDIR, src + ".o")
        if src.endswith(".hip"):
            cmds.append([HIPCC] + HIP_FLAGS + inc + ["-c", path, "-o", obj])
        else:
            cmds.append(["gcc"] + C_FLAGS + inc + ["-c", path, "-o", obj])
        objs.append(obj)
    # translation units are independent: compile them side by side
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp]
         + objs + ["-lpthread", "-lm"], verbose)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], verbose)


def build(verbose=False):
    build_oracle(verbose)
    return build_native(verbose)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
