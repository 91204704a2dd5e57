"""Builds the HIP C-ABI library lib/libnsgpu.so for gfx950 (MI355X) with hipcc.

ns_kernels.hip is compiled three times: with -ffp-contract=off (EXACT: the reference's operation
order, bitwise), with -ffp-contract=on (FAST: fused multiply-add within each source expression, so
every instantiation of a kernel body fuses the same operations and the single-block and ensemble
kernels agree bitwise) and with single-precision reconstruction (FP32, same contraction).  Every HIP compile disables packed FP32
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib")
ARCH = os.environ.get("NSGPU_ARCH", "gfx950")

COMMON = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-fno-gpu-flush-denormals-to-zero",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed: " + " ".join(cmd))
    return r


def build(verbose=False, defines=(), out=None):
    """Compiles and links the library; `defines` are extra -D flags for the kernel units (tuning
    variants), `out` an alternative output directory."""
    OUT_ = out or OUT
    os.makedirs(OUT_, exist_ok=True)
    objs = []
    units = [
        ("ns_kernels.hip", "kernels_exact.o", ["-DNS_FAST=0", "-ffp-contract=off", *defines]),
        ("ns_kernels.hip", "kernels_fast.o", ["-DNS_FAST=1", "-ffp-contract=on", *defines]),
        ("ns_kernels.hip", "kernels_fp32.o", ["-DNS_FAST=1", "-DNS_FP32=1", "-ffp-contract=on",
                                              *defines]),
        ("nsgpu_capi.hip", "capi.o", ["-ffp-contract=off"]),
        ("ns_mlp.hip", "mlp.o", ["-ffp-contract=on"]),
    ]
    cmds = []
    for src, obj, extra in units:
        o = os.path.join(OUT_, obj)
        cmds.append(["hipcc", "-c", *COMMON, *extra, os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd))
    # the units are independent: compile them concurrently
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    lib = os.path.join(OUT_, "libnsgpu.so")
    tmp = lib + ".tmp"
    _run(["hipcc", "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser()
    ap.add_argument("-D", dest="defines", action="append", default=[])
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    print(build(verbose=True, defines=[f"-D{d}" for d in a.defines], out=a.out))
