"""Build the native library lib/libgpuqo_amd.so for gfx950 (in-tree).

hipcc flags:
  -ffp-contract=off                       no FMA contraction: the float cost
                                          model must round like the reference
  -fhip-fp32-correctly-rounded-divide-sqrt IEEE float division
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                          no packed-FP32 instructions
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib", "libgpuqo_amd.so")
SOURCES = ["gq_dp.hip", "gq_wide.hip", "gq_host.cpp", "gq_drivers.cpp", "gq_server.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
         "-fhip-fp32-correctly-rounded-divide-sqrt", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-result"]


BENCH_LIB = os.path.join(HERE, "lib", "libgq_bench.so")
BENCH_SOURCES = ["gq_gather_bench.hip"]


def _compile(out, sources, verbose):
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in sources]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(tmp, out)


def build(verbose=False):
    """Compile the planner library and the roofline microbenchmark."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    _compile(LIB, SOURCES, verbose)
    _compile(BENCH_LIB, BENCH_SOURCES, verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
