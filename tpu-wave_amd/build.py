"""Build libtwave.so (HIP, gfx950) in-tree: tpu-wave_amd/twave/libtwave.so."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["csrc/tw_host.cpp", "csrc/tw_io.cpp", "csrc/tw_kernels.hip"]
OUT = os.path.join(HERE, "twave", "libtwave.so")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # IEEE per-operation rounding: the reference's operation order is kept
    # and no a*b+c is fused, so results match the CPU restatement bit for bit
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]


def build(verbose=False):
    cmd = [HIPCC] + FLAGS + ["-o", OUT] + [os.path.join(HERE, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        sys.stderr.write("\n".join(msgs) + "\n")
        raise RuntimeError("hipcc failed (%d)" % r.returncode)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
