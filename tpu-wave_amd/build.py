"""Build libtwave.so (HIP, gfx950) in-tree: tpu-wave_amd/twave/libtwave.so.

The kernels (csrc/tw_kernels.hip) are compiled once per field storage type
(-DTW_PREC=0..3) and, with the host sources, in parallel; objects go to
tpu-wave_amd/_build/."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OUT = os.path.join(HERE, "twave", "libtwave.so")
OBJ = os.path.join(HERE, "_build")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # IEEE per-operation rounding: the reference's operation order is kept
    # and no a*b+c is fused, so results match the CPU restatement bit for bit
    "-ffp-contract=off",
    # no packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
UNITS = [("csrc/tw_kernels.hip", "tw_kernels_p%d.o" % k, ["-DTW_PREC=%d" % k]) for k in range(4)] + [
    ("csrc/tw_dispatch.cpp", "tw_dispatch.o", []),
    ("csrc/tw_noalloc.hip", "tw_noalloc.o", []),
    ("csrc/tw_host.cpp", "tw_host.o", []),
    ("csrc/tw_io.cpp", "tw_io.o", []),
]


def _clean(text):
    return "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    for src, obj, extra in UNITS:
        cmd = [HIPCC] + FLAGS + extra + ["-c", os.path.join(HERE, src), "-o", os.path.join(OBJ, obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, cwd=HERE, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    failed = []
    for src, pr in procs:
        out, _ = pr.communicate()
        msg = _clean(out)
        if pr.returncode != 0:
            failed.append(src)
            sys.stderr.write(msg + "\n")
    if failed:
        raise RuntimeError("hipcc failed: %s" % ", ".join(failed))
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", OUT] + [os.path.join(OBJ, o) for _, o, _ in UNITS]
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(_clean(r.stdout + r.stderr) + "\n")
        raise RuntimeError("link failed (%d)" % r.returncode)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
