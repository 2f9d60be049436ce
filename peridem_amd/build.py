"""In-tree build of the HIP library (hipcc, gfx950).

The library is built next to this file as libperidem_gpu.so.  Compile flags:
  -ffp-contract=off   no fused multiply-add: the FP64 results stay bitwise
                      equal to the reference's operation order
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions anywhere in the library
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libperidem_gpu.so")
ARCH = os.environ.get("PD_OFFLOAD_ARCH", "gfx950")

FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-ffp-contract=off",
         f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value"]


def build(verbose=False):
    cmd = ["hipcc", *FLAGS, "-o", OUT + ".tmp", os.path.join(SRC, "pd_api.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
