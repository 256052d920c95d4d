"""Build libcm3d_nb.so (gfx950) in-tree with hipcc.

    python -m cm3d_amd.build

No GPU is needed to build: hipcc cross-compiles for gfx950.  Flags:
  -ffp-contract=off  no FMA contraction by default, so the FP64 minimum
                     image, r^2, cut-off tests and table index round exactly
                     as cm3d's x86-64 build (list membership and the table
                     node of a pair are decided bit for bit on them); the
                     interpolation, the sums and the Ewald kernels opt in to
                     FMA with `#pragma clang fp contract(fast)`;
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 ops.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcm3d_nb.so")
SOURCES = ["nb_api.hip", "nb_kernels.hip", "nb_mixed.hip", "nb_md.hip", "nb_batch.hip",
           "nb_setup.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def command(out=OUT):
    srcs = [os.path.join(SRC, s) for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    return [HIPCC, "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
            "-packed-fp32-ops", "-Wno-unused-command-line-argument",
            "-o", out] + srcs


def build(verbose=False):
    cmd = command()
    if verbose:
        print(" ".join(cmd), flush=True)
    tmp = OUT + ".tmp"
    cmd[cmd.index("-o") + 1] = tmp
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libcm3d_nb.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
