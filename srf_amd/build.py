"""Build libsrf_hip.so (gfx950) in-tree with hipcc.

    python -m srf_amd.build        # or srf_amd.build.build()

Flags:
  --offload-arch=gfx950            MI355X (CDNA4) only
  -ffp-contract=off                the marcher's index arithmetic must round
                                   every operation separately (explicit fmaf
                                   where fusion is wanted)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                   no packed-FP32 instructions are generated
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsrf_hip.so")

SOURCES = ["srf_project.hip", "srf_elementwise.hip", "srf_tiled.hip", "srf_sss.hip", "srf_cnn.hip", "srf_geometry.hip"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return [
        "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    ]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    cmd = [hipcc()] + flags() + srcs + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed (%d)" % r.returncode)
    # keep only real diagnostics (the host pass does not know the packed-fp32 feature)
    noise = ("packed-fp32-ops", "warning: unknown target feature", "1 warning generated",
             "warnings generated")
    msgs = [l for l in (r.stdout + r.stderr).splitlines() if l.strip() and not any(n in l for n in noise)]
    if msgs and verbose:
        sys.stderr.write("\n".join(msgs) + "\n")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
