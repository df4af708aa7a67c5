"""Builds the native libraries in-tree: liblitiv_bgs.so and liblitiv_synth.so (HIP, gfx950).

Every HIP source is compiled with -ffp-contract=off (no FMA contraction: the FP32 feedback maps
must round like the sequential definition) and with packed-FP32 instruction selection disabled.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ARCH = os.environ.get("LITIV_AMD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

TARGETS = {
    "liblitiv_bgs.so": ["bgs_kernels.hip", "bgs_api.hip", "bgs_serve.hip", "bgs_eval.hip", "bgs_archive.hip"],
    "liblitiv_synth.so": ["synth.hip"],
}


def build(verbose=False):
    outs = []
    for lib, srcs in TARGETS.items():
        out = os.path.join(HERE, lib)
        tmp = out + ".tmp"
        cmd = [HIPCC] + HIP_FLAGS + ["-o", tmp] + [os.path.join(CSRC, s) for s in srcs]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError("hipcc failed for %s:\n%s" % (lib, res.stderr[-4000:]))
        os.replace(tmp, out)
        if verbose:
            print("built", out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    build(verbose=True)
