"""Builds libmtgpu.so (HIP, gfx950) in-tree, and the CPU oracle used by the tests."""
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
LIB = os.path.join(PKG, "libmtgpu.so")

# -fwrapv: i32 arithmetic wraps as in the reference's release build.
# -ffp-contract=off: every f32/f64 product and sum rounds on its own, as the reference's do.
# -packed-fp32-ops off: no v_pk_* f32 instructions.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-fwrapv",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
SOURCES = [os.path.join(PKG, "csrc", f) for f in ("mtgpu.hip", "mt_rate.hip", "mt_pred.hip", "mt_dist.hip", "mt_eval.hip")]


def build_library(verbose=False):
    cmd = [HIPCC] + HIP_FLAGS + ["-o", LIB + ".tmp"] + SOURCES
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out)
    if verbose and out.strip():
        print(out)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_oracle():
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-B", "liboracle.so"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def build_all(verbose=False):
    build_oracle()
    return build_library(verbose)


if __name__ == "__main__":
    print(build_all(verbose=True))
