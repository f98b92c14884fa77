"""Builds libmtgpu.so (HIP, gfx950) in-tree, and the CPU oracle used by the tests."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

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
SOURCES = [os.path.join(PKG, "csrc", f) for f in ("mtgpu.hip", "mt_rate.hip", "mt_pred.hip", "mt_dist.hip", "mt_eval.hip", "mt_rdo.hip", "mt_nn.hip", "mt_encode.hip", "mt_deblock.hip", "mt_price.hip", "mt_cdef.hip", "mt_av1enc.cpp")]


def _compile(src):
    obj = os.path.join(PKG, "csrc", "_obj", os.path.basename(src) + ".o")
    cmd = [HIPCC] + [f for f in HIP_FLAGS if f != "-shared"] + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return obj, r


def build_library(verbose=False):
    """Compiles every source for gfx950 (one process per file) and links libmtgpu.so."""
    os.makedirs(os.path.join(PKG, "csrc", "_obj"), exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        results = list(ex.map(_compile, SOURCES))
    msgs = []
    for obj, r in results:
        out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed:\n" + out)
        if out.strip():
            msgs.append(out)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    if verbose and msgs:
        print("\n".join(msgs))
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
