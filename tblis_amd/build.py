"""Builds libtblis_amd.so (host C++ + gfx950 HIP kernels) in-tree with hipcc."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtblis_amd.so")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("TBLIS_AMD_ARCH", "gfx950")

SOURCES = ["canon.cpp", "scatter.cpp", "dpd.cpp", "plan.cpp", "contract.cpp", "dpd_plan.cpp",
           "api.cpp", "panel.cpp", "level1.cpp", "indexed.cpp", "indexed_dpd.cpp", "complex.cpp", "gemm_panel.hip", "gemm_panel_bf16.hip", "gemm_lds.hip",
           "gemm_ozaki.hip", "level1_kernels.hip", "noalloc.hip"]

# Per-source device flags.  The panel GEMM keeps its MFMA accumulators in
# ordinary VGPRs (the unified gfx950 register file); the default AGPR form
# copies every accumulator in and out of AGPRs once per k iteration.
EXTRA = {"gemm_panel.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
         "gemm_lds.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
         "gemm_panel_bf16.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
         "gemm_ozaki.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}

# Packed-FP32 instructions are disabled for every device pass.
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
            "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src, build_dir=None, defines=()):
    obj = os.path.join(build_dir or BUILD, os.path.splitext(src)[0] + ".o")
    flags = HIPFLAGS + EXTRA.get(src, []) + ["-D" + d for d in defines]
    cmd = ["hipcc"] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = ["hipcc", "-x", "hip"] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), err))
    return obj, err


def build(verbose=False, out=OUT, defines=()):
    """Compile every source and link `out`.  `defines` (NAME=VALUE strings)
    build a variant of the kernels for tuning runs."""
    build_dir = BUILD if not defines else BUILD + "_" + "_".join(
        d.replace("=", "") for d in defines)
    os.makedirs(build_dir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(lambda s: _compile(s, build_dir, defines), SOURCES))
    for _, err in results:
        if verbose and err.strip():
            print(err, file=sys.stderr)
    tmp = out + ".tmp"
    cmd = ["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + \
        [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
