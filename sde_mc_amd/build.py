"""Build libsdemc.so (the HIP kernels and the C ABI) for gfx950 with hipcc.

The library is compiled in-tree into ``sde_mc_amd/lib/`` so it travels with
the source tree.  Every .hip file under ``csrc/`` is one translation unit;
they are compiled in parallel and linked into one shared object.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIBDIR = os.path.join(HERE, "lib")
BUILDDIR = os.path.join(ROOT, "_build", "sdemc")
LIBNAME = "libsdemc.so"

ARCH = os.environ.get("SDEMC_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no packed fp32 instructions anywhere in the device code.
DEVICE_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE,
]


# Per-file additions.  The control-variate kernels keep MFMA accumulators in
# ordinary VGPRs (no accumulator-register round trips before the ReLU).
FILE_FLAGS = {
    "k_cv_uniform.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "k_cv_adapted.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def _compile(src):
    obj = os.path.join(BUILDDIR, os.path.splitext(os.path.basename(src))[0] + ".o")
    cmd = [HIPCC] + DEVICE_FLAGS + FILE_FLAGS.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr))
    return obj


def build(verbose=False):
    """Compile every translation unit and link libsdemc.so; returns its path."""
    os.makedirs(BUILDDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    out = lib_path()
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr)
    os.replace(tmp, out)
    # launch shapes tuned on the previous build no longer apply
    tuned = os.path.join(LIBDIR, "autotune.json")
    if os.path.exists(tuned):
        os.remove(tuned)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
