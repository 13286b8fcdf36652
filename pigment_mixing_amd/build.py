"""Build libmpmx.so (HIP kernels for gfx950 + the C-ABI) in-tree.

    python -m pigment_mixing_amd.build          # or pigment_mixing_amd.build.build()

Compiles every translation unit with hipcc for --offload-arch=gfx950 and links
pigment_mixing_amd/lib/libmpmx.so.  Always recompiles (the caller asked for a
build); objects go to pigment_mixing_amd/_build/.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmpmx.so")
DEMO = os.path.join(OBJ, "simulation_api_demo")
ADAPTER_DEMO = os.path.join(OBJ, "adapter_style_demo")
COPY_PROBE = os.path.join(OBJ, "hbm_copy")   # HBM-counter calibration (plain copy)
PHASE_PROBE = os.path.join(OBJ, "phase_barrier")   # grid barrier vs graph kernel boundary

SOURCES = [
    "kernels/mpm_kernels.hip",
    "kernels/mlp_kernels.hip",
    "host/mpmx_api.cpp",
    "host/closure.cpp",
    "host/kinematics.cpp",
    "host/scene.cpp",
    "host/sampling.cpp",
    "host/blobs.cpp",
]

ARCH = os.environ.get("MPMX_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 math (v_pk_*_f32) anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(REPO, 'include')}",
]


def kernel_source_hash():
    """SHA-256 over the device sources (kernels/ and common/ headers, sorted by
    path): identifies the kernel build that a stored counter measurement
    belongs to."""
    import hashlib
    hs = hashlib.sha256()
    for sub in ("kernels", "common"):
        d = os.path.join(CSRC, sub)
        for name in sorted(os.listdir(d)):
            if name.endswith((".hip", ".hpp", ".h")):
                hs.update(f"{sub}/{name}\n".encode())
                with open(os.path.join(d, name), "rb") as f:
                    hs.update(f.read())
    return hs.hexdigest()


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src.replace("/", "_") + ".o")
    cmd = [hipcc(), "-x", "hip", *COMMON, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *COMMON, "-shared", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # C++ drivers of the Simulation facade (used by GPU tests): the facade's own
    # API, and a host written as the reference's renderer adapter is
    for name, exe in (("simulation_api_demo.cpp", DEMO), ("adapter_style_demo.cpp", ADAPTER_DEMO)):
        demo_src = os.path.join(REPO, "tests", "cpp", name)
        if not os.path.exists(demo_src):
            continue
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(REPO, 'include')}", demo_src, "-o", exe,
               f"-L{LIBDIR}", "-lmpmx", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath,$ORIGIN/../lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"demo build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    for name, exe in (("hbm_copy.hip", COPY_PROBE), ("phase_barrier.hip", PHASE_PROBE)):
        probe_src = os.path.join(CSRC, "bench", name)
        cmd = [hipcc(), "-x", "hip", *COMMON, probe_src, "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"probe build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
