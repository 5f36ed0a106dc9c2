"""Build invdft_amd/_lib/libidft.so with hipcc for gfx950 (in-tree)."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
BUILD_DIR = os.path.join(HERE, "_lib", "obj")
SOURCES = ["kernels.hip", "noalloc.hip", "setup.hip", "fe.cpp", "context.cpp", "bfgs.cpp",
           "vxc_io.cpp", "ao.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _stale(src, obj):
    """True when obj is missing or older than src or any header it may include."""
    if not os.path.exists(obj):
        return True
    deps = [os.path.join(SRC, src)]
    deps += [os.path.join(SRC, f) for f in os.listdir(SRC) if f.endswith(".h")]
    inc = os.path.join(os.path.dirname(HERE), "include")
    deps += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src, incremental=False):
    obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
    if incremental and not _stale(src, obj):
        return obj
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(verbose=True, incremental=False):
    """Compile every source (incremental=True: only those older than a
    source or header they may include) and link libidft.so."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda f: _compile(f, incremental), SOURCES))
    lib = os.path.join(OUT_DIR, "libidft.so")
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-lrocblas", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
    sys.exit(0)
