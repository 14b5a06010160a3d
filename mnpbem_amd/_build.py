"""Build libmnpbem_hip.so from mnpbem_amd/csrc/*.hip with hipcc for gfx950.

The library is built in-tree (mnpbem_amd/libmnpbem_hip.so) so that it travels with the
source tree. Objects go to mnpbem_amd/_build/.
"""
import concurrent.futures as cf
import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmnpbem_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for all device code (no v_pk_* f32 forms).
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-value"]


# roctx (rocprofiler-sdk) for the per-stage trace ranges of the C-ABI entry points
ROCM_LIB = "/opt/rocm/lib"
LIBS = ["-L" + ROCM_LIB, "-Wl,-rpath," + ROCM_LIB, "-lrocprofiler-sdk-roctx"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + FLAGS + ["-I", CSRC, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for {}:\n{}\n{}".format(src, " ".join(cmd), r.stderr))
    return obj


def _stale(src):
    """True when src or a header is newer than its object."""
    obj = os.path.join(OBJ, os.path.basename(src)[:-4] + ".o")
    if not os.path.exists(obj):
        return True
    deps = [src] + glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(HERE, "..", "include", "*.h"))
    return max(os.path.getmtime(d) for d in deps) > os.path.getmtime(obj)


def build(verbose=False, changed_only=False):
    """Compile every HIP source (``changed_only``: those newer than their objects) and link the
    shared library. Returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    todo = [s_ for s_ in srcs if _stale(s_)] if changed_only else srcs
    if todo:
        with cf.ThreadPoolExecutor(max_workers=min(8, len(todo))) as ex:
            list(ex.map(_compile, todo))
    objs = [os.path.join(OBJ, os.path.basename(s_)[:-4] + ".o") for s_ in srcs]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n{}\n{}".format(" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    import sys
    build(verbose=True, changed_only="--changed" in sys.argv)
