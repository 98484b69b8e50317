"""Build libserac_amd.so (HIP for gfx950) in-tree with hipcc.

Objects go to serac_amd/_build/, the shared library to serac_amd/libserac_amd.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libserac_amd.so")

ARCH = "gfx950"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function"]
# packed FP32 ops are disabled for all device code (gfx950 v_pk_*_f32 hazard)
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    "sx_kernels_p1.hip",
    "sx_kernels_p2.hip",
    "sx_kernels_p3.hip",
    "sx_kernels_p4.hip",
    "sx_vector.hip",
    "sx_boundary.hip",
    "sx_qoi.hip",
    "sx_mg.hip",
    "sx_unet.hip",
    "sx_learn.cpp",
    "sx_capi.cpp",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc()] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass does not know the packed-fp32 feature flag: harmless warning
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"compiling {src} failed:\n{err}")
    return obj, err


def build(jobs=None, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for obj, err in results:
            if err:
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("linking libserac_amd.so failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))


def build_cpp_tests():
    """Compile the C++ front-end test programs under tests/cpp into tests/cpp/bin
    (host C++ over the C-ABI; linked against the in-tree libraries)."""
    root = os.path.dirname(HERE)
    src_dir = os.path.join(root, "tests", "cpp")
    out_dir = os.path.join(src_dir, "bin")
    os.makedirs(out_dir, exist_ok=True)
    built = []
    for src in sorted(os.listdir(src_dir)):
        if not src.endswith(".cpp"):
            continue
        exe = os.path.join(out_dir, src[:-4])
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", os.path.join(src_dir, src), "-I", os.path.join(root, "include"),
               "-L", HERE, "-lserac_amd", "-L", os.path.join(root, "oracle"), "-loracle",
               "-Wl,-rpath,$ORIGIN/../../../serac_amd:$ORIGIN/../../../oracle", "-o", exe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compiling {src} failed:\n{r.stderr}")
        built.append(exe)
    return built
