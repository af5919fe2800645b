"""Build libquac.so in-tree: hipcc for the gfx950 kernels, gcc for the C host layer.

    python -m quac_amd.build            # or quac_amd.build.build()

The library is written to quac_amd/libquac.so (git-ignored; it travels with the
source tree).  Objects go to quac_amd/_build/.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libquac.so")
ARCH = os.environ.get("QUAC_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
             # packed FP32 instructions are kept out of every kernel
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", os.path.join(ROOT, "include"), "-Wno-unused-result"]
C_FLAGS = ["-O2", "-fPIC", "-std=c11", "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include")]


def _sources():
    hips = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    host = os.path.join(CSRC, "host")
    cs = sorted(os.path.join(host, f) for f in os.listdir(host) if f.endswith(".c")) if os.path.isdir(host) else []
    return hips, cs


def _compile(cmd, src):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def build(verbose=False, jobs=None):
    os.makedirs(OUT, exist_ok=True)
    hips, cs = _sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs, cmds = [], []
    for src in hips:
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        cmds.append(([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj], src))
        objs.append(obj)
    for src in cs:
        obj = os.path.join(OUT, os.path.basename(src) + ".o")
        cmds.append((["gcc"] + C_FLAGS + ["-c", src, "-o", obj], src))
        objs.append(obj)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for out in ex.map(lambda a: _compile(*a), cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _compile([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lrocsolver", "-lrocblas", "-lm"], "link")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
