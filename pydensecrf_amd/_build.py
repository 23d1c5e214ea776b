"""In-tree build of the native parts.

    python -m pydensecrf_amd._build

1. libdcrf_hip.so   - the HIP kernels and the C ABI (hipcc, gfx950)
2. eigen / densecrf - the Cython front end, linked against libdcrf_hip.so
3. oracle/          - the test-only CPU checkers (make -C oracle [ref])

Object files and generated C++ go to _build/ at the repository root.
"""
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
INCLUDE = os.path.join(ROOT, "include")
ORACLE = os.path.join(ROOT, "oracle")
REFERENCE = "/root/reference"

HIP_SOURCES = ["pool.hip", "lattice.hip", "filter.hip", "adjoint.hip", "kgrad.hip", "unary.hip", "objective.hip",
               "capi.hip"]
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
CY_MODULES = ["eigen", "densecrf"]
# Sources whose kernels must not use scratch (per-lane private memory): the per-pixel
# and per-vertex filter, adjoint and kernel-gradient kernels keep every array in
# registers or LDS.  The compiler's resource-usage remarks are read at every build and
# the build fails on a kernel with ScratchSize > 0.
NO_SCRATCH_SOURCES = ["filter.hip", "adjoint.hip", "kgrad.hip"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd=None, verbose=False):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def scratch_kernels(remarks):
    """(kernel, bytes per lane) of every kernel with scratch in hipcc's
    -Rpass-analysis=kernel-resource-usage remarks."""
    out, name = [], None
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and int(m.group(1)) > 0:
            out.append((name, int(m.group(1))))
    return out


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    with cf.ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        futs = {}
        for src in HIP_SOURCES:
            obj = os.path.join(BUILD, src.replace(".hip", ".o"))
            objs.append(obj)
            extra = ["-Rpass-analysis=kernel-resource-usage"] if src in NO_SCRATCH_SOURCES else []
            futs[src] = ex.submit(_run, [hipcc] + HIPCC_FLAGS + extra + ["-c", os.path.join(CSRC, src), "-o", obj],
                                  None, verbose)
        for src, f in futs.items():
            remarks = f.result()
            if src in NO_SCRATCH_SOURCES:
                bad = scratch_kernels(remarks)
                if bad:
                    raise RuntimeError("%s: kernels with scratch memory: %s" % (src, bad))
    out = os.path.join(PKG, "libdcrf_hip.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose=verbose)
    os.replace(tmp, out)
    return out


def build_cython(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    pyinc = sysconfig.get_paths()["include"]
    cython = shutil.which("cython") or "cython"
    for m in CY_MODULES:
        cpp = os.path.join(BUILD, m + ".cpp")
        _run([cython, "--cplus", "-3", "-I", ROOT, os.path.join(PKG, m + ".pyx"), "-o", cpp], cwd=ROOT,
             verbose=verbose)
        out = os.path.join(PKG, m + ext)
        tmp = out + ".tmp"
        _run(["g++", "-O2", "-fPIC", "-shared", "-w", "-I" + pyinc, "-I" + INCLUDE, cpp, "-L" + PKG, "-ldcrf_hip",
              "-Wl,-rpath,$ORIGIN", "-o", tmp], verbose=verbose)
        os.replace(tmp, out)


def build_oracle(verbose=False):
    _run(["make", "-C", ORACLE, "-s"], verbose=verbose)
    if os.path.isdir(os.path.join(REFERENCE, "pydensecrf", "densecrf", "src")):
        _run(["make", "-C", ORACLE, "-s", "ref", "REF_ROOT=" + REFERENCE], verbose=verbose)


def build(verbose=False):
    build_hip(verbose)
    build_cython(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", flush=True)
