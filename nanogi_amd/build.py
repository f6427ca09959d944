"""Build the native library (libngi.so) and the nanogi command-line tool.

Host C++ is compiled with g++, HIP sources with hipcc for gfx950, all in
parallel, and linked in-tree so the shared object ships with the sources.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libngi.so")
CLI = os.path.join(HERE, "nanogi")
ARCH = os.environ.get("NGI_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-Wall", "-Wextra", "-I", CSRC]
HIP_FLAGS = ["-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-I", CSRC,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src, out_dir=OUT_DIR, extra=()):
    base = os.path.splitext(os.path.basename(src))[0]
    obj = os.path.join(out_dir, base + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + list(extra) + ["-c", src, "-o", obj]
    else:
        cmd = [CXX] + HOST_FLAGS + ["-c", src, "-o", obj]
    _run(cmd)
    return obj


def build(verbose=False, jobs=None, lib_path=LIB, out_dir=OUT_DIR, extra_hip_flags=()):
    """Compile every source; `extra_hip_flags`/`lib_path` build variants."""
    os.makedirs(out_dir, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".cpp", ".hip")))
    lib_srcs = [s for s in srcs if not os.path.basename(s).startswith("cli_")]
    cli_srcs = [s for s in srcs if os.path.basename(s).startswith("cli_")]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, out_dir, extra_hip_flags), lib_srcs + cli_srcs))
    lib_objs = objs[:len(lib_srcs)]
    cli_objs = objs[len(lib_srcs):]
    tmp = lib_path + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + lib_objs + ["-lz", "-lpthread"])
    os.replace(tmp, lib_path)
    if cli_objs and lib_path == LIB:
        _run([CXX, "-o", CLI] + cli_objs + ["-L", HERE, "-lngi", "-Wl,-rpath,$ORIGIN", "-lpthread"])
    if verbose:
        print("built", lib_path)
    return lib_path


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
