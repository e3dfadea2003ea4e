"""Build the HIP library liblatticeflow_amd.so in-tree with hipcc (gfx950)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblatticeflow_amd.so")
SOURCES = ["lf_mincut.hip", "lf_mincut_canon.hip", "lf_mincut_generic.hip", "lf_mincut_wide.hip", "lf_mincut_wide32.hip", "lf_mincut_wide48.hip", "lf_tvnet.hip",
           "lf_tv.cpp", "lf_tvpath_core.cpp", "lf_tvpath.cpp", "lf_bk.cpp", "lf_verify.cpp", "lf_capi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 ops are disabled for all device code of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def _newest_header() -> float:
    heads = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    heads.append(os.path.join(HERE, "..", "include", "latticeflow_amd.h"))
    return max(os.path.getmtime(h) for h in heads if os.path.exists(h))


def build(verbose: bool = False, jobs: int = 8, incremental: bool = False) -> str:
    """Compile every source to an object, then link the shared library.

    ``incremental``: skip sources whose object is newer than the source and
    every header (a development shortcut; the default rebuilds everything)."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    head_t = _newest_header()
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        if incremental and os.path.exists(obj):
            ot = os.path.getmtime(obj)
            deps = [os.path.join(CSRC, src)]
            if src.startswith(("lf_mincut_wide", "lf_mincut_canon")):
                deps.append(os.path.join(CSRC, "lf_mincut.hip"))  # they include it
            if ot > max(os.path.getmtime(d) for d in deps) and ot > head_t:
                continue
        lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
        cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            failed.append(src)
            sys.stderr.write(out.decode(errors="replace"))
        elif verbose and out:
            sys.stderr.write(out.decode(errors="replace"))
    if failed:
        raise RuntimeError("hipcc failed for: %s" % ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_ext(verbose: bool = False) -> str:
    """The Cython extension ext/pylatticeflow (over include/latticeflow_amd/*.hpp),
    linked against the library built above with a relative run path."""
    import sysconfig
    from Cython.Build import cythonize  # noqa: F401  (Cython 3 is the translator)
    from Cython.Compiler import Main as cymain
    ext = os.path.join(HERE, "ext")
    pyx = os.path.join(ext, "pylatticeflow.pyx")
    cpp = os.path.join(HERE, "_build", "pylatticeflow.cpp")
    os.makedirs(os.path.dirname(cpp), exist_ok=True)
    opts = cymain.CompilationOptions(cplus=True, language_level=3, output_file=cpp)
    res = cymain.compile(pyx, opts)
    if res.num_errors:
        raise RuntimeError("cython failed for pylatticeflow.pyx")
    out = os.path.join(ext, "pylatticeflow" + sysconfig.get_config_var("EXT_SUFFIX"))
    inc = [sysconfig.get_paths()["include"], os.path.join(HERE, "..", "include")]
    cmd = ["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-w"] + ["-I" + i for i in inc] + [
        cpp, "-o", out + ".tmp", "-L" + HERE, "-llatticeflow_amd", "-Wl,-rpath,$ORIGIN/..",
        "-Wl,-rpath-link,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
    print(build_ext(verbose="-v" in sys.argv))
