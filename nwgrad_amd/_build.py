"""Build the package's native parts in-tree (no GPU needed; hipcc cross-compiles gfx950).

* ``libnwgrad_hip.so``: every HIP kernel plus the C ABI of include/nwgrad_hip.h, built by
  hipcc for gfx950.  Compiled with -ffp-contract=off (the hard path is bit-exact only if no
  multiply-add is fused, as in the reference build, CMakeLists.txt:75) and with packed FP32
  instructions disabled.
* ``_nwgrad_amd*.so``: the pybind11 module over the header-only C++ engine; host code only,
  linked against libnwgrad_hip.so through an $ORIGIN rpath.

    python nwgrad_amd/_build.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# Translation units, compiled in parallel and linked into one library (each holds its own
# kernels; no device code crosses units).
HIP_SOURCES = ["nwg_api.hip", "hirschberg.hip"]
# Per-unit flags.  The Hirschberg unit's values are finite or -inf and never NaN (every
# operation is a max, or a sum / difference with a finite operand), so fmax needs no NaN
# quieting of its operands: -fno-honor-nans drops those extra instructions (143 VALU a row
# step instead of 170).  Infinities keep their IEEE meaning.
UNIT_FLAGS = {"hirschberg.hip": ["-fno-honor-nans"]}
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def lib_path() -> str:
    return os.path.join(HERE, "libnwgrad_hip.so")


def ext_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_nwgrad_amd" + suffix)


def _deps(path, seen=None):
    """The file and every file it includes with #include "...", transitively."""
    seen = set() if seen is None else seen
    path = os.path.normpath(path)
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("#include \""):
                _deps(os.path.join(os.path.dirname(path), line.split('"')[1]), seen)
    return seen


def build_hip(verbose: bool = False, force: bool = True) -> str:
    """Compile every HIP unit and link the library.  force=False recompiles only the units
    whose object is older than the unit or a header it includes (a developer's rebuild)."""
    out = lib_path()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        if not force and os.path.exists(obj) and os.path.getmtime(obj) > max(
                os.path.getmtime(d) for d in _deps(os.path.join(CSRC, src))):
            continue
        cmd = [HIPCC, *HIP_FLAGS, *UNIT_FLAGS.get(src, []), "-c", "-o", obj, os.path.join(CSRC, src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, cwd=CSRC, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
    failed = None
    for cmd, p in procs:
        text = p.communicate()[0]
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l]
        if p.returncode != 0:
            sys.stderr.write("\n".join(lines) + "\n")
            failed = failed or cmd
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("build failed: " + " ".join(failed))
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", out + ".tmp", *objs], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_ext(verbose: bool = False) -> str:
    import pybind11

    out = ext_path()
    cmd = ["g++", "-O2", "-std=c++20", "-fPIC", "-shared", "-fvisibility=hidden",
           "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
           os.path.join(CSRC, "py_module.cpp"), "-o", out + ".tmp",
           "-L" + HERE, "-lnwgrad_hip", "-Wl,-rpath,$ORIGIN"]
    _run(cmd, verbose)
    os.replace(out + ".tmp", out)
    return out


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=CSRC, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 target feature and says so: harmless.
    lines = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    if verbose and lines:
        print("\n".join(lines))


def build(verbose: bool = False, force: bool = True) -> None:
    build_hip(verbose, force)
    build_ext(verbose)


if __name__ == "__main__":
    build(verbose=True, force="--changed" not in sys.argv)
