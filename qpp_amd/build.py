"""Build the gfx950 engine library (libqpp_amd.so) in-tree.

Every HIP source under csrc/ is compiled for gfx950 only and linked into one
shared library next to this file.  Packed FP32 instructions are disabled
(-target-feature -packed-fp32-ops) for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqpp_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "qs_api.hip",
    "qs_gates.hip",
    "qs_maps.hip",
    "qs_measure.hip",
    "qs_fused.hip",
    "qs_program.hip",
    "qs_engine.hip",
    "qs_precision.hip",
    "qs_batch.hip",
    "qs_serve.hip",
    "qs_io.hip",
    "qs_kraus.hip",
    "qs_native.hip",
]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose=False, jobs=None):
    """Compile each source to an object in parallel, then link."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()
    procs = []
    objs = []
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [HIPCC, *FLAGS, "-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                          stderr=subprocess.STDOUT)))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((s, text))
        elif verbose and text.strip():
            print(text, flush=True)
    if failed:
        for s, text in failed:
            sys.stderr.write(f"--- {s}\n{text}\n")
        raise RuntimeError("HIP compilation failed: " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lhiprtc"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_pyqpp(verbose)
    return LIB


def pyqpp_path():
    import sysconfig
    return os.path.join(HERE, "pyqpp" + sysconfig.get_config_var("EXT_SUFFIX"))


def build_pyqpp(verbose=False):
    """The pybind11 module qpp_amd.pyqpp (host C++ over libqpp_amd.so)."""
    import sysconfig
    import pybind11
    out = pyqpp_path()
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
           "-I", os.path.join(os.path.dirname(HERE), "include"),
           os.path.join(CSRC, "pyqpp.cpp"), "-L", HERE, "-lqpp_amd",
           "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
