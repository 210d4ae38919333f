"""Builds librimp2_hip.so (HIP, gfx950) in-tree and the Fortran host driver.

Invoked by __graft_entry__.build() and usable directly:
    python -m rimp2_gpu_amd.build [--no-fortran]
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "build")
LIB = os.path.join(HERE, "librimp2_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fopenmp",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def build_library(jobs=8):
    os.makedirs(OUT, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(OUT, os.path.basename(s) + ".o") for s in srcs]

    def compile_one(so):
        s, o = so
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        _run([HIPCC] + HIP_FLAGS + lang + ["-c", s, "-o", o + ".tmp"])
        os.replace(o + ".tmp", o)

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(compile_one, zip(srcs, objs)))
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fopenmp", "-o", LIB + ".tmp"] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lrocprofiler-sdk-roctx", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(LIB + ".tmp", LIB)
    return LIB


MEASURE_LIB = os.path.join(HERE, "librimp2_measure.so")


def build_measure():
    """Builds librimp2_measure.so: the benchmark's probe of the int8 MFMA rate
    the card sustains (measure/mfma_ceiling.hip); not part of the energy path."""
    src = os.path.join(HERE, "measure", "mfma_ceiling.hip")
    _run([HIPCC] + HIP_FLAGS + ["-x", "hip", "-shared", src, "-o", MEASURE_LIB + ".tmp"])
    os.replace(MEASURE_LIB + ".tmp", MEASURE_LIB)
    return MEASURE_LIB


def build_fortran():
    """Builds the Fortran host driver (rimp2_driver.x) against the library."""
    fdir = os.path.join(HERE, "fortran")
    if not os.path.isdir(fdir):
        return None
    flang = os.path.join(ROCM, "bin", "amdflang")
    exe = os.path.join(fdir, "rimp2_driver.x")
    mods = os.path.join(OUT, "fmod")
    os.makedirs(mods, exist_ok=True)
    srcs = [os.path.join(fdir, "rimp2_hip_interfaces.F90"), os.path.join(fdir, "rimp2_driver.F90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    _run([flang, "-O2", "-J", mods, "-o", exe + ".tmp"] + srcs +
         ["-L" + HERE, "-lrimp2_hip", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(exe + ".tmp", exe)
    return exe


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    print("built", build_library())
    print("built", build_measure())
    if "--no-fortran" not in argv:
        exe = build_fortran()
        if exe:
            print("built", exe)


if __name__ == "__main__":
    main()
