"""Build recipe of libmrmg.so (hipcc, gfx950).  Run by __graft_entry__.build()."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "mrmg.hip")
OUT = os.path.join(HERE, "libmrmg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 arithmetic
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "mrmg_drive")


def build_fortran_driver(verbose=False):
    """The Fortran binding module and its example driver, linked against libmrmg.so with
    an rpath relative to the executable.  Skipped (returns None) without flang."""
    if not os.path.exists(FLANG):
        return None
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        cmd = [FLANG, "-O2", "-J", d, "-o", DRIVER + ".tmp", os.path.join(FORTRAN, "mrmg_m.f90"),
               os.path.join(FORTRAN, "mrmg_drive.f90"), "-L" + HERE, "-lmrmg", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=d)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stderr[-4000:])
    os.replace(DRIVER + ".tmp", DRIVER)
    if verbose:
        print("built", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(verbose=True)
    build_fortran_driver(verbose=True)
