"""Build recipe of libmrmg.so (hipcc, gfx950).  Run by __graft_entry__.build().

The library is several translation units compiled in parallel and linked into one shared
object; the demag units are compiled once per precision form, the column unit twice per
form (two halves of its instantiations)."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
# (source, object, extra flags): the demag units once per precision form
UNITS = [("mrmg.hip", "mrmg.o", []), ("grain.hip", "grain.o", []), ("tensor.hip", "tensor.o", []),
         ("demag.hip", "demag_f32.o", ["-DMRMG_FORM=float"]), ("demag.hip", "demag_f64.o", ["-DMRMG_FORM=double"]),
         ("column.hip", "column_f32a.o", ["-DMRMG_FORM=float", "-DMRMG_COL_PART=1"]),
         ("column.hip", "column_f32b.o", ["-DMRMG_FORM=float", "-DMRMG_COL_PART=2"]),
         ("column.hip", "column_f64a.o", ["-DMRMG_FORM=double", "-DMRMG_COL_PART=1"]),
         ("column.hip", "column_f64b.o", ["-DMRMG_FORM=double", "-DMRMG_COL_PART=2"])]
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libmrmg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 arithmetic
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _sources_newer(obj, src):
    """True when obj is missing or older than its unit or any header of csrc/ or include/."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    deps.append(os.path.join(HERE, "..", "include", "mrmg.h"))
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=False, reuse=False):
    """Compile every unit (or, with reuse, only those whose sources changed) and link."""
    os.makedirs(OBJ, exist_ok=True)

    def compile_unit(unit):
        u, o, extra = unit
        src = os.path.join(CSRC, u)
        obj = os.path.join(OBJ, o)
        if reuse and not _sources_newer(obj, src):
            return obj
        r = subprocess.run([HIPCC] + FLAGS + extra + ["-c", "-o", obj + ".tmp", src], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {u}:\n" + r.stderr[-4000:])
        os.replace(obj + ".tmp", obj)
        return obj

    with ThreadPoolExecutor(max_workers=len(UNITS)) as ex:
        objs = list(ex.map(compile_unit, UNITS))
    tmp = OUT + ".tmp"
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs,
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
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
    import sys
    build(verbose=True, reuse="--reuse" in sys.argv)
    build_fortran_driver(verbose=True)
