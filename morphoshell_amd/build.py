"""Build recipe of the native library (hipcc for gfx950) and of the oracle.

``python -m morphoshell_amd.build`` compiles

* ``morphoshell_amd/_lib/libmorphoshell_amd.so``: host set-up (C++/OpenMP),
  the device context and the C-ABI, and the HIP kernels for gfx950;
* ``morphoshell_amd/_lib/ms_run``: the host driver (settings file and VTK in,
  VTK out) around the C-ABI;
* ``oracle/liboracle.so``: the FP64 CPU checker (test infrastructure only).

Every call recompiles (the build is a few seconds of hipcc per file).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIBNAME = "libmorphoshell_amd.so"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("MS_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_*_f32 in device code.
DEVICE_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}", "-Wall",
          "-Wno-unused-result", "-Wno-unknown-pragmas", "-Wno-unused-value"]

SOURCES = [
    # (file, kind, extra flags)
    ("precompute.cpp", "host", ["-fopenmp"]),
    ("context.cpp", "hip", []),
    ("tri_kernels.hip", "hip", []),
    ("node_kernels.hip", "hip", ["-ffp-contract=off"]),
    ("tile_kernels.hip", "hip", []),
    ("nn_kernels.hip", "hip", []),
    ("nn_model.cpp", "hip", []),
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    # surface warnings other than the host pass not knowing the device feature
    msg = "\n".join(
        ln for ln in (r.stdout + r.stderr).splitlines()
        if ln.strip() and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln
    )
    if msg and os.environ.get("MS_BUILD_VERBOSE"):
        print(msg, file=sys.stderr)


def compile_objects(objdir: str) -> list[str]:
    os.makedirs(objdir, exist_ok=True)
    jobs = []
    for src, kind, extra in SOURCES:
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + ".o")
        path = os.path.join(CSRC, src)
        if kind == "host":
            cmd = [HIPCC, "-x", "c++", *COMMON, *extra, "-c", path, "-o", obj]
        else:
            cmd = [HIPCC, "-x", "hip", *DEVICE_FLAGS, *COMMON, *extra, "-c", path, "-o", obj]
        jobs.append((cmd, obj))
    with ThreadPoolExecutor(max_workers=min(4, os.cpu_count() or 1)) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    return [o for _, o in jobs]


def build_library() -> str:
    objdir = os.path.join(LIBDIR, "obj")
    objs = compile_objects(objdir)
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    rpath = f"-Wl,-rpath,{ROCM}/lib,-rpath,{ROCM}/lib/llvm/lib"
    _run([HIPCC, *DEVICE_FLAGS, "-shared", "-fopenmp", "-o", tmp, *objs, rpath, f"-L{ROCM}/lib/llvm/lib"])
    os.replace(tmp, out)
    return out


DRIVER_SOURCES = ["run_settings.cpp", "vtk_io.cpp", "ms_run.cpp"]


def build_driver() -> str:
    """The host driver executable ``_lib/ms_run`` (C++17, links the library)."""
    d = os.path.join(HERE, "driver")
    out = os.path.join(LIBDIR, "ms_run")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-pthread", f"-I{os.path.join(ROOT, 'include')}", f"-I{d}",
          *[os.path.join(d, s) for s in DRIVER_SOURCES], f"-L{LIBDIR}", "-lmorphoshell_amd",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    d = os.path.join(ROOT, "oracle")
    subprocess.run(["make", "-s", "-B", "-C", d], check=True, capture_output=True)
    return os.path.join(d, "liboracle.so")


def main() -> None:
    lib = build_library()
    drv = build_driver()
    orc = build_oracle()
    print(lib)
    print(drv)
    print(orc)


if __name__ == "__main__":
    main()
