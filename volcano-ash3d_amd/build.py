"""Build libash3d_tr.so (the gfx950 HIP library behind include/ash3d_tr.h).

Run ``python volcano-ash3d_amd/build.py`` or call ``build()``.  The library
is written next to this file so that it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libash3d_tr.so")
SOURCES = ["ash3d_tr.hip", "bc.hip", "advect.hip", "advect_x.hip", "diffuse.hip", "totals.hip", "transpose.hip",
           "adjoint.hip", "meteo.hip", "linesolve.hip", "products.hip", "source.hip", "advect_hz.hip"]
ARCH = os.environ.get("ASH3D_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the kernels evaluate every expression with the
# reference's rounding (no fused multiply-add), which makes the results
# bitwise those of the Fortran routines.  Packed FP32 is disabled as a
# matter of policy (the library is FP64 throughout).
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose: bool = False, jobs: int | None = None) -> str:
    """Compile each source to an object (in parallel, up to jobs at once)
    and link the shared library."""
    import concurrent.futures
    import tempfile
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    jobs = jobs or min(8, os.cpu_count() or 1)
    flags = [f for f in HIPFLAGS if f != "-shared"]
    tmp = LIB + ".tmp"
    with tempfile.TemporaryDirectory(prefix="ash3d_build_") as work:
        def compile_one(src):
            obj = os.path.join(work, os.path.splitext(src)[0] + ".o")
            cmd = [hipcc, *flags, "-I", os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            return obj, r
        with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
            results = list(ex.map(compile_one, SOURCES))
        for obj, r in results:
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError(f"hipcc failed compiling {os.path.basename(obj)}")
        cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *[o for o, _ in results], "-o", tmp]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("hipcc failed linking libash3d_tr.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
