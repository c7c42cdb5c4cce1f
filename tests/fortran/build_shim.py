"""Build the shim test driver: the drop-in Fortran modules of
volcano-ash3d_amd/fortran compiled against the test host modules, linked to
libash3d_tr.so.  Variants, as the reference's DIFFMETH switch:
_build/cn (-DCRANKNIC), _build/ex (explicit diffusion), _build/sg
(-DCRANKNIC -DFAST_SUBGRID) and _build/cnp (-DCRANKNIC with the driver
selecting the partitioned line solver)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
PKG = os.path.join(ROOT, "volcano-ash3d_amd")
SHIMS = ["ash3d_tr_c.f90", "Diffusion.F90", "ash3d_tr_attach.F90", "AdvectionHorz_DCU.F90",
         "AdvectionHorz.F90", "AdvectionVert_DCU.F90", "Set_BC.F90", "Ash3d_Adjoint.F90"]
# test programs: the reference driver's hot-path sequence, and a source
# inversion through the adjoint entry points (Crank-Nicolson build only)
PROGRAMS = {"shim_driver": ("cn", "ex", "sg", "cnp"), "inversion_driver": ("cn",)}
VARIANTS = (("cn", ["-DCRANKNIC", "-DLIM_SUPERBEE"]), ("ex", ["-DLIM_SUPERBEE"]),
            ("sg", ["-DCRANKNIC", "-DLIM_SUPERBEE", "-DFAST_SUBGRID"]),
            ("cnp", ["-DCRANKNIC", "-DLIM_SUPERBEE", "-DLINE_SOLVER_PARTITIONED"]))
FC = os.environ.get("ASH3D_FC", "/opt/rocm/llvm/bin/amdflang")


def build(verbose=False):
    outs = {}
    for variant, defs in VARIANTS:
        bdir = os.path.join(HERE, "_build", variant)
        os.makedirs(bdir, exist_ok=True)
        for prog, variants in PROGRAMS.items():
            if variant not in variants:
                continue
            srcs = [os.path.join(HERE, "host_modules.F90")] + \
                [os.path.join(PKG, "fortran", s) for s in SHIMS] + [os.path.join(HERE, prog + ".F90")]
            exe = os.path.join(bdir, prog)
            cmd = [FC, "-O2", "-ffp-contract=off", "-cpp", *defs, "-J", bdir, "-I", bdir, *srcs, "-o", exe,
                   "-L", PKG, "-lash3d_tr", f"-Wl,-rpath,{PKG}"]
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True, cwd=bdir)
            if r.returncode != 0:
                sys.stderr.write(r.stdout + r.stderr)
                raise RuntimeError(f"building {prog} failed")
            outs[(prog, variant)] = exe
    return outs


if __name__ == "__main__":
    print(build(verbose=True))
